"""Two-stage batch plans (TwoStageBatch): the batched two-stage launch
sequence replayed from HIP graphs, with autotuned stage-1 launch shapes.

Bar: the same as the one-call two-stage search -- candidate and top-k ids
bit-exact against the oracle, f32 scores bit-identical (1e-5 relative
contract tolerance asserted too).  Graph replays are checked one by one, with
the program's own eager GPU work and synchronisations between them.
"""
import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _assert_scores(got, want):
    np.testing.assert_allclose(got, want, rtol=RTOL, atol=1e-7)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "scores not bit-identical"


def _np(x):
    return x.cpu().numpy() if hasattr(x, "cpu") else x


def _oracle(oracle, rq, sb, qs, m, k):
    dim = qs.shape[1]
    cand = oracle.sign_top_m(sb.words(), oracle.sign_encode(qs), m, 8)
    offs = np.arange(len(qs) + 1, dtype=np.uint64) * cand.shape[1]
    return oracle.subset_batched(rq.codes(), dim, 2, qs, offs, cand.reshape(-1), k, 8)


@pytest.fixture(scope="module")
def mixture_60k(engine):
    d, n = 1024, 60_000
    gen = S.MixtureNP(d, seed=23)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 400)
    sb = engine.SignBitmap(d)
    rq = engine.RankQuant(d, 2)
    engine.two_stage_add(sb, rq, docs)
    return docs, qs, sb, rq


def test_graph_replays_equal_oracle_and_eager(engine, oracle, mixture_60k):
    """Query counts on every scan form (1: vector-ALU scan; 40, 100, 300:
    matrix-core tiles of 64 and 128 queries, 300 with the side-stream query
    preparation); device tensors through graphs and eagerly, host arrays, and
    three replays of each graph with eager work in between."""
    import torch

    _, qs, sb, rq = mixture_60k
    m, k = 1000, 10
    plan = engine.TwoStageBatch(sb, rq, 400, m, k)
    for nq in (1, 40, 100, 300):
        sel = qs[:nq]
        ws, wi = _oracle(oracle, rq, sb, sel, m, k)
        tq = torch.from_numpy(sel.copy()).cuda()
        out = (torch.empty((nq, k), dtype=torch.float32, device="cuda"),
               torch.empty((nq, k), dtype=torch.int64, device="cuda"))
        for rep in range(3):
            out[0].fill_(7.0)
            out[1].fill_(-7)
            s, i = plan.run(tq, out=out)
            assert np.array_equal(_np(i), wi), (nq, rep)
            _assert_scores(_np(s), ws)
            # the program's own eager GPU work between replays
            es, ei = engine.two_stage_search(sb, rq, tq, m, k)
            assert np.array_equal(_np(ei), wi)
            torch.cuda.synchronize()
        s, i = plan.run(tq, use_graph=False)
        assert np.array_equal(_np(i), wi), nq
        _assert_scores(_np(s), ws)
        s, i = plan.run(sel)
        assert np.array_equal(i, wi), nq
        _assert_scores(s, ws)
    st = plan.stats()
    assert st["eager_calls"] == 4 and st["exact_queries"] == 0
    # one graph pair per (query count, buffers), the last four kept
    assert st["graph_captures"] >= 8 and st["graph_calls"] == 16


def test_shapes_do_not_change_results(engine, oracle, mixture_60k):
    import torch

    _, qs, sb, rq = mixture_60k
    m, k = 700, 16
    sel = qs[:300]
    ws, wi = _oracle(oracle, rq, sb, sel, m, k)
    tq = torch.from_numpy(sel.copy()).cuda()
    plan = engine.TwoStageBatch(sb, rq, 300, m, k)
    shapes = [dict(scan_tiles=4, scan_pairs=512, scan_flat=0, sample_wgs=128, side_prep=1),
              dict(scan_tiles=2, scan_pairs=2048, scan_flat=1, sample_wgs=512, side_prep=0),
              dict(scan_tiles=1, scan_pairs=1024, scan_flat=0, sample_wgs=256, side_prep=1),
              dict(scan_tiles=4, scan_pairs=4096, scan_flat=1, sample_wgs=0, side_prep=0),
              dict(scan_tiles=0, scan_pairs=0, scan_flat=0, sample_wgs=0, side_prep=1)]
    for sh in shapes:
        assert plan.set_shape(**sh) == sh
        for _ in range(2):
            s, i = plan.run(tq)
            assert np.array_equal(_np(i), wi), sh
            _assert_scores(_np(s), ws)
    with pytest.raises(ValueError):
        plan.set_shape(scan_tiles=3)
    with pytest.raises(ValueError):
        plan.set_shape(bogus=1)


def test_autotune_keeps_a_measured_shape(engine, oracle, mixture_60k):
    import torch

    _, qs, sb, rq = mixture_60k
    m, k = 1000, 10
    sel = qs[:256]
    ws, wi = _oracle(oracle, rq, sb, sel, m, k)
    tq = torch.from_numpy(sel.copy()).cuda()
    plan = engine.TwoStageBatch(sb, rq, 256, m, k)
    rows = plan.autotune(tq, reps=2)
    assert len(rows) >= 8 and all(r["ms"] > 0 for r in rows)
    best = min(rows, key=lambda r: r["ms"])
    assert plan.shape() == {key: best[key] for key in plan._SHAPE_KEYS}
    s, i = plan.run(tq)
    assert np.array_equal(_np(i), wi)
    _assert_scores(_np(s), ws)
    # host queries tune too
    assert len(plan.autotune(sel[:130], reps=1)) >= 8
    s, i = plan.run(sel[:130])
    assert np.array_equal(i, wi[:130])


def test_flagged_queries_are_recomputed_behind_the_graph(engine, oracle):
    """A corpus with 20k identical documents: queries near them overflow the
    survivor buffer and take the exact stage-1 pipeline after the graphs."""
    import torch

    rng = np.random.default_rng(29)
    dim = 256
    gen = S.MixtureNP(dim, seed=31)
    docs = gen.docs(40_000)
    dup = docs[77].copy()
    docs[2000:22000] = dup
    qs, _ = gen.queries(docs, 200)
    qs[::9] = dup + 0.001 * rng.standard_normal((len(qs[::9]), dim)).astype(np.float32)
    sb = engine.SignBitmap(dim)
    sb.add(docs)
    rq = engine.RankQuant(dim, 2)
    rq.add(docs)
    bm, qb = oracle.sign_encode(docs), oracle.sign_encode(qs)
    cand = oracle.sign_top_m(bm, qb, 700)
    offs = np.arange(len(qs) + 1, dtype=np.uint64) * 700
    ws, wi = oracle.subset_batched(oracle.rq_encode(docs, 2), dim, 2, qs, offs, cand.reshape(-1), 10)
    plan = engine.TwoStageBatch(sb, rq, 200, 700, 10)
    tq = torch.from_numpy(qs.copy()).cuda()
    for _ in range(2):
        s, i = plan.run(tq)
        assert np.array_equal(_np(i), wi)
        _assert_scores(_np(s), ws)
    assert plan.stats()["exact_queries"] >= 2 * len(qs[::9])
    s, i = plan.run(qs)
    assert np.array_equal(i, wi)


def test_buffers_indexes_and_errors(engine, oracle, mixture_60k):
    import torch

    from ordvec_amd import _native as N

    docs, qs, sb, rq = mixture_60k
    m, k = 500, 10
    plan = engine.TwoStageBatch(sb, rq, 64, m, k)
    ws, wi = _oracle(oracle, rq, sb, qs[:64], m, k)
    # five distinct query buffers of one size: each captured, the oldest evicted
    bufs = [torch.from_numpy(qs[:64].copy()).cuda() for _ in range(5)]
    for b in bufs + bufs[:1]:
        s, i = plan.run(b)
        assert np.array_equal(_np(i), wi)
    assert plan.stats()["graph_captures"] == 6
    # the plan pins its indexes
    with pytest.raises(ValueError):
        rq.add(docs[:3])
    with pytest.raises(ValueError):
        sb.swap_remove(0)
    with pytest.raises(ValueError):
        plan.run(qs[:65])
    bad = torch.from_numpy(qs[:4].copy()).cuda()
    bad[2, 5] = float("nan")
    with pytest.raises(ValueError):
        plan.run(bad)
    plan.run(bad, assume_finite=True)  # the caller vouched: no check
    with pytest.raises(NotImplementedError):
        engine.TwoStageBatch(sb, rq, 8, N.MAX_STAGE1_M + 1, 10)
    # freeing the plan unpins the indexes (this is the module's last use of them)
    plan.close()
    n0 = len(rq)
    engine.two_stage_add(sb, rq, docs[:5])
    assert len(rq) == n0 + 5 and len(sb) == n0 + 5
