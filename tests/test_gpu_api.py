"""Python API surface of the reference binding beyond search
(ordvec-python/src/lib.rs:1488-1760, src/sign_bitmap.rs:204-214,485-560):
SignBitmap.score_all / score_all_batched / build_query_bitmap, the module
rank_transform and rankquant_eval_search, and the rank encoders at widths
above the per-row LDS sort (the per-row segmented sort, up to 65535) --
all bit-exact against the oracle."""
import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _tricky(d, rng):
    v = rng.standard_normal(d).astype(np.float32)
    v[rng.random(d) < 0.2] = 0.0
    v[rng.random(d) < 0.1] = -0.0
    v[: d // 5] = v[d // 5: 2 * (d // 5)]          # ties
    return v


@pytest.mark.parametrize("dim", [64, 1024, 4096])
def test_score_all_matches_oracle(engine, oracle, dim):
    import torch

    x = S.iid_unit(5000, dim, seed=dim)
    q = S.iid_unit(7, dim, seed=dim + 1)
    sb = engine.SignBitmap(dim)
    sb.add(x)
    words = oracle.sign_encode(x)
    qb = oracle.sign_encode(q)
    want = np.stack([dim - oracle.hamming_all(words, qb[t]) for t in range(len(q))]).astype(np.uint32)
    got = sb.score_all_batched(q)
    assert got.dtype == np.uint32 and got.shape == (7, 5000)
    assert np.array_equal(got, want)
    gd = sb.score_all_batched(torch.from_numpy(q).cuda())
    assert np.array_equal(gd.cpu().numpy().view(np.uint32) if gd.dtype != torch.uint32 else gd.cpu().numpy(), want)
    assert np.array_equal(sb.score_all(q[3]), want[3])
    assert sb.score_all_batched(q[:0]).shape == (0, 5000)
    empty = engine.SignBitmap(dim)
    assert empty.score_all_batched(q).shape == (7, 0)
    bad = q.copy()
    bad[2, 5] = np.nan
    with pytest.raises(ValueError):
        sb.score_all_batched(bad)


def test_build_query_bitmap_matches_oracle(engine, oracle):
    import torch

    rng = np.random.default_rng(3)
    for dim in (64, 768, 2048):
        sb = engine.SignBitmap(dim)
        for v in (_tricky(dim, rng), (rng.standard_normal(dim) * 1e-42).astype(np.float32),
                  np.full(dim, -0.0, np.float32)):
            want = oracle.sign_encode(v[None])[0]
            assert np.array_equal(sb.build_query_bitmap(v), want)
            got = sb.build_query_bitmap(torch.from_numpy(v).cuda())
            assert np.array_equal(got.cpu().numpy().view(np.uint64), want)
    with pytest.raises(ValueError):
        engine.SignBitmap(64).build_query_bitmap(np.ones(63, np.float32))


@pytest.mark.parametrize("d", [1, 2, 8, 1024, 4096, 4097, 5000, 65535])
def test_rank_transform_matches_oracle(engine, oracle, d):
    import torch

    rng = np.random.default_rng(d)
    v = _tricky(d, rng) if d > 4 else rng.standard_normal(d).astype(np.float32)
    want = oracle.rank_transform(v)
    got = engine.rank_transform(v)
    assert got.dtype == np.uint16 and np.array_equal(got, want)
    gd = engine.rank_transform(torch.from_numpy(v).cuda())
    assert np.array_equal(gd.cpu().numpy().view(np.uint16), want)


def test_rank_transform_kats_and_errors(engine):
    assert engine.rank_transform(np.array([3, 1, 4, 1.5, 5, 9, 2, 6], np.float32)).tolist() == [3, 0, 4, 1, 5, 7, 2, 6]
    assert engine.rank_transform(np.array([0.0, -0.0, 1.0, -0.0, 0.0], np.float32)).tolist() == [0, 1, 4, 2, 3]
    assert engine.rank_transform(np.zeros(0, np.float32)).shape == (0,)
    with pytest.raises(ValueError):
        engine.rank_transform(np.array([1.0, np.nan], np.float32))
    with pytest.raises(ValueError):
        engine.rank_transform(np.zeros(65536, np.float32))


@pytest.mark.parametrize("dim,bits", [(8192, 2), (8192, 4), (65528, 1)])
def test_rankquant_add_above_lds_sort_width(engine, oracle, dim, bits):
    import torch

    rng = np.random.default_rng(dim + bits)
    n = 40 if dim < 65535 else 6
    x = rng.standard_normal((n, dim)).astype(np.float32)
    x[1] = _tricky(dim, rng)
    x[2] = 0.25
    rq = engine.RankQuant(dim, bits)
    rq.add(x[: n // 2])
    rq.add(torch.from_numpy(x[n // 2:]).cuda())
    want = oracle.rq_encode(x, bits)
    assert np.array_equal(rq.codes(), want)
    if dim <= 8192:
        q = rng.standard_normal((3, dim)).astype(np.float32)
        s, i = rq.search(q, 5)
        ws, wi = oracle.rq_search_symmetric(want, dim, bits, q, 5)
        assert np.array_equal(i, wi)
        assert np.array_equal(s, ws)


@pytest.mark.parametrize("dim", [5000, 65535])
def test_rank_add_above_lds_sort_width(engine, oracle, dim):
    rng = np.random.default_rng(dim)
    n = 24 if dim < 65535 else 4
    x = rng.standard_normal((n, dim)).astype(np.float32)
    x[0] = _tricky(dim, rng)
    r = engine.Rank(dim)
    r.add(x)
    assert np.array_equal(r.ranks(), oracle.rank_encode(x))
    if dim <= 8192:
        q = rng.standard_normal((2, dim)).astype(np.float32)
        s, i = r.search(q, 4)
        ws, wi = oracle.rank_search(oracle.rank_encode(x), q, 4)
        assert np.array_equal(i, wi)
        assert np.array_equal(s, ws)


@pytest.mark.parametrize("bits", [1, 2, 3, 4, 5, 6, 7])
def test_rankquant_eval_search_matches_oracle(engine, oracle, bits):
    import torch

    rng = np.random.default_rng(90 + bits)
    d = 384 if bits != 5 else 100
    corpus = rng.standard_normal((3000, d)).astype(np.float32)
    corpus[10] = corpus[20]
    queries = rng.standard_normal((9, d)).astype(np.float32)
    queries[0] = corpus[5]
    ws, wi = oracle.eval_search(corpus, queries, bits, 12)
    s, i = engine.rankquant_eval_search(corpus, queries, bits, 12)
    assert np.array_equal(i, wi)
    assert np.array_equal(s, ws)
    sd, idd = engine.rankquant_eval_search(torch.from_numpy(corpus).cuda(), torch.from_numpy(queries).cuda(), bits, 12)
    assert np.array_equal(idd.cpu().numpy(), wi)
    assert np.array_equal(sd.cpu().numpy(), ws)
    if bits in (1, 2, 4):
        rq = engine.RankQuant(d, bits)
        rq.add(corpus)
        ps, pi = rq.search(queries, 12)
        assert np.array_equal(pi, i)
        np.testing.assert_allclose(ps, s, rtol=0, atol=1e-6)


def test_rankquant_eval_search_edges(engine, oracle):
    rng = np.random.default_rng(7)
    c = rng.standard_normal((5, 64)).astype(np.float32)
    s, i = engine.rankquant_eval_search(c, c[:2], 3, 10)       # k above n clamps to n
    assert s.shape == (2, 5) and i.shape == (2, 5)
    ws, wi = oracle.eval_search(c, c[:2], 3, 10)
    assert np.array_equal(i, wi) and np.array_equal(s, ws)
    assert engine.rankquant_eval_search(c, c[:0], 3, 4)[0].shape == (0, 4)
    for bad_bits in (0, 8):
        with pytest.raises(ValueError):
            engine.rankquant_eval_search(c, c[:1], bad_bits, 4)
    with pytest.raises(ValueError):
        engine.rankquant_eval_search(c[:, :1], c[:1, :1], 2, 4)
    bad = c.copy()
    bad[1, 1] = np.inf
    with pytest.raises(ValueError):
        engine.rankquant_eval_search(bad, c[:1], 2, 4)
