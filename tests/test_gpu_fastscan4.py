"""FastScan b=4 on the GPU against the oracle (bit-exact ids and scores), and
the batched multi-index form: graph replays, eager runs, autotuned configs
and per-index plain searches all give identical rows."""
import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,n", [(768, 5000), (64, 1000), (1024, 70_000), (2048, 3000), (768, 37)])
def test_plain_b4_matches_oracle(engine, oracle, dim, n):
    gen = S.MixtureNP(dim, seed=dim + n)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 40)
    ix = engine.RankQuantFastscan(dim, bits=4)
    ix.add(docs)
    assert len(ix) == n and ix.bits == 4
    codes = oracle.rq_encode(docs, 4)
    # 17 and 40: whole 16-query sets on matrix-core tiles, the rest on LUT tiles
    for nq in (1, 5, 17, 40):
        for k in (1, 10, 100, 256):
            s, i = ix.search(qs[:nq], k)
            ws, wi = oracle.fs4_search(codes, dim, qs[:nq], k)
            assert np.array_equal(i, wi), (nq, k)
            assert np.array_equal(s.view(np.uint32), ws.view(np.uint32)), (nq, k)


def test_b4_validation(engine):
    with pytest.raises(ValueError):
        engine.RankQuantFastscan(100, bits=4)  # dim % 8
    with pytest.raises(ValueError):
        engine.RankQuantFastscan(64, bits=3)
    ix = engine.RankQuantFastscan(64, bits=4)
    ix.add(S.iid_unit(400, 64, seed=1))
    with pytest.raises(NotImplementedError):
        ix.search(S.iid_unit(2, 64, seed=2), 300)  # k > 256
    with pytest.raises(NotImplementedError):
        ix.write("/tmp/unused.ovfs")


def _batch(engine, dim, sizes, nqs, ks, seed):
    rng = np.random.default_rng(seed)
    indexes, queries, docs_all = [], [], []
    for j, n in enumerate(sizes):
        gen = S.MixtureNP(dim, seed=seed * 100 + j, n_centroids=64)
        docs = gen.docs(n) if n else np.zeros((0, dim), np.float32)
        ix = engine.RankQuantFastscan(dim, bits=4)
        if n:
            ix.add(docs)
        indexes.append(ix)
        docs_all.append(docs)
        src = docs if n else S.iid_unit(max(1, nqs[j]), dim, seed=j)
        q = src[rng.integers(0, len(src), size=nqs[j])] + 0.1 * rng.standard_normal((nqs[j], dim)).astype(np.float32)
        queries.append(q.astype(np.float32))
    return indexes, np.concatenate(queries), docs_all


def _plain(indexes, queries, nqs, ks):
    out, r = [], 0
    for ix, nq, k in zip(indexes, nqs, ks):
        out.append(ix.search(queries[r:r + nq], k) if nq else None)
        r += nq
    return out


def _check(batch, scores, ids, plain):
    for (s, i), p, nq in zip(batch.split(scores, ids), plain, batch.queries_per_index):
        if nq == 0:
            continue
        assert np.array_equal(i, p[1]) and np.array_equal(s.view(np.uint32), p[0].view(np.uint32))


def test_batch_equals_plain_and_oracle(engine, oracle):
    dim = 768
    sizes = [1000, 70, 0, 20_000, 5, 3000, 64, 65, 128, 4000, 257, 9000]
    nqs = [3, 16, 2, 21, 1, 0, 7, 4, 9, 16, 5, 12]
    ks = [10, 100, 10, 100, 10, 100, 100, 10, 10, 100, 256, 10]
    indexes, queries, docs = _batch(engine, dim, sizes, nqs, ks, seed=3)
    b = engine.FastscanBatch(indexes, nqs, ks)
    cfg = b.config()
    assert cfg["total_queries"] == sum(nqs) and cfg["kout"] == 256
    plain = _plain(indexes, queries, nqs, ks)
    s, i = b.run(queries)
    _check(b, s, i, plain)
    # padding past each row's min(k, n)
    for (ss, ii), k, n in zip(b.split(s, i), ks, sizes):
        if len(ii):
            assert ii.shape[1] == min(k, n)
    r0 = 0
    for nq, n in zip(nqs, sizes):
        if min(n, 256) < 256 and nq:
            assert np.all(i[r0:r0 + nq, min(n, 256):] == -1)
        r0 += nq
    # the oracle on every index that has queries
    r = 0
    for j, (nq, k) in enumerate(zip(nqs, ks)):
        if nq and sizes[j]:
            ws, wi = oracle.fs4_search(oracle.rq_encode(docs[j], 4), dim, queries[r:r + nq], k)
            ss, ii = b.split(s, i)[j]
            assert np.array_equal(ii, wi) and np.array_equal(ss.view(np.uint32), ws.view(np.uint32)), j
        r += nq
    # every tile shape gives the same rows: LDS-LUT tiles of 4 and 8 queries,
    # matrix-core tiles of 16 (ragged query sets, segments and last blocks)
    # and the mixed form (20 / 24: whole 16-query sets on the matrix cores, the
    # rest on LUT tiles of 4 / 8)
    for qpt, bpt, grid in ((16, 16, 256), (16, 5, 64), (16, 256, 512), (8, 32, 256), (4, 7, 100), (20, 64, 256),
                           (24, 13, 128), (20, 256, 512)):
        cfg = b.set_config(qpt, bpt, grid)
        assert (cfg["queries_per_tile"], cfg["blocks64_per_tile"], cfg["grid"]) == (qpt, bpt, grid)
        for use_graph in (True, False):
            s2, i2 = b.run(queries, use_graph=use_graph)
            assert np.array_equal(i2, i) and np.array_equal(s2.view(np.uint32), s.view(np.uint32)), (qpt, bpt, grid)


def test_graph_replays_eager_runs_and_autotune(engine):
    import torch

    dim = 256
    sizes = [5000, 300, 12_000, 64, 900]
    nqs = [9, 3, 16, 1, 6]
    ks = [10, 100, 10, 100, 10]
    indexes, queries, _ = _batch(engine, dim, sizes, nqs, ks, seed=4)
    b = engine.FastscanBatch(indexes, nqs, ks)
    plain = _plain(indexes, queries, nqs, ks)
    eager_s, eager_i = b.run(queries, use_graph=False)
    _check(b, eager_s, eager_i, plain)
    other = queries[::-1].copy()
    plain_other = _plain(indexes, other, nqs, ks)
    for rep in range(4):  # replays with unrelated eager work and syncs between them
        s, i = b.run(queries if rep % 2 == 0 else other)
        _check(b, s, i, plain if rep % 2 == 0 else plain_other)
        x = torch.randn(1000, 1000, device="cuda")
        (x @ x).sum().item()
        torch.cuda.synchronize()
    rows = b.autotune(queries, reps=1)
    # 30 LDS-LUT shapes + 10 matrix-core shapes (dim % 64 == 0) + 12 mixed
    assert len(rows) == 52 and all(r["ms"] > 0 for r in rows)
    assert sum(r["queries_per_tile"] == 16 for r in rows) == 10
    assert sum(r["queries_per_tile"] in (20, 24) for r in rows) == 12
    best = min(rows, key=lambda r: r["ms"])
    cfg = b.config()
    assert (cfg["queries_per_tile"], cfg["blocks64_per_tile"], cfg["grid"]) == (
        best["queries_per_tile"], best["blocks64_per_tile"], best["grid"])
    s, i = b.run(queries)
    _check(b, s, i, plain)
    # device tensors in, device tensors out
    s, i = b.run(torch.tensor(other, device="cuda"))
    _check(b, s.cpu().numpy(), i.cpu().numpy(), plain_other)
    # a second batch of the same shape class starts from the tuned configuration
    b2 = engine.FastscanBatch(indexes, nqs, ks)
    cfg2 = b2.config()
    assert (cfg2["queries_per_tile"], cfg2["blocks64_per_tile"], cfg2["grid"]) == (
        cfg["queries_per_tile"], cfg["blocks64_per_tile"], cfg["grid"])


def test_batch_validation(engine):
    a = engine.RankQuantFastscan(64, bits=4)
    a.add(S.iid_unit(100, 64, seed=1))
    b2 = engine.RankQuantFastscan(64, bits=2)
    b2.add(S.iid_unit(100, 64, seed=1))
    c = engine.RankQuantFastscan(128, bits=4)
    c.add(S.iid_unit(100, 128, seed=1))
    with pytest.raises(ValueError):
        engine.FastscanBatch([a, b2], [1, 1], [10, 10])
    with pytest.raises(ValueError):
        engine.FastscanBatch([a, c], [1, 1], [10, 10])
    with pytest.raises(ValueError):
        engine.FastscanBatch([a], [1], [257])
    bt = engine.FastscanBatch([a], [2], [10])
    with pytest.raises(ValueError):
        bt.run(S.iid_unit(3, 64, seed=3))
    for bad_cfg in ((12, 16, 64), (28, 16, 64), (16, 0, 64), (16, 16, 0)):
        with pytest.raises(ValueError):
            bt.set_config(*bad_cfg)
    d72 = engine.RankQuantFastscan(72, bits=4)  # matrix-core tiles need dim % 64 == 0
    d72.add(S.iid_unit(100, 72, seed=1))
    with pytest.raises(ValueError):
        engine.FastscanBatch([d72], [2], [10]).set_config(16, 16, 64)
    bad = S.iid_unit(2, 64, seed=3)
    bad[0, 0] = np.nan
    with pytest.raises(ValueError):
        bt.run(bad)
