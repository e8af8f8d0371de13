"""Top-bucket Bitmap (AND-popcount) on the GPU against the oracle: encode,
query bitmaps, overlaps and every selection bit-exact; .ovbm through the
Part 2 API and through ABI v1 (ordvec_index_load / ordvec_index_search)."""
import ctypes as C

import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _corpus(n, d, seed, ties=False):
    gen = S.MixtureNP(d, seed=seed)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 48)
    if ties:
        docs = np.round(docs * 8).astype(np.float32)  # many equal coordinates
        docs[: n // 4] = docs[0]                      # many identical documents
        qs = np.round(qs * 8).astype(np.float32)
        qs[0, :20] = -0.0
    return docs, qs


@pytest.mark.parametrize("d,n_top", [(64, 16), (128, 32), (768, 192), (1024, 256), (4096, 1000)])
def test_encode_and_query_bits(engine, oracle, d, n_top):
    docs, qs = _corpus(700, d, seed=d, ties=True)
    bm = engine.Bitmap(d, n_top)
    bm.add(docs)
    assert len(bm) == 700 and bm.n_top == n_top and bm.bytes_per_vec == d // 8
    assert np.array_equal(bm.words(), oracle.bitmap_encode(docs, n_top))
    want_q = oracle.bitmap_query(qs, n_top)
    for r in range(8):
        assert np.array_equal(bm.build_query_bitmap_fp32(qs[r]), want_q[r])


@pytest.mark.parametrize("n,ties", [(3000, False), (60_000, False), (60_000, True)])
def test_search_and_top_m(engine, oracle, n, ties):
    d, n_top = 1024, 256
    docs, qs = _corpus(n, d, seed=11, ties=ties)
    bm = engine.Bitmap(d, n_top)
    bm.add(docs)
    words = oracle.bitmap_encode(docs, n_top)
    qb = oracle.bitmap_query(qs, n_top)
    for k in (1, 10, 100):
        s, i = bm.search(qs, k)
        ws, wi = oracle.bitmap_search(words, qb, k)
        assert np.array_equal(i, wi), k
        assert np.array_equal(s, ws), k
    for nq in (1, 5, 48):
        c = bm.top_m_candidates_batched(qs[:nq], 1000)
        assert np.array_equal(c, oracle.bitmap_top_m(words, qb[:nq], 1000)), nq
    assert np.array_equal(bm.top_m_candidates(qs[3], 50), oracle.bitmap_top_m(words, qb[3:4], 50)[0])
    chunked = bm.top_m_candidates_batched_chunked(qs[:10], 64, 3)
    assert np.array_equal(chunked, bm.top_m_candidates_batched(qs[:10], 64))


def test_subset_and_overlaps(engine, oracle):
    d, n_top = 256, 64
    docs, qs = _corpus(5000, d, seed=12, ties=True)
    bm = engine.Bitmap(d, n_top)
    bm.add(docs)
    words = oracle.bitmap_encode(docs, n_top)
    rng = np.random.default_rng(3)
    rows = rng.integers(0, 5000, size=900).astype(np.uint32)
    rows[:40] = rows[40:80]  # duplicates
    for r in range(6):
        qb = oracle.bitmap_query(qs[r], n_top)
        for k in (1, 7, 100, 900, 2000):
            s, i = bm.search_subset(qs[r], rows, min(k, 1024))
            ws, wi = oracle.bitmap_search_subset(words, qb, rows, min(k, 1024))
            assert np.array_equal(i, wi) and np.array_equal(s, ws), (r, k)
        ov = bm.body_overlap_scores_subset(qb[0], rows)
        assert np.array_equal(ov, oracle.and_popcount_all(words[rows], qb)[0])
    s, i = bm.search_subset(qs[0], np.zeros(0, np.uint32), 10)
    assert s.shape == (0,) and i.shape == (0,)
    with pytest.raises(ValueError):
        bm.body_overlap_scores_subset(oracle.bitmap_query(qs[0], n_top)[0], np.array([5000], np.uint32))


def test_persistence_and_abi_v1(engine, oracle, tmp_path):
    from ordvec_amd import _native as N

    # the reference FFI fixture (ordvec-ffi/src/lib.rs:1035-1049, 1227-1270)
    doc = np.zeros(64, np.float32)
    doc[:4] = 10 + np.arange(4)
    bm = engine.Bitmap(64, 4)
    bm.add(np.repeat(doc[None], 4, axis=0))
    path = tmp_path / "f.ovbm"
    bm.write(str(path))
    raw = path.read_bytes()
    assert raw[:17] == b"OVBM\x01" + (64).to_bytes(4, "little") + (4).to_bytes(4, "little") + (4).to_bytes(4, "little")
    assert raw[17:25] == (0xF).to_bytes(8, "little")
    lib = N.lib()
    h = C.c_void_p()
    assert lib.ordvec_index_load(str(path).encode(), 0, C.byref(h)) == N.OK
    info = N.IndexInfo()
    lib.ordvec_index_info_init(C.byref(info))
    assert lib.ordvec_index_info(h, C.byref(info)) == N.OK
    assert (info.kind, info.dim, info.n_top, info.vector_count, info.bytes_per_vec) == (2, 64, 4, 4, 8)
    params = N.SearchParams()
    lib.ordvec_search_params_init(C.byref(params))
    q = doc.copy()
    cands = (C.c_uint32 * 4)(3, 1, 1, 2)
    params.query = q.ctypes.data_as(C.POINTER(C.c_float))
    params.dim = 64
    params.k = 3
    params.candidate_rows = C.cast(cands, C.POINTER(C.c_uint32))
    params.candidate_count = 4
    hits = (N.Hit * 3)()
    ret = C.c_uint64()
    stats = N.SearchStats()
    lib.ordvec_search_stats_init(C.byref(stats))
    assert lib.ordvec_index_search(h, C.byref(params), hits, 3, C.byref(ret), C.byref(stats)) == N.OK
    assert ret.value == 3 and [hits[j].row_id for j in range(3)] == [1, 1, 2]
    assert [hits[j].score for j in range(3)] == [4.0, 4.0, 4.0]
    assert stats.kind == 2 and stats.candidate_count == 4 and stats.vectors_scored == 4
    params.candidate_rows = None
    params.candidate_count = 0
    params.k = 2
    assert lib.ordvec_index_search(h, C.byref(params), hits, 2, C.byref(ret), None) == N.OK
    assert ret.value == 2 and [hits[j].row_id for j in range(2)] == [0, 1]
    lib.ordvec_index_free(h)

    # round trip of a larger index through the Part 2 loader
    docs, qs = _corpus(5000, 512, seed=13)
    big = engine.Bitmap(512, 100)
    big.add(docs)
    p2 = tmp_path / "b.ovbm"
    big.write(str(p2))
    loaded = engine.Bitmap.load(str(p2))
    assert np.array_equal(loaded.words(), big.words())
    s1, i1 = big.search(qs, 10)
    s2, i2 = loaded.search(qs, 10)
    assert np.array_equal(i1, i2) and np.array_equal(s1, s2)


def test_device_tensor_path(engine, oracle):
    import torch

    d, n_top = 1024, 256
    docs, qs = _corpus(20_000, d, seed=14)
    bm = engine.Bitmap(d, n_top)
    bm.add(torch.tensor(docs, device="cuda"))
    assert np.array_equal(bm.words(), oracle.bitmap_encode(docs, n_top))
    s, i = bm.search(torch.tensor(qs, device="cuda"), 10)
    assert s.is_cuda and i.is_cuda
    ws, wi = oracle.bitmap_search(oracle.bitmap_encode(docs, n_top), oracle.bitmap_query(qs, n_top), 10)
    assert np.array_equal(i.cpu().numpy(), wi) and np.array_equal(s.cpu().numpy(), ws)
    bad = torch.tensor(qs[:2], device="cuda")
    bad[1, 5] = float("nan")
    with pytest.raises(ValueError):
        bm.search(bad, 10)
