"""Searches past the fast paths' limits against the oracle: stage-1 candidate
counts above 4096 (every query on the exact pipeline's ordered scatter), k
above 1024 (whole candidate rows sorted), the two-stage candidate policy and
RankQuant's sign-probe helpers (src/quant.rs:235-262,1453-1508), and ABI v1
searches with a large k.  Bar: ids and distances bit-exact, score bits equal."""
import ctypes as C

import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _same_scores(got, want):
    assert np.array_equal(np.asarray(got).view(np.uint32), np.asarray(want).view(np.uint32))


def _uniform_corpus(seed, n, d):
    """The reference's two-stage test corpus shape (tests/index/main.rs:39-48:
    uniform [-1, 1), D = 128, N = 256), from this build's own generator."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, size=(n, d)).astype(np.float32)


@pytest.fixture(scope="module")
def two_stage_128(engine):
    docs = _uniform_corpus(15_001, 256, 128)
    sb = engine.SignBitmap(128)
    sb.add(docs)
    rq = engine.RankQuant(128, 2)
    rq.add(docs)
    return docs, sb, rq


def test_sign_probe_helper_matches_manual_candidates(engine, oracle, two_stage_128):
    """tests/index/two_stage.rs:177-204: helper == top_m_candidates + subset."""
    docs, sb, rq = two_stage_128
    q = _uniform_corpus(15_004, 1, 128)[0]
    policy = engine.TwoStageCandidatePolicy(min_candidates=0, k_multiplier=4, max_candidates=37)
    k = 5
    m = policy.candidate_count(k, len(rq))
    assert m == 20
    cands = sb.top_m_candidates(q, m)
    ms, mi = rq.search_asymmetric_subset(q, cands, k)
    hs, hi = rq.try_search_with_sign_probe(sb, q, k, policy)
    assert np.array_equal(hi, mi)
    _same_scores(hs, ms)
    ws, wi = oracle.search_subset(oracle.rq_encode(docs, 2), 128, 2, q,
                                  oracle.sign_top_m(oracle.sign_encode(docs), oracle.sign_encode(q[None]), m)[0], k)
    assert np.array_equal(hi, wi)
    _same_scores(hs, ws)
    dt = rq.try_search_with_sign_probe(sb, q, k)
    dp = rq.search_with_sign_probe(sb, q, k)
    assert np.array_equal(dt[1], dp[1])
    # a cap below k: min(k, m) results
    s, i = rq.try_search_with_sign_probe(sb, q, 50, engine.TwoStageCandidatePolicy(0, 1, 7))
    assert len(i) == 7 and len(s) == 7
    cands7 = sb.top_m_candidates(q, 7)
    assert np.array_equal(i, rq.search_asymmetric_subset(q, cands7, 50)[1])
    # k = 0 and the default policy at k = 10 (m = N)
    assert len(rq.try_search_with_sign_probe(sb, q, 0)[1]) == 0
    s, i = rq.search_with_sign_probe(sb, q, 10)
    fs, fi = rq.search_asymmetric(q[None], 10)
    assert np.array_equal(i, fi[0])  # m = N: the full candidate set equals full search


def test_sign_probe_helper_validates_probe_and_query(engine, two_stage_128):
    """tests/index/two_stage.rs:206-245: wrong probe dim, short probe, query of
    the wrong length and a non-finite query are errors."""
    from ordvec_amd import _native as N

    docs, sb, rq = two_stage_128
    q = _uniform_corpus(15_005, 1, 128)[0]
    wrong_dim = engine.SignBitmap(256)
    with pytest.raises(ValueError, match="sign_probe.dim"):
        rq.try_search_with_sign_probe(wrong_dim, q, 5)
    short = engine.SignBitmap(128)
    short.add(_uniform_corpus(15_006, 255, 128))
    with pytest.raises(ValueError, match="sign_probe.len"):
        rq.try_search_with_sign_probe(short, q, 5)
    with pytest.raises(N.OrdvecError) as e:
        rq.try_search_with_sign_probe(sb, q[:127], 5)
    assert e.value.status == N.DIM_MISMATCH
    bad = q.copy()
    bad[0] = np.nan
    with pytest.raises(N.OrdvecError) as e:
        rq.try_search_with_sign_probe(sb, bad, 5)
    assert e.value.status == N.NONFINITE_QUERY


@pytest.fixture(scope="module")
def corpus_256(engine):
    gen = S.MixtureNP(256, seed=41)
    docs = gen.docs(80_000)
    qs, _ = gen.queries(docs, 40)
    sb = engine.SignBitmap(256)
    sb.add(docs)
    rq = engine.RankQuant(256, 2)
    rq.add(docs)
    return docs, qs, sb, rq


@pytest.mark.parametrize("m", [4097, 5000, 32_768])
def test_top_m_above_4096_is_exact(engine, oracle, corpus_256, m):
    docs, qs, sb, _ = corpus_256
    nq = 5 if m > 20_000 else 9
    c, h = sb.top_m_candidates_batched(qs[:nq], m, return_hamming=True)
    bm, qb = oracle.sign_encode(docs), oracle.sign_encode(qs[:nq])
    want = oracle.sign_top_m_bruteforce(bm, qb, m)
    assert np.array_equal(c, want)
    for qi in range(nq):
        assert np.array_equal(h[qi], oracle.hamming_all(bm, qb[qi])[want[qi]])


def test_top_m_above_4096_ties_and_m_equal_n(engine, oracle):
    """Duplicate-run corpus (every boundary tier huge) and m = n."""
    rng = np.random.default_rng(44)
    base = rng.standard_normal((5, 64)).astype(np.float32)
    docs = base[rng.integers(0, 5, size=12_000)]
    qs = base[:3] + 0.05 * rng.standard_normal((3, 64)).astype(np.float32)
    sb = engine.SignBitmap(64)
    sb.add(docs)
    bm, qb = oracle.sign_encode(docs), oracle.sign_encode(qs)
    for m in (6000, 12_000, 50_000):
        assert np.array_equal(sb.top_m_candidates_batched(qs, m), oracle.sign_top_m_bruteforce(bm, qb, m)), m


@pytest.mark.parametrize("m,k", [(5000, 10), (32_768, 10), (2000, 1500), (40_000, 1300)])
def test_two_stage_large_m_and_k(engine, oracle, corpus_256, m, k):
    docs, qs, sb, rq = corpus_256
    nq = 6
    s, i = engine.two_stage_search(sb, rq, qs[:nq], m, k)
    cand = oracle.sign_top_m(oracle.sign_encode(docs), oracle.sign_encode(qs[:nq]), m, 8)
    offs = np.arange(nq + 1, dtype=np.uint64) * cand.shape[1]
    ws, wi = oracle.subset_batched(rq.codes(), 256, 2, qs[:nq], offs, cand.reshape(-1), k, 8)
    assert np.array_equal(i, wi)
    _same_scores(s, ws)


def test_two_stage_default_policy_large_k(engine, oracle, corpus_256):
    """The default policy at k = 1024 asks for m = 32,768 candidates."""
    docs, qs, sb, rq = corpus_256
    pol = engine.TwoStageCandidatePolicy()
    assert pol.candidate_count(1024, len(rq)) == 32_768
    s, i = engine.two_stage_search_policy(sb, rq, qs[:4], 1024)
    ds, di = engine.two_stage_search(sb, rq, qs[:4], 32_768, 1024)
    assert np.array_equal(i, di)
    _same_scores(s, ds)
    hs, hi = rq.search_with_sign_probe(sb, qs[7], 1100)  # m = 35,200
    cand = oracle.sign_top_m(oracle.sign_encode(docs), oracle.sign_encode(qs[7:8]), 35_200, 8)[0]
    ws, wi = oracle.search_subset(rq.codes(), 256, 2, qs[7], cand, 1100)
    assert np.array_equal(hi, wi)
    _same_scores(hs, ws)


@pytest.mark.parametrize("dim,bits", [(256, 2), (256, 4), (128, 1), (1024, 2)])
def test_full_search_k_above_1024(engine, oracle, dim, bits):
    gen = S.MixtureNP(dim, seed=45)
    docs = gen.docs(9000)
    qs, _ = gen.queries(docs, 5)
    rq = engine.RankQuant(dim, bits)
    rq.add(docs)
    codes = oracle.rq_encode(docs, bits)
    for k in (1025, 3000, 9000, 20_000):
        s, i = rq.search_asymmetric(qs, k)
        ws, wi = oracle.search_asymmetric(codes, dim, bits, qs, k)
        assert np.array_equal(i, wi), k
        _same_scores(s, ws)
    # the same rows through device tensors
    import torch

    ds, di = rq.search_asymmetric(torch.from_numpy(qs).cuda(), 2500)
    ws, wi = oracle.search_asymmetric(codes, dim, bits, qs, 2500)
    assert np.array_equal(di.cpu().numpy(), wi)
    _same_scores(ds.cpu().numpy(), ws)


def test_subset_k_above_1024_with_duplicates(engine, oracle):
    rng = np.random.default_rng(46)
    gen = S.MixtureNP(256, seed=46)
    docs = gen.docs(6000)
    qs, _ = gen.queries(docs, 7)
    rq = engine.RankQuant(256, 2)
    rq.add(docs)
    codes = oracle.rq_encode(docs, 2)
    lens = np.array([0, 1, 1500, 5999, 3000, 2047, 4100])
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    cands = rng.integers(0, 6000, size=int(offs[-1])).astype(np.uint32)
    cands[10_000:10_500] = cands[9000:9500]  # duplicates
    for k in (1025, 2500, 7000):
        s, i = rq.search_asymmetric_subset_batched(qs, offs, cands, k)
        ws, wi = oracle.subset_batched(codes, 256, 2, qs, offs, cands, k)
        assert np.array_equal(i, wi), k
        _same_scores(s, ws)
    rows = cands[:3000]
    s, i = rq.search_asymmetric_subset(qs[2], rows, 2000)
    ws, wi = oracle.search_subset(codes, 256, 2, qs[2], rows, 2000)
    assert np.array_equal(i, wi)
    _same_scores(s, ws)


def test_bitmap_search_and_subset_k_above_limits(engine, oracle):
    gen = S.MixtureNP(256, seed=47)
    docs = gen.docs(9000)
    qs, _ = gen.queries(docs, 4)
    bmx = engine.Bitmap(256, 64)
    bmx.add(docs)
    bm = oracle.bitmap_encode(docs, 64)
    qb = np.stack([oracle.bitmap_query(q, 64) for q in qs])
    for k in (4097, 8000):
        s, i = bmx.search(qs, k)
        ws, wi = oracle.bitmap_search(bm, qb, k)
        assert np.array_equal(i, wi), k
        assert np.array_equal(s, ws)
    rows = np.random.default_rng(48).integers(0, 9000, size=5000).astype(np.uint32)
    s, i = bmx.search_subset(qs[1], rows, 3000)
    ws, wi = oracle.bitmap_search_subset(bm, qb[1], rows, 3000)
    assert np.array_equal(i, wi)
    assert np.array_equal(s, ws)


def test_abi_v1_rankquant_search_with_k_above_1024(engine, oracle, tmp_path):
    """ordvec_index_search answers every k the reference answers."""
    from ordvec_amd import _native as N

    gen = S.MixtureNP(128, seed=49)
    docs = gen.docs(5000)
    q = gen.queries(docs, 1)[0][0]
    rq = engine.RankQuant(128, 2)
    rq.add(docs)
    path = str(tmp_path / "big_k.ovrq")
    rq.write(path)
    lib = N.lib()
    h = C.c_void_p()
    N.check(lib.ordvec_index_load(path.encode(), 0, C.byref(h)))
    try:
        codes = oracle.rq_encode(docs, 2)
        for rows, k in ((None, 4000), (np.arange(4999, -1, -1, dtype=np.uint32)[:3500], 2000)):
            p = N.SearchParams()
            lib.ordvec_search_params_init(C.byref(p))
            p.query = q.ctypes.data_as(C.POINTER(C.c_float))
            p.dim = 128
            p.k = k
            if rows is not None:
                p.candidate_rows = rows.ctypes.data_as(C.POINTER(C.c_uint32))
                p.candidate_count = len(rows)
            hits = (N.Hit * k)()
            got = C.c_uint64()
            N.check(lib.ordvec_index_search(h, C.byref(p), hits, k, C.byref(got), None))
            assert got.value == k
            ids = np.array([hits[j].row_id for j in range(k)], np.int64)
            sc = np.array([hits[j].score for j in range(k)], np.float32)
            if rows is None:
                ws, wi = oracle.search_asymmetric(codes, 128, 2, q[None], k)
                ws, wi = ws[0], wi[0]
            else:
                ws, wi = oracle.search_subset(codes, 128, 2, q, rows, k)
            assert np.array_equal(ids, wi)
            _same_scores(sc, ws)
    finally:
        lib.ordvec_index_free(h)
