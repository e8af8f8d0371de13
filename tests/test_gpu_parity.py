"""GPU parity: the HIP engine against the CPU oracle on the same seeded inputs.

Bar: codes, Hamming distances, candidate ids and top-k ids bit-exact; f32
scores bit-exact as well (the engine reproduces the reference tier's
summation order), with the 1e-5 relative tolerance of the contract asserted
on top so a failure message shows how far off a score is.
"""
import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu

RTOL = 1e-5  # contract tolerance for f32 asymmetric scores


def _assert_scores(got, want):
    np.testing.assert_allclose(got, want, rtol=RTOL, atol=1e-7)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "scores not bit-identical"


def test_c1_iid128_b1_codes_and_search(engine, oracle):
    """20,000 iid 128-d docs (seed 0), 256 queries (seed 1), b=1, k=10."""
    docs = S.iid_unit(20_000, 128, seed=0)
    qs = S.iid_unit(256, 128, seed=1)
    rq = engine.RankQuant(128, 1)
    rq.add(docs)
    assert len(rq) == 20_000
    assert np.array_equal(rq.codes(), oracle.rq_encode(docs, 1))
    s, i = rq.search_asymmetric(qs, 10)
    ws, wi = oracle.search_asymmetric(oracle.rq_encode(docs, 1), 128, 1, qs, 10)
    assert np.array_equal(i, wi)
    _assert_scores(s, ws)


@pytest.mark.parametrize("dim,bits", [(1024, 2), (1024, 4), (768, 4), (256, 2), (48, 2), (80, 4), (20, 2), (128, 4), (64, 1)])
def test_full_search_all_tiers(engine, oracle, dim, bits):
    gen = S.MixtureNP(dim, seed=7)
    docs = gen.docs(3000)
    qs, _ = gen.queries(docs, 24)
    rq = engine.RankQuant(dim, bits)
    rq.add(docs)
    codes = oracle.rq_encode(docs, bits)
    assert np.array_equal(rq.codes(), codes)
    for k in (1, 10, 100):
        s, i = rq.search_asymmetric(qs, k)
        ws, wi = oracle.search_asymmetric(codes, dim, bits, qs, k)
        assert np.array_equal(i, wi), (dim, bits, k)
        _assert_scores(s, ws)


def test_encode_ties_and_signed_zeros(engine, oracle):
    rows = np.array([
        [0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0],
        [1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0],
        [0.0, -0.0, 1.0, -0.0, 0.0, 2.0, -1.0, 0.0],
        [3.0, 1.0, 3.0, 2.0, 1.0, 5.0, 5.0, -7.0],
    ], dtype=np.float32)
    for bits in (1, 2):
        rq = engine.RankQuant(8, bits)
        rq.add(rows)
        assert np.array_equal(rq.codes(), oracle.rq_encode(rows, bits)), bits
    rows16 = np.concatenate([rows, rows[:, ::-1]], axis=1)  # b=4 needs dim % 16 == 0
    rq = engine.RankQuant(16, 4)
    rq.add(rows16)
    assert np.array_equal(rq.codes(), oracle.rq_encode(rows16, 4))
    rq = engine.RankQuant(8, 2)
    rq.add(rows[:1])
    assert rq.codes()[0].tolist() == [0x05, 0xAF]  # tests/persistence_compat.rs:130-157


def test_quickstart_kat(engine):
    docs = np.array([[8, 7, 6, 5, 4, 3, 2, 1], [1, 2, 3, 4, 5, 6, 7, 8], [8, 1, 7, 2, 6, 3, 5, 4]], np.float32)
    rq = engine.RankQuant(8, 1)
    rq.add(docs)
    s, i = rq.search_asymmetric(docs[:1], 1)
    assert int(i[0, 0]) == 0 and round(float(s[0, 0]), 3) == 0.396


def test_zero_query_and_tie_kats(engine):
    """tests/determinism_contract.rs:15-125."""
    docs = np.ones((12, 64), np.float32)
    rq = engine.RankQuant(64, 2)
    rq.add(docs)
    z = np.zeros(64, np.float32)
    s, i = rq.search_asymmetric(z[None], 4)
    assert i[0].tolist() == [0, 1, 2, 3] and np.all(s == 0.0)
    s, i = rq.search_asymmetric_subset(z, np.array([9, 3, 7, 1]), 2)
    assert i.tolist() == [1, 3] and s.tolist() == [0.0, 0.0]
    s, i = rq.search_asymmetric_subset(z, np.array([7, 8, 7]), 2)
    assert i.tolist() == [7, 7]
    s, i = rq.search_asymmetric_subset_batched(z[None], np.array([0, 4]), np.array([7, 7, 3, 7]), 4)
    assert i[0].tolist() == [3, 7, 7, 7] and s[0].tolist() == [0.0] * 4
    sb = engine.SignBitmap(64)
    sb.add(np.ones((10, 64), np.float32))
    assert sb.top_m_candidates(np.ones(64, np.float32), 5).tolist() == [0, 1, 2, 3, 4]
    c = sb.top_m_candidates_batched(np.ones((2, 64), np.float32), 5)
    assert c.tolist() == [[0, 1, 2, 3, 4]] * 2


def test_empty_and_zero_k_shapes(engine):
    rq = engine.RankQuant(64, 2)
    s, i = rq.search_asymmetric(np.ones((1, 64), np.float32), 10)
    assert s.shape == (1, 0) and i.shape == (1, 0)
    rq.add(np.ones((2, 64), np.float32))
    s, i = rq.search_asymmetric(np.ones((1, 64), np.float32), 0)
    assert s.shape == (1, 0)
    sb = engine.SignBitmap(64)
    assert sb.top_m_candidates(np.ones(64, np.float32), 10).shape == (0,)


@pytest.mark.parametrize("dim,n,nq,m", [(1024, 50_000, 16, 1000), (768, 20_011, 40, 300), (64, 5000, 3, 64),
                                        (128, 777, 9, 1000), (1024, 4096, 1, 4096), (192, 3001, 20, 1),
                                        # matrix-core scan shapes (>= 32 queries, over 8192 documents with
                                        # sparse survivors): full and partial query tiles, a ragged last
                                        # document block, one- and eight-word rows
                                        (1024, 30_001, 100, 1000), (64, 40_000, 70, 64), (512, 12_345, 32, 500),
                                        # batches over tiny corpora (every document survives: VALU scan)
                                        (64, 5000, 70, 64), (192, 7001, 33, 300), (1024, 65, 64, 10),
                                        # VALU scan without a word-count specialisation
                                        (2048, 9000, 5, 500), (320, 10_000, 1, 700)])
def test_sign_top_m_exact(engine, oracle, dim, n, nq, m):
    gen = S.MixtureNP(dim, seed=11)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, nq)
    sb = engine.SignBitmap(dim)
    sb.add(docs)
    bm = oracle.sign_encode(docs)
    assert np.array_equal(sb.words(), bm)
    qb = oracle.sign_encode(qs)
    c, h = sb.top_m_candidates_batched(qs, m, return_hamming=True)
    want = oracle.sign_top_m_bruteforce(bm, qb, m)
    assert np.array_equal(c, want)
    for qi in range(nq):
        assert np.array_equal(h[qi], oracle.hamming_all(bm, qb[qi])[want[qi]])


def test_sign_top_m_tie_heavy(engine, oracle):
    """Boundary tiers far larger than the tie buffer: duplicate-run corpora
    (tests/tiled_candgen.rs:97-127)."""
    rng = np.random.default_rng(5)
    base = rng.standard_normal((7, 256)).astype(np.float32)
    docs = base[rng.integers(0, 7, size=30_000)]
    qs = base[:3] + 0.01 * rng.standard_normal((3, 256)).astype(np.float32)
    sb = engine.SignBitmap(256)
    sb.add(docs)
    bm = oracle.sign_encode(docs)
    c = sb.top_m_candidates_batched(qs, 2500)
    assert np.array_equal(c, oracle.sign_top_m_bruteforce(bm, oracle.sign_encode(qs), 2500))
    ident = engine.SignBitmap(64)
    ident.add(np.ones((9000, 64), np.float32))
    assert ident.top_m_candidates(np.ones(64, np.float32), 777).tolist() == list(range(777))


def test_two_stage_mixture_exact(engine, oracle):
    dim, n = 1024, 60_000
    gen = S.MixtureNP(dim, seed=2)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 48)
    sb = engine.SignBitmap(dim)
    sb.add(docs)
    rq = engine.RankQuant(dim, 2)
    rq.add(docs)
    s, i = engine.two_stage_search(sb, rq, qs, 1000, 10)
    cand = oracle.sign_top_m(oracle.sign_encode(docs), oracle.sign_encode(qs), 1000)
    offs = np.arange(49, dtype=np.uint64) * 1000
    ws, wi = oracle.subset_batched(oracle.rq_encode(docs, 2), dim, 2, qs, offs, cand.reshape(-1), 10)
    assert np.array_equal(i, wi)
    _assert_scores(s, ws)


def test_subset_batched_csr_random(engine, oracle):
    rng = np.random.default_rng(9)
    dim, n = 256, 4000
    gen = S.MixtureNP(dim, seed=3)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 30)
    lens = rng.integers(0, 700, size=30)
    lens[3] = 0
    lens[7] = 1
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    cands = rng.integers(0, n, size=int(offs[-1])).astype(np.uint32)
    cands[:50] = cands[50:100]  # duplicates
    for bits in (1, 2, 4):
        rq = engine.RankQuant(dim, bits)
        rq.add(docs)
        for k in (1, 10, 64):
            s, i = rq.search_asymmetric_subset_batched(qs, offs, cands, k)
            ws, wi = oracle.subset_batched(oracle.rq_encode(docs, bits), dim, bits, qs, offs, cands, k)
            assert np.array_equal(i, wi), (bits, k)
            _assert_scores(s, ws)
        # one query whose row repeats every candidate in another split's slice:
        # each of the best rows comes back twice
        row = np.concatenate([np.arange(300), np.arange(300)[::-1]]).astype(np.uint32)
        one = np.array([0, len(row)], np.uint64)
        for k in (7, 10, 32):
            s, i = rq.search_asymmetric_subset_batched(qs[:1], one, row, k)
            ws, wi = oracle.subset_batched(oracle.rq_encode(docs, bits), dim, bits, qs[:1], one, row, k)
            assert np.array_equal(i, wi), (bits, k)
            assert i[0, 0] == i[0, 1]
            _assert_scores(s, ws)


def test_full_candidates_equal_full_search(engine):
    """tests/index/two_stage.rs:238-258."""
    gen = S.MixtureNP(128, seed=4)
    docs = gen.docs(900)
    qs, _ = gen.queries(docs, 5)
    rq = engine.RankQuant(128, 2)
    rq.add(docs)
    fs, fi = rq.search_asymmetric(qs, 10)
    for qi in range(5):
        s, i = rq.search_asymmetric_subset(qs[qi], np.random.default_rng(qi).permutation(900), 10)
        assert np.array_equal(i, fi[qi]) and np.array_equal(s, fs[qi])


def test_fastscan_exact(engine, oracle):
    dim, n = 256, 5003
    gen = S.MixtureNP(dim, seed=8)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 16)
    fs = engine.RankQuantFastscan(dim)
    fs.add(docs)
    buckets = np.stack([oracle.unpack_buckets(r, dim, 2) for r in oracle.rq_encode(docs, 2)])
    packed = oracle.fs_pack(buckets, dim)
    for k in (1, 10, 100):
        s, i = fs.search(qs, k)
        ws, wi = oracle.fs_search(packed, n, dim, qs, k)
        assert np.array_equal(i, wi)
        _assert_scores(s, ws)


def test_persistence_roundtrip(engine, oracle, tmp_path):
    gen = S.MixtureNP(128, seed=12)
    docs = gen.docs(1000)
    qs, _ = gen.queries(docs, 4)
    rq = engine.RankQuant(128, 4)
    rq.add(docs)
    rq.write(tmp_path / "a.ovrq")
    raw = (tmp_path / "a.ovrq").read_bytes()
    assert raw[:14] == b"OVRQ\x01\x04" + (128).to_bytes(4, "little") + (1000).to_bytes(4, "little")
    assert raw[14:] == oracle.rq_encode(docs, 4).tobytes()
    rq2 = engine.RankQuant.load(tmp_path / "a.ovrq")
    assert np.array_equal(rq2.search_asymmetric(qs, 7)[1], rq.search_asymmetric(qs, 7)[1])
    sb = engine.SignBitmap(128)
    sb.add(docs)
    sb.write(tmp_path / "a.ovsb")
    raw = (tmp_path / "a.ovsb").read_bytes()
    assert raw[13:] == oracle.sign_encode(docs).tobytes()
    sb2 = engine.SignBitmap.load(tmp_path / "a.ovsb")
    assert np.array_equal(sb2.words(), sb.words())
    fs = engine.RankQuantFastscan(128)
    fs.add(docs)
    fs.write(tmp_path / "a.ovfs")
    fs2 = engine.RankQuantFastscan.load(tmp_path / "a.ovfs")
    assert np.array_equal(fs2.search(qs, 5)[1], fs.search(qs, 5)[1])


def test_device_tensor_path(engine, oracle):
    import torch

    gen = S.MixtureNP(1024, seed=21)
    docs = gen.docs(10_000)
    qs, _ = gen.queries(docs, 64)
    td, tq = torch.from_numpy(docs).cuda(), torch.from_numpy(qs).cuda()
    sb = engine.SignBitmap(1024)
    sb.add(td)
    rq = engine.RankQuant(1024, 2)
    rq.add(td)
    s, i = engine.two_stage_search(sb, rq, tq, 500, 10)
    torch.cuda.synchronize()
    hs, hi = engine.two_stage_search(sb, rq, qs, 500, 10)
    assert np.array_equal(i.cpu().numpy(), hi)
    assert np.array_equal(s.cpu().numpy(), hs)
    assert np.array_equal(rq.codes(), oracle.rq_encode(docs, 2))


def test_two_stage_large_batch_and_mixed_fallback(engine, oracle):
    """>=128 queries (single-split rerank) and a batch where only some
    queries overflow the survivor buffer (exact recomputation of those rows)."""
    rng = np.random.default_rng(17)
    dim = 256
    gen = S.MixtureNP(dim, seed=19)
    docs = gen.docs(40_000)
    dup = docs[123].copy()
    docs[1000:21000] = dup  # 20k identical docs: queries near them tie massively
    qs, _ = gen.queries(docs, 200)
    qs[::7] = dup + 0.001 * rng.standard_normal((len(qs[::7]), dim)).astype(np.float32)
    sb = engine.SignBitmap(dim)
    sb.add(docs)
    rq = engine.RankQuant(dim, 2)
    rq.add(docs)
    s, i = engine.two_stage_search(sb, rq, qs, 700, 10)
    bm, qb = oracle.sign_encode(docs), oracle.sign_encode(qs)
    cand = oracle.sign_top_m(bm, qb, 700)
    assert np.array_equal(sb.top_m_candidates_batched(qs, 700), cand)
    offs = np.arange(len(qs) + 1, dtype=np.uint64) * 700
    ws, wi = oracle.subset_batched(oracle.rq_encode(docs, 2), dim, 2, qs, offs, cand.reshape(-1), 10)
    assert np.array_equal(i, wi)
    _assert_scores(s, ws)
    # device tensors: the call returns once stage 1's status is back, with the
    # rerank (and the exact rows' second rerank) still in stream order
    import torch

    e0 = engine.stage1_exact_queries()
    ds, di = engine.two_stage_search(sb, rq, torch.from_numpy(qs).cuda(), 700, 10)
    assert engine.stage1_exact_queries() > e0
    assert np.array_equal(di.cpu().numpy(), wi)
    assert np.array_equal(ds.cpu().numpy(), s)


def test_nonfinite_device_query_is_rejected(engine):
    import torch

    sb = engine.SignBitmap(64)
    rq = engine.RankQuant(64, 2)
    x = S.iid_unit(500, 64, 0)
    sb.add(x)
    rq.add(x)
    q = torch.from_numpy(x[:3].copy()).cuda()
    q[1, 5] = float("nan")
    with pytest.raises(ValueError):
        engine.two_stage_search(sb, rq, q, 50, 5)
    with pytest.raises(ValueError):
        rq.search_asymmetric(q, 5)
    with pytest.raises(ValueError):
        engine.two_stage_search(sb, rq, x[:3] * np.array([1, np.inf, 1], np.float32)[:, None], 50, 5)
    big = torch.from_numpy(x[:100].copy()).cuda()  # a batch that waits for stage 1's status only
    big[77, 3] = float("inf")
    with pytest.raises(ValueError):
        engine.two_stage_search(sb, rq, big, 50, 5)
    big[77, 3] = 0.5
    s, i = engine.two_stage_search(sb, rq, big, 50, 5)
    s1, i1 = engine.two_stage_search(sb, rq, big[77:78], 50, 5)
    assert torch.equal(i[77:78], i1) and torch.equal(s[77:78], s1)


@pytest.fixture(scope="module")
def mixture_200k(engine):
    """200,000 x 1024 mixture corpus (seed 21) with its sign bitmap and b=2
    RankQuant index resident, shared by the C2-shape tests below."""
    d, n = 1024, 200_000
    gen = S.MixtureNP(d, seed=21)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 1100)
    sb = engine.SignBitmap(d)
    sb.add(docs)
    rq = engine.RankQuant(d, 2)
    rq.add(docs)
    return docs, qs, sb, rq


def test_fast_stage1_path_needs_no_exact_fallback_on_partial_tiles(engine, oracle, mixture_200k):
    """Batches that fill their matrix-core query tiles (32, 64 or 128 queries)
    partially, and the one- and few-query VALU scan whose workgroups derive
    the sampled threshold themselves (the sample is strided at 200k docs); on
    ordinary data every query must finish on the fast path (a wrong sampled
    threshold or survivor list would still give exact results, but only
    through the exact fallback, which the counter exposes)."""
    docs, qs, sb, _ = mixture_200k
    words = sb.words()
    for nq in (1, 3, 2, 8, 31, 32, 33, 48, 64, 65, 100, 128, 129, 200):
        e0 = engine.stage1_exact_queries()
        got = sb.top_m_candidates_batched(qs[:nq], 1000)
        assert engine.stage1_exact_queries() == e0, nq
        assert np.array_equal(got, oracle.sign_top_m(words, oracle.sign_encode(qs[:nq]), 1000, 8)), nq


@pytest.mark.parametrize("n,m", [(4096, 100), (8192, 1000), (3000, 3000), (9000, 1000), (50000, 2000)])
def test_dense_survivor_batches_stay_on_the_fast_path(engine, oracle, n, m):
    """Batches whose stage 1 keeps many survivors per document block: an index
    of at most 8192 documents (the survivor buffer's size) admits every
    document, and m near n / 10 keeps about one in ten.  The matrix-core
    scan's LDS survivor lists cannot hold that many per block, so such plans
    take the vector-ALU scan; every query must finish on the fast path and
    match the oracle, also with m equal to the corpus size."""
    d = 1024
    gen = S.MixtureNP(d, seed=31)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 200)
    sb = engine.SignBitmap(d)
    rq = engine.RankQuant(d, 2)
    engine.two_stage_add(sb, rq, docs)
    words = sb.words()
    e0 = engine.stage1_exact_queries()
    got = sb.top_m_candidates_batched(qs, m)
    s, i = engine.two_stage_search(sb, rq, qs, m, 10)
    assert engine.stage1_exact_queries() == e0
    assert np.array_equal(got, oracle.sign_top_m(words, oracle.sign_encode(qs), m, 8))
    ws, wi = _two_stage_oracle(oracle, rq.codes(), words, qs, m, 10)
    assert np.array_equal(i, wi)
    assert np.array_equal(s.view(np.uint32), ws.view(np.uint32))


def test_large_m_batches(engine, oracle, mixture_200k):
    """m near the 4096 limit leaves little room between m and the 8192-key
    survivor buffer; the sampled threshold aims at the geometric mean of the
    two.  Up to m = 3000 every query of a 256-query batch finishes on the fast
    path; at m = 4000 this corpus's Hamming bins near the threshold hold
    thousands of documents each (the count jumps past m and past 8192 in one
    distance step for most queries), so those queries take the exact path,
    and the candidates must match the oracle either way."""
    docs, qs, sb, _ = mixture_200k
    words = sb.words()
    qb = oracle.sign_encode(qs[:256])
    for m in (2500, 3000, 4000):
        e0 = engine.stage1_exact_queries()
        got = sb.top_m_candidates_batched(qs[:256], m)
        if m <= 3000:
            assert engine.stage1_exact_queries() == e0, m
        assert np.array_equal(got, oracle.sign_top_m(words, qb, m, 8)), m


def _two_stage_oracle(oracle, docs_codes, words, qs, m, k):
    dim = qs.shape[1]
    cand = oracle.sign_top_m(words, oracle.sign_encode(qs), m, 8)
    offs = np.arange(len(qs) + 1, dtype=np.uint64) * cand.shape[1]
    return oracle.subset_batched(docs_codes, dim, 2, qs, offs, cand.reshape(-1), k, 8)


def test_c2_one_query_calls_match_oracle(engine, oracle, mixture_200k):
    """The plain one-query-per-call two-stage path at the C2 shape (1024-d,
    m=1000 < n, k=10): VALU scan with the query prep and threshold inside its
    grid, split rerank with the one-wave merge; host arrays and device
    tensors, every query on the fast stage-1 path."""
    import torch

    docs, qs, sb, rq = mixture_200k
    codes, words = rq.codes(), sb.words()
    sel = qs[1000:1012]
    ws, wi = _two_stage_oracle(oracle, codes, words, sel, 1000, 10)
    e0 = engine.stage1_exact_queries()
    for j in range(len(sel)):
        s, i = engine.two_stage_search(sb, rq, sel[j:j + 1], 1000, 10)
        assert np.array_equal(i[0], wi[j]), j
        _assert_scores(s[0], ws[j])
        tq = torch.from_numpy(sel[j:j + 1].copy()).cuda()
        ds, di = engine.two_stage_search(sb, rq, tq, 1000, 10)
        assert np.array_equal(di.cpu().numpy()[0], wi[j]), j
        _assert_scores(ds.cpu().numpy()[0], ws[j])
    assert engine.stage1_exact_queries() == e0
    # the same queries in one call give the same rows
    s, i = engine.two_stage_search(sb, rq, sel, 1000, 10)
    assert np.array_equal(i, wi)
    _assert_scores(s, ws)


def test_c2_batch_1024_matches_oracle(engine, oracle, mixture_200k):
    """A 1024-query two-stage batch (128-query matrix-core tiles, batched
    rerank, the early stage-1 status wait on device tensors) against the
    oracle at 200k x 1024, m=1000, k=10."""
    import torch

    docs, qs, sb, rq = mixture_200k
    batch = qs[:1024]
    ws, wi = _two_stage_oracle(oracle, rq.codes(), sb.words(), batch, 1000, 10)
    e0 = engine.stage1_exact_queries()
    s, i = engine.two_stage_search(sb, rq, batch, 1000, 10)
    assert np.array_equal(i, wi)
    _assert_scores(s, ws)
    ds, di = engine.two_stage_search(sb, rq, torch.from_numpy(batch.copy()).cuda(), 1000, 10)
    assert np.array_equal(di.cpu().numpy(), wi)
    _assert_scores(ds.cpu().numpy(), ws)
    assert engine.stage1_exact_queries() == e0
