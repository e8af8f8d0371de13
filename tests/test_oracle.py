"""Pins the CPU oracle before it is trusted as the checker.

Each check comes from the reference's own tests (KATs / fixtures), a closed
form, or an independent formulation (brute force, float64, numpy), chosen so
that a plausible slip in the restatement (dropped term, wrong index, wrong
tie rule, wrong summation order) fails one of them.
"""
import numpy as np
import pytest

from ordvec_amd import synthetic as S


# --- src/rank.rs KATs (src/rank.rs:591-643,777-820) -------------------------

def test_rank_transform_kats(oracle):
    assert oracle.rank_transform([3, 1, 4, 1.5, 5, 9, 2, 6]).tolist() == [3, 0, 4, 1, 5, 7, 2, 6]
    assert oracle.rank_transform([1, 1, 1, 1]).tolist() == [0, 1, 2, 3]
    assert oracle.rank_transform([3, 1, 3, 2, 1]).tolist() == [3, 0, 4, 2, 1]
    assert oracle.rank_transform(np.array([0.0, -0.0, 1.0, -0.0, 0.0], np.float32)).tolist() == [0, 1, 4, 2, 3]


def test_rank_transform_matches_numpy_stable_argsort(oracle):
    rng = np.random.default_rng(0)
    for d in (2, 7, 64, 257, 1024):
        v = rng.integers(-5, 5, size=d).astype(np.float32)  # heavy ties
        want = np.argsort(np.argsort(v, kind="stable"), kind="stable")
        assert np.array_equal(oracle.rank_transform(v), want)
        assert np.array_equal(oracle.rank_transform_qsort(v), want)


def test_bucket_pack_centre_norm(oracle):
    assert [oracle.rank_to_bucket(r, 1024, 2) for r in (0, 255, 256, 511, 512, 1023)] == [0, 0, 1, 1, 2, 3]
    b = np.array([0, 0, 1, 1, 2, 2, 3, 3], np.uint8)
    assert oracle.pack_buckets(b, 2).tolist() == [0x05, 0xAF]
    assert np.array_equal(oracle.unpack_buckets(oracle.pack_buckets(b, 2), 8, 2), b)
    assert [oracle.bucket_centre(i, 2) for i in range(4)] == [-1.5, -0.5, 0.5, 1.5]
    assert [oracle.bucket_centre(i, 1) for i in range(2)] == [-0.5, 0.5]
    for d, bits in ((1024, 2), (768, 4), (128, 1)):
        nb = 1 << bits
        closed = np.float32(np.sqrt(d * (nb * nb - 1) / 12.0))
        assert oracle.rankquant_norm(d, bits) == closed
        # equals the empirical norm of any permutation's centred buckets
        c = np.array([oracle.bucket_centre(oracle.rank_to_bucket(r, d, bits), bits) for r in range(d)], np.float64)
        assert abs(np.sqrt((c * c).sum()) - float(closed)) < 1e-4


def test_ovrq_payload_fixture(oracle):
    """tests/persistence_compat.rs:130-157: RankQuant(8,2).add([0..7]) -> 05 AF."""
    assert oracle.rq_encode(np.arange(8, dtype=np.float32)[None], 2).tolist() == [[0x05, 0xAF]]


def test_ovsb_payload_fixture(oracle):
    """tests/persistence_compat.rs:190-222: bits 0 and 63 set."""
    v = -np.ones((1, 64), np.float32)
    v[0, 0] = v[0, 63] = 1.0
    assert int(oracle.sign_encode(v)[0, 0]) == 0x8000000000000001
    z = np.zeros((1, 64), np.float32)
    assert int(oracle.sign_encode(z)[0, 0]) == 0  # threshold: 0.0 is unset


def test_encode_bucket_composition(oracle):
    x = S.iid_unit(50, 1024, 3)
    for bits in (1, 2, 4):
        codes = oracle.rq_encode(x, bits)
        for r in codes[:5]:
            u = oracle.unpack_buckets(r, 1024, bits)
            assert np.all(np.bincount(u, minlength=1 << bits) == 1024 >> bits)


# --- scoring (src/quant_kernels.rs) -------------------------------------------

def test_quickstart_kat(oracle):
    """ordvec-python/tests/test_rank_quant.py:120-137: id 0, score 0.396."""
    docs = np.array([[8, 7, 6, 5, 4, 3, 2, 1], [1, 2, 3, 4, 5, 6, 7, 8], [8, 1, 7, 2, 6, 3, 5, 4]], np.float32)
    s, i = oracle.search_asymmetric(oracle.rq_encode(docs, 1), 8, 1, docs[:1], 1)
    assert int(i[0, 0]) == 0 and round(float(s[0, 0]), 3) == 0.396
    assert abs(float(s[0, 0]) - 8 / np.sqrt(204) / np.sqrt(2)) < 1e-6  # closed form


@pytest.mark.parametrize("dim,bits", [(1024, 2), (1024, 4), (48, 2), (80, 4), (20, 2), (128, 1)])
def test_scores_match_float64_closed_form(oracle, dim, bits):
    """Every tier computes sum(q_hat * centre(code)) / R; a float64 evaluation
    of that closed form bounds each tier's rounding."""
    gen = S.MixtureNP(dim, seed=1)
    docs = gen.docs(200)
    q = gen.queries(docs, 1)[0][0]
    codes = oracle.rq_encode(docs, bits)
    qu = oracle.l2_normalise(q)
    got = oracle.rq_scores(codes, bits, qu)
    cent = np.array([oracle.bucket_centre(b, bits) for b in range(1 << bits)], np.float64)
    want = np.array([(qu.astype(np.float64) * cent[oracle.unpack_buckets(r, dim, bits)]).sum() for r in codes])
    want /= float(oracle.rankquant_norm(dim, bits))
    assert np.max(np.abs(got - want)) < 2e-6


def test_tiers_agree_within_contract(oracle):
    """Scalar, AVX2 and AVX-512 orders differ only in rounding (1e-4 abs,
    docs/determinism.md:37-47)."""
    gen = S.MixtureNP(1024, seed=4)
    docs = gen.docs(300)
    qu = oracle.l2_normalise(gen.queries(docs, 1)[0][0])
    for bits in (2, 4):
        codes = oracle.rq_encode(docs, bits)
        t = [oracle.rq_scores(codes, bits, qu, tier=tr) for tr in (0, 1, 2)]
        assert np.max(np.abs(t[0] - t[1])) < 1e-5 and np.max(np.abs(t[0] - t[2])) < 1e-5
        assert not np.array_equal(t[0], t[2])  # the orders really differ


def test_tier_dispatch(oracle):
    """src/quant.rs:300-329 on an AVX-512 host."""
    assert oracle.select_tier(1024, 2) == oracle.TIER_AVX512
    assert oracle.select_tier(768, 4) == oracle.TIER_AVX512
    assert oracle.select_tier(48, 2) == oracle.TIER_AVX2
    assert oracle.select_tier(80, 4) == oracle.TIER_AVX2
    assert oracle.select_tier(20, 2) == oracle.TIER_SCALAR
    assert oracle.select_tier(1024, 1) == oracle.TIER_SCALAR


def test_l2_normalise(oracle):
    v = np.array([3.0, 4.0], np.float32)
    assert oracle.l2_normalise(v).tolist() == [np.float32(3.0) * (np.float32(1) / np.float32(5)),
                                               np.float32(4.0) * (np.float32(1) / np.float32(5))]
    assert np.all(oracle.l2_normalise(np.full(8, 1e-20, np.float32)) == 0.0)


# --- TopK and tie contracts (src/util.rs:386-600, tests/determinism_contract.rs)

def test_topk_matches_full_sort(oracle):
    rng = np.random.default_rng(1)
    s = rng.integers(0, 20, size=500).astype(np.float32) / 4
    for k in (1, 7, 100, 500):
        os_, oi = oracle.topk(s, k)
        order = np.lexsort((np.arange(500), -s))[:k]
        assert np.array_equal(oi, order) and np.array_equal(os_, s[order])


def test_topk_signed_zero_total_order(oracle):
    s = np.array([0.0, -0.0, 0.0, -0.0], np.float32)
    _, oi = oracle.topk(s, 4)
    assert oi.tolist() == [0, 2, 1, 3]  # +0 before -0 under total_cmp


def test_subset_tie_kats(oracle):
    docs = np.ones((12, 64), np.float32)
    codes = oracle.rq_encode(docs, 2)
    z = np.zeros(64, np.float32)
    s, i = oracle.search_subset(codes, 64, 2, z, [9, 3, 7, 1], 2)
    assert i.tolist() == [1, 3] and s.tolist() == [0.0, 0.0]
    _, i = oracle.search_subset(codes, 64, 2, z, [7, 8, 7], 2)
    assert i.tolist() == [7, 7]
    s, i = oracle.subset_batched(codes, 64, 2, z[None], [0, 4], [7, 7, 3, 7], 4)
    assert i[0].tolist() == [3, 7, 7, 7]
    s, i = oracle.search_asymmetric(codes, 64, 2, z[None], 4)
    assert i[0].tolist() == [0, 1, 2, 3] and np.all(s == 0.0)


def test_subset_full_equals_full_search(oracle):
    gen = S.MixtureNP(256, seed=5)
    docs = gen.docs(400)
    qs = gen.queries(docs, 3)[0]
    codes = oracle.rq_encode(docs, 2)
    fs, fi = oracle.search_asymmetric(codes, 256, 2, qs, 10)
    perm = np.random.default_rng(0).permutation(400)
    for q in range(3):
        s, i = oracle.search_subset(codes, 256, 2, qs[q], perm, 10)
        assert np.array_equal(i, fi[q]) and np.array_equal(s, fs[q])


def test_subset_batched_empty_rows_padded(oracle):
    docs = S.iid_unit(50, 64, 0)
    codes = oracle.rq_encode(docs, 2)
    s, i = oracle.subset_batched(codes, 64, 2, docs[:2], [0, 0, 3], [1, 2, 3], 5)
    assert i[0].tolist() == [-1] * 5 and np.all(np.isneginf(s[0]))
    assert i[1, 3:].tolist() == [-1, -1] and sorted(i[1, :3].tolist()) == [1, 2, 3]


# --- stage 1 (src/sign_bitmap.rs) ---------------------------------------------

def test_hamming_matches_numpy(oracle):
    rng = np.random.default_rng(2)
    bm = rng.integers(0, 2**63, size=(100, 4), dtype=np.uint64)
    q = rng.integers(0, 2**63, size=4, dtype=np.uint64)
    want = np.array([sum(bin(int(a) ^ int(b)).count("1") for a, b in zip(r, q)) for r in bm])
    assert np.array_equal(oracle.hamming_all(bm, q), want)


@pytest.mark.parametrize("threads", [1, 3])
def test_streamed_top_m_equals_bruteforce(oracle, threads):
    """tests/tiled_candgen.rs:41-175: random, tie-heavy and duplicate-run corpora."""
    rng = np.random.default_rng(3)
    for dim, n, nq, m in ((256, 3000, 5, 100), (768, 2500, 33, 2500), (64, 9000, 4, 700)):
        x = rng.standard_normal((n, dim)).astype(np.float32)
        if dim == 64:
            x = x[rng.integers(0, 5, size=n)]  # duplicate runs
        bm = oracle.sign_encode(x)
        qb = oracle.sign_encode(rng.standard_normal((nq, dim)).astype(np.float32))
        assert np.array_equal(oracle.sign_top_m(bm, qb, m, threads), oracle.sign_top_m_bruteforce(bm, qb, m))


def test_identical_docs_give_lowest_ids(oracle):
    """tests/determinism_contract.rs:127-150."""
    bm = oracle.sign_encode(np.ones((10, 64), np.float32))
    qb = oracle.sign_encode(np.ones((2, 64), np.float32))
    assert oracle.sign_top_m(bm, qb, 5).tolist() == [[0, 1, 2, 3, 4]] * 2


# --- FastScan (src/fastscan.rs) -----------------------------------------------

def test_fastscan_layout_and_lut(oracle):
    buckets = np.array([[0, 1, 2, 3, 3, 2, 1, 0]], np.uint8)
    packed = oracle.fs_pack(buckets, 8)
    assert packed.shape == (4 * 32,)
    assert [packed[p * 32] for p in range(4)] == [0b0001, 0b1011, 0b1110, 0b0100]
    assert np.all(packed.reshape(4, 32)[:, 1:] == 0)  # tail lanes zero
    q = oracle.l2_normalise(np.array([1, -2, 3, 0.5], np.float32))
    lut, bias, invq = oracle.fs_query(q)
    cen = np.array([-1.5, -0.5, 0.5, 1.5], np.float32)
    vals = np.array([[q[2 * p] * cen[a] + q[2 * p + 1] * cen[c] for a in range(4) for c in range(4)]
                     for p in range(2)], np.float32)
    gmin, gmax = vals.min(), vals.max()
    assert bias == np.float32(2) * gmin
    # each dequantised entry within half a quantisation step of the float LUT
    step = (gmax - gmin) / 255
    assert np.max(np.abs(gmin + lut.reshape(2, 16) * invq - vals)) <= 0.5 * step * (1 + 1e-5) + 1e-7


@pytest.mark.parametrize("dim", [64, 256, 1024])
def test_fastscan_scores_within_quantisation_bound(oracle, dim):
    """score = (bias_sum + acc * inv_q) / R (src/fastscan.rs:316-347) against
    the float64 closed form sum_j q_hat[j] * centre(code_j) / R: each of the
    d/2 pair entries is off by at most half a quantisation step, so every score
    lies within pairs * 0.5 * inv_q / R of it.  A dropped 1/R or bias_sum
    lands far outside that bound (checked on the same scores)."""
    gen = S.MixtureNP(dim, seed=61)
    n = 700
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 5)
    codes = oracle.rq_encode(docs, 2)
    buckets = np.stack([oracle.unpack_buckets(r, dim, 2) for r in codes])
    packed = oracle.fs_pack(buckets, dim)
    R = float(np.sqrt(dim * 15 / 12.0))
    assert np.float32(R) == oracle.rankquant_norm(dim, 2)
    cen = oracle.decode_centres(codes, dim, 2)
    pairs = dim // 2
    for q in qs:
        qu = oracle.l2_normalise(q)
        _, bias, invq = oracle.fs_query(qu)
        s, i = oracle.fs_search(packed, n, dim, q[None], n)
        assert sorted(i[0].tolist()) == list(range(n))
        exact = cen[i[0]] @ qu.astype(np.float64) / R
        bound = pairs * 0.5 * float(invq) / R * (1 + 1e-5) + 1e-6
        err = np.abs(s[0].astype(np.float64) - exact)
        assert np.max(err) <= bound
        # sensitivity: the plausible slips fail the same bound
        assert np.max(np.abs(s[0] * np.float64(R) - exact)) > bound
        assert np.max(np.abs(s[0] - float(bias) / R - exact)) > bound
        # order: score desc, id asc on exact ties
        assert np.all(np.diff(s[0]) <= 0)


def test_fastscan_top10_close_to_exact(oracle):
    """tests/index/fastscan.rs:24-70: FastScan top-10 overlaps exact b=2."""
    gen = S.MixtureNP(256, seed=6)
    docs = gen.docs(2000)
    qs = gen.queries(docs, 8)[0]
    codes = oracle.rq_encode(docs, 2)
    buckets = np.stack([oracle.unpack_buckets(r, 256, 2) for r in codes])
    _, fi = oracle.fs_search(oracle.fs_pack(buckets, 256), 2000, 256, qs, 10)
    _, ei = oracle.search_asymmetric(codes, 256, 2, qs, 10)
    overlap = np.mean([len(set(a) & set(b)) for a, b in zip(fi, ei)])
    assert overlap >= 9


# ---- score gradient (differentiable rerank checker) ----------------------------

def _loss_f64(centres, x, G, R):
    """L(x) = sum_i G_i * (centres_i . x/||x||) / R in float64."""
    qh = x / np.linalg.norm(x)
    return float(G @ (centres @ qh) / R)


@pytest.mark.parametrize("dim,bits", [(64, 2), (48, 4), (32, 1)])
def test_score_ids_closed_form_and_grad_vs_finite_differences(oracle, dim, bits):
    rng = np.random.default_rng(dim + bits)
    docs = S.iid_unit(300, dim, seed=dim)
    codes = oracle.rq_encode(docs, bits)
    x = rng.standard_normal((3, dim)).astype(np.float32) * 2.5
    ids = rng.integers(0, 300, size=(3, 12)).astype(np.int64)
    ids[1, 4] = -1
    G = rng.standard_normal((3, 12))
    R = float(np.sqrt(dim * (4 ** bits - 1) / 12.0))
    # score_ids equals the f64 closed form (within f32 rounding)
    s = oracle.score_ids(codes, bits, x, ids)
    for b in range(3):
        ok = ids[b] >= 0
        want = oracle.decode_centres(codes[ids[b, ok]], dim, bits) @ (x[b] / np.linalg.norm(x[b])) / R
        np.testing.assert_allclose(s[b, ok], want, rtol=1e-5, atol=1e-6)
        assert np.all(np.isneginf(s[b, ~ok]))
    grad = oracle.score_ids_grad(codes, bits, x, ids, G)
    h = 1e-3
    for b in range(3):
        ok = ids[b] >= 0
        cen = oracle.decode_centres(codes[ids[b, ok]], dim, bits)
        xb = x[b].astype(np.float64)
        fd = np.empty(dim)
        for j in range(dim):
            e = np.zeros(dim)
            e[j] = h
            fd[j] = (_loss_f64(cen, xb + e, G[b, ok], R) - _loss_f64(cen, xb - e, G[b, ok], R)) / (2 * h)
        np.testing.assert_allclose(grad[b], fd, rtol=1e-5, atol=1e-8)
        # the projection term matters: leaving it out is caught by the same check
        g = (G[b, ok] @ cen) / R
        assert np.max(np.abs(g / np.linalg.norm(xb) - fd)) > 1e-4


def test_score_grad_zero_query_is_zero(oracle):
    docs = S.iid_unit(10, 16, seed=1)
    codes = oracle.rq_encode(docs, 2)
    x = np.zeros((1, 16), np.float32)
    ids = np.arange(5, dtype=np.int64)[None]
    assert np.all(oracle.score_ids_grad(codes, 2, x, ids, np.ones((1, 5))) == 0.0)
    assert np.all(oracle.score_ids(codes, 2, x, ids) == 0.0)


# ---- Bitmap (top-bucket, AND-popcount) ------------------------------------------

def _stable_top_bits(order, n_top, d):
    w = np.zeros(d // 64, dtype=np.uint64)
    for j in order:
        w[j // 64] |= np.uint64(1) << np.uint64(j % 64)
    return w


def test_bitmap_fixture_kats(oracle):
    # tests/persistence_compat.rs:160-166: Bitmap(64, 2) of 0..63 -> bits 62, 63
    w = oracle.bitmap_encode(np.arange(64, dtype=np.float32)[None], 2)
    assert int(w[0, 0]) == 0xC000000000000000
    # ordvec-ffi/src/lib.rs:1035-1049, 1227-1270: Bitmap(64, 4), docs with
    # 10..13 in coordinates 0..3; subset [3, 1, 1, 2], k = 3 -> [1, 1, 2]
    doc = np.zeros(64, np.float32)
    doc[:4] = 10 + np.arange(4)
    bm = oracle.bitmap_encode(np.repeat(doc[None], 4, axis=0), 4)
    assert np.all(bm[:, 0] == 0xF)
    qb = oracle.bitmap_query(doc, 4)
    assert int(qb[0, 0]) == 0xF
    s, i = oracle.bitmap_search_subset(bm, qb, [3, 1, 1, 2], 3)
    assert list(i) == [1, 1, 2] and list(s) == [4.0, 4.0, 4.0]
    s, i = oracle.bitmap_search(bm, qb, 2)
    assert list(i[0]) == [0, 1]


def test_bitmap_encode_and_query_bruteforce(oracle):
    rng = np.random.default_rng(41)
    for d, n_top in [(64, 16), (128, 1), (192, 100), (1024, 256)]:
        x = np.round(rng.standard_normal((20, d)) * 2).astype(np.float32)  # heavy ties
        x[0, :10] = -0.0
        x[1, :10] = 0.0
        bm = oracle.bitmap_encode(x, n_top)
        qb = oracle.bitmap_query(x, n_top)
        for r in range(20):
            canon = np.where(x[r] == 0, np.float32(0), x[r])
            asc = np.argsort(canon, kind="stable")  # rank order: ties by index
            assert np.array_equal(bm[r], _stable_top_bits(asc[d - n_top:], n_top, d)), (d, r)
            desc = np.argsort(-canon, kind="stable")  # value desc, ties by index
            assert np.array_equal(qb[r], _stable_top_bits(desc[:n_top], n_top, d)), (d, r)
            assert sum(bin(int(v)).count("1") for v in bm[r]) == n_top


def test_bitmap_search_and_top_m_orders(oracle):
    rng = np.random.default_rng(42)
    d, n_top = 128, 32
    docs = np.round(rng.standard_normal((500, d))).astype(np.float32)
    docs[100:140] = docs[7]  # duplicates tie
    bm = oracle.bitmap_encode(docs, n_top)
    qb = oracle.bitmap_query(docs[:6], n_top)
    ov = oracle.and_popcount_all(bm, qb)
    s, i = oracle.bitmap_search(bm, qb, 50)
    tm = oracle.bitmap_top_m(bm, qb, 50)
    for q in range(6):
        want = sorted(range(500), key=lambda j: (-int(ov[q, j]), j))[:50]
        assert list(i[q]) == want and list(tm[q]) == want
        assert np.array_equal(s[q], ov[q, want].astype(np.float32))
    # identical docs: ids 0..k-1 with equal scores (tests/determinism_contract.rs:40-49, 135-150)
    same = oracle.bitmap_encode(np.repeat(docs[:1], 10, axis=0), n_top)
    s, i = oracle.bitmap_search(same, oracle.bitmap_query(docs[3], n_top), 4)
    assert list(i[0]) == [0, 1, 2, 3] and len(set(s[0])) == 1
    assert list(oracle.bitmap_top_m(same, oracle.bitmap_query(docs[3], n_top), 5)[0]) == [0, 1, 2, 3, 4]


# ---- FastScan b=4 (extension of the b=2 recipe) -----------------------------------

def test_fs4_lut_structure(oracle):
    rng = np.random.default_rng(51)
    q = oracle.l2_normalise(rng.standard_normal(64).astype(np.float32))
    lut, bias, invq = oracle.fs4_query(q)
    t = lut.reshape(64, 16).astype(int)
    assert t.min() == 0 and t.max() == 255
    for j in range(64):
        d = np.diff(t[j])
        assert np.all(d >= 0) if q[j] > 0 else np.all(d <= 0)
    gmin = np.float32(np.min(q[:, None] * (np.arange(16, dtype=np.float32) - np.float32(7.5))))
    assert bias == np.float32(64) * gmin
    # zero query: every entry equal, every score 0, ids ascending
    lut0, b0, _ = oracle.fs4_query(np.zeros(64, np.float32))
    assert np.all(lut0 == 0) and b0 == 0
    codes = oracle.rq_encode(S.iid_unit(20, 64, seed=2), 4)
    s, i = oracle.fs4_search(codes, 64, np.zeros((1, 64), np.float32), 5)
    assert list(i[0]) == [0, 1, 2, 3, 4] and np.all(s == 0)


@pytest.mark.parametrize("dim", [64, 768])
def test_fs4_scores_within_quantisation_bound(oracle, dim):
    gen = S.MixtureNP(dim, seed=52)
    docs = gen.docs(400)
    qs, _ = gen.queries(docs, 6)
    codes = oracle.rq_encode(docs, 4)
    R = float(np.sqrt(dim * 255 / 12.0))
    cen = oracle.decode_centres(codes, dim, 4)
    for q in qs:
        qu = oracle.l2_normalise(q)
        _, _, invq = oracle.fs4_query(qu)
        s, i = oracle.fs4_search(codes, dim, q[None], 400)
        exact = cen[i[0]] @ qu.astype(np.float64) / R
        # each of the dim entries is off by at most half a quantisation step
        bound = (dim * 0.5 * float(invq)) / R * (1 + 1e-5) + 1e-6
        assert np.max(np.abs(s[0] - exact)) <= bound


def test_fs4_top10_overlaps_exact_b4(oracle):
    gen = S.MixtureNP(256, seed=53)
    docs = gen.docs(3000)
    qs, _ = gen.queries(docs, 20)
    codes = oracle.rq_encode(docs, 4)
    fs, fi = oracle.fs4_search(codes, 256, qs, 10)
    es, ei = oracle.search_asymmetric(codes, 256, 4, qs, 10)
    overlap = np.mean([len(set(a) & set(b)) for a, b in zip(fi, ei)])
    assert overlap >= 9.0, overlap


# --- Rank and symmetric RankQuant search (src/rank.rs:249-252, 364-480,
# src/quant.rs:629-695; tests/index/rank.rs) --------------------------------

def _uniform(n, d, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, size=(n, d)).astype(np.float32)


def _ranks64(x):
    """Ranks by stable argsort of argsort (ties by index), float64."""
    return np.argsort(np.argsort(x, axis=1, kind="stable"), axis=1, kind="stable").astype(np.float64)


def test_rank_norm_is_the_centred_permutation_norm(oracle):
    for d in (2, 3, 64, 100, 768, 1024, 4096, 65535):
        centred = np.arange(d, dtype=np.float64) - (d - 1) / 2.0
        assert oracle.rank_norm(d) == np.float32(np.sqrt(np.sum(centred * centred))), d


def test_rank_search_is_spearman(oracle):
    """tests/index/rank.rs:10-48: top-10 of brute-force rank cosine, scores
    within 1e-4; here the rank cosine is scipy's Spearman correlation."""
    from scipy.stats import spearmanr

    d, n = 128, 256
    docs = _uniform(n, d, 1)
    qs = _uniform(5, d, 99)
    ranks = oracle.rank_encode(docs)
    s, i = oracle.rank_search(ranks, qs, 10)
    for qi in range(len(qs)):
        rho = np.array([spearmanr(qs[qi], docs[j]).statistic for j in range(n)])
        order = np.lexsort((np.arange(n), -rho))[:10]
        assert np.array_equal(i[qi], order)
        assert np.allclose(s[qi], rho[order], rtol=0, atol=1e-5)


def test_rank_search_asymmetric_closed_form(oracle):
    """tests/index/rank.rs:51-86 (ref_asymmetric): <q/|q|, r - mean> / norm."""
    d, n = 128, 256
    docs = _uniform(n, d, 2)
    qs = _uniform(5, d, 100)
    ranks = oracle.rank_encode(docs)
    s, i = oracle.rank_search_asymmetric(ranks, qs, 10)
    centred = _ranks64(docs) - (d - 1) / 2.0
    norm = np.sqrt(d * (d * d - 1) / 12.0)
    for qi in range(len(qs)):
        q = qs[qi].astype(np.float64)
        sc = centred @ (q / np.linalg.norm(q)) / norm
        order = np.lexsort((np.arange(n), -sc))[:10]
        assert np.array_equal(i[qi], order)
        assert np.allclose(s[qi], sc[order], rtol=0, atol=1e-5)


def test_rank_search_recall_against_fp32_cosine(oracle):
    """tests/index/rank.rs:89-138: top-10 overlap with fp32 cosine >= 0.7."""
    d, n = 128, 256
    docs = _uniform(n, d, 7)
    qs = _uniform(20, d, 8)
    _, i = oracle.rank_search(oracle.rank_encode(docs), qs, 10)
    dn = docs / np.linalg.norm(docs, axis=1, keepdims=True)
    cos = (qs / np.linalg.norm(qs, axis=1, keepdims=True)) @ dn.T
    overlap = sum(len(set(i[r]) & set(np.argsort(-cos[r])[:10])) for r in range(20))
    assert overlap / 200 >= 0.7


def test_rank_search_self_reverse_ties_and_zero_query(oracle):
    d = 96
    docs = _uniform(40, d, 5)
    docs[7] = docs[3]                      # identical documents: lower id first
    ranks = oracle.rank_encode(docs)
    s, i = oracle.rank_search(ranks, docs[3:4], 40)
    assert i[0, 0] == 3 and i[0, 1] == 7 and s[0, 0] == s[0, 1]
    assert abs(float(s[0, 0]) - 1.0) < 1e-6
    s, i = oracle.rank_search(ranks, -docs[11:12], 40)
    assert i[0, -1] == 11 and abs(float(s[0, -1]) + 1.0) < 1e-6
    # a zero query normalises to zeros: every asymmetric score is +0.0
    s, i = oracle.rank_search_asymmetric(ranks, np.zeros((1, d), np.float32), 5)
    assert i[0].tolist() == [0, 1, 2, 3, 4] and np.all(s.view(np.uint32) == 0)
    s, i = oracle.rank_search(ranks, docs[:2], 100)  # k clamps to n
    assert s.shape == (2, 40)


@pytest.mark.parametrize("bits", [1, 2, 4])
def test_rq_search_symmetric_closed_form(oracle, bits):
    """Bucket-centre correlation: <centre(bq), centre(bd)> / R^2, R^2 =
    d (4^b - 1) / 12 (constant composition makes every row's norm R)."""
    d, n = 128, 300
    docs = _uniform(n, d, 10 + bits)
    qs = _uniform(4, d, 20 + bits)
    nb = 1 << bits
    codes = oracle.rq_encode(docs, bits)
    s, i = oracle.rq_search_symmetric(codes, d, bits, qs, 10)
    cd = np.floor(_ranks64(docs) * nb / d) - (nb - 1) / 2.0
    cq = np.floor(_ranks64(qs) * nb / d) - (nb - 1) / 2.0
    r2 = d * (nb * nb - 1) / 12.0
    assert np.allclose(np.sum(cd * cd, axis=1), r2)
    sc = cq @ cd.T / r2
    for qi in range(len(qs)):
        order = np.lexsort((np.arange(n), -sc[qi]))[:10]
        assert np.allclose(s[qi], sc[qi][order], rtol=0, atol=1e-5)
        # ties are common at b=1: the id sets may differ only within equal scores
        assert np.allclose(sc[qi][i[qi]], sc[qi][order], rtol=0, atol=1e-9)
        for j in range(9):  # equal scores: ascending ids
            assert s[qi, j] > s[qi, j + 1] or (s[qi, j] == s[qi, j + 1] and i[qi, j] < i[qi, j + 1])
    s, i = oracle.rq_search_symmetric(codes, d, bits, docs[5:6], 1)
    assert i[0, 0] == 5 and abs(float(s[0, 0]) - 1.0) < 1e-6


# ---- AVX-512 forms (the CPU baseline's kernels) -------------------------------

@pytest.mark.skipif("not __import__('oracle').simd_supported()", reason="CPU lacks AVX-512 F/DQ/VPOPCNTDQ")
@pytest.mark.parametrize("dim", [64, 128, 576, 1024])
def test_simd_forms_equal_scalar_restatements(oracle, dim):
    """The AVX-512 VPOPCNTDQ scan (src/sign_bitmap.rs:845-955) and the AVX-512
    scoring tier (src/quant_kernels.rs:406-581) give the scalar restatements'
    results bit for bit, including masked tail words (qpv % 8 != 0)."""
    gen = S.MixtureNP(dim, seed=71)
    docs = gen.docs(3000)
    qs, _ = gen.queries(docs, 13)
    bm, qb = oracle.sign_encode(docs), oracle.sign_encode(qs)
    try:
        assert oracle.simd(True)
        fast = oracle.sign_top_m(bm, qb, 300, 2)
        assert oracle.simd(False) is False
        slow = oracle.sign_top_m(bm, qb, 300, 2)
        assert np.array_equal(fast, slow)
        assert np.array_equal(fast, oracle.sign_top_m_bruteforce(bm, qb, 300))
        for bits in (2, 4):
            codes = oracle.rq_encode(docs, bits)
            oracle.simd(False)
            ss, si = oracle.search_asymmetric(codes, dim, bits, qs, 50)
            oracle.simd(True)
            fs, fi = oracle.search_asymmetric(codes, dim, bits, qs, 50)
            assert np.array_equal(si, fi)
            assert np.array_equal(ss.view(np.uint32), fs.view(np.uint32))
    finally:
        oracle.simd(True)


def _eval_bruteforce(corpus, queries, bits, k):
    """Eval search restated with numpy: stable-argsort ranks, bucket
    floor(r * 2^bits / d), centres b - (2^bits - 1)/2, f64 dot products over
    the f64 empirical norm, top-k by (score desc, id asc)."""
    d = corpus.shape[1]

    def centres(x):
        r = np.empty(x.shape, np.int64)
        order = np.argsort(np.where(x == 0, 0, x), axis=1, kind="stable")
        np.put_along_axis(r, order, np.arange(d)[None, :].repeat(len(x), 0), axis=1)
        return (r << bits) // d - ((1 << bits) - 1) / 2.0

    cc, qc = centres(corpus), centres(queries)
    norm2 = float(np.sum((((np.arange(d) << bits) // d) - ((1 << bits) - 1) / 2.0) ** 2))
    s = qc @ cc.T / norm2
    ids = np.stack([np.lexsort((np.arange(len(corpus)), -row))[:k] for row in s])
    return np.take_along_axis(s, ids, 1), ids


@pytest.mark.parametrize("bits", [1, 2, 3, 4, 5, 7, 8])
def test_eval_search_matches_bruteforce_and_symmetric(oracle, bits):
    rng = np.random.default_rng(60 + bits)
    d = 128
    corpus = rng.standard_normal((45, d)).astype(np.float32)
    corpus[3] = corpus[7]                      # duplicate documents tie on score
    queries = rng.standard_normal((5, d)).astype(np.float32)
    s, i = oracle.eval_search(corpus, queries, bits, 9)
    ws, wi = _eval_bruteforce(corpus, queries, bits, 9)
    assert np.array_equal(i, wi)
    np.testing.assert_allclose(s, ws, rtol=0, atol=1e-6)
    if bits in (1, 2, 4):  # tests/index/quant.rs:91-117: equal to the packed symmetric search
        ps, pi = oracle.rq_search_symmetric(oracle.rq_encode(corpus, bits), d, bits, queries, 9)
        assert np.array_equal(i, pi)
        np.testing.assert_allclose(s, ps, rtol=0, atol=1e-6)
    # a plausible slip (closed-form norm where occupancy is uneven) is visible at d = 100
    if bits == 3:
        c100 = rng.standard_normal((20, 100)).astype(np.float32)
        s100, _ = oracle.eval_search(c100, c100[:2], 3, 1)
        assert abs(float(s100[0, 0]) - 1.0) < 1e-6  # a document scores exactly 1 against itself
    assert oracle.eval_search(corpus, queries[:0], bits, 9)[0].shape == (0, 9)


def test_eval_search_b8_any_dim(oracle):
    # tests/index/quant_b8.rs:527-556: b = 8 at dim 384 returns valid ids
    dim, n, nq = 384, 32, 2
    corpus = ((np.arange(n * dim) * 7 % 101).astype(np.float32) - 50.0).reshape(n, dim)
    queries = ((np.arange(nq * dim) * 13 % 97).astype(np.float32) - 48.0).reshape(nq, dim)
    s, i = oracle.eval_search(corpus, queries, 8, 5)
    assert i.shape == (2, 5) and ((i >= 0) & (i < n)).all()
    ws, wi = _eval_bruteforce(corpus, queries, 8, 5)
    np.testing.assert_allclose(s, ws, rtol=0, atol=1e-6)
