"""CPU oracle for the ordinal retrieval hot path -- TEST INFRASTRUCTURE ONLY.

A numpy front end over ``oracle/liboracle.so`` (built from
``oracle/ordvec_oracle.c``), the plain-C restatement of the reference's CPU
path.  It is the checker the HIP product path is compared against and the
CPU baseline timed beside it.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg may import it; the product package never
does.

Each wrapper names the reference function it restates; the C file carries
the file:line citations.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

TIER_SCALAR, TIER_AVX2, TIER_AVX512 = 0, 1, 2


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
U64 = C.c_uint64
U32 = C.c_uint32


def _declare(L):
    sig = {
        "orc_simd_supported": (C.c_int, []),
        "orc_eval_search": (None, [P, U64, P, U64, U32, U32, U64, P, P]),
        "orc_set_simd": (C.c_int, [C.c_int]),
        "orc_rank_transform": (None, [P, U32, P]),
        "orc_rank_transform_qsort": (None, [P, U32, P]),
        "orc_rank_to_bucket": (C.c_uint8, [U32, U32, U32]),
        "orc_pack_buckets": (None, [P, U32, U32, P]),
        "orc_unpack_buckets": (None, [P, U32, U32, P]),
        "orc_bucket_centre": (C.c_float, [U32, U32]),
        "orc_rankquant_norm": (C.c_float, [U32, U32]),
        "orc_rq_encode": (None, [P, U64, U32, U32, P]),
        "orc_sign_encode": (None, [P, U64, U32, P]),
        "orc_l2_normalise": (None, [P, U32, P]),
        "orc_select_tier": (C.c_int, [U32, U32]),
        "orc_rq_scores": (None, [P, P, U64, U32, U32, P, C.c_int, P]),
        "orc_score_ids_grad": (None, [P, U32, U32, P, U64, P, U64, P, P]),
        "orc_bitmap_encode": (None, [P, U64, U32, U32, P]),
        "orc_fs4_query": (None, [P, U32, P, P, P]),
        "orc_rq_search_symmetric": (None, [P, U64, U32, U32, P, U64, U64, P, P]),
        "orc_rank_norm": (C.c_float, [U32]),
        "orc_rank_search": (None, [P, U64, U32, P, U64, U64, P, P]),
        "orc_rank_search_asymmetric": (None, [P, U64, U32, P, U64, U64, P, P]),
        "orc_fs4_search": (None, [P, U64, U32, P, U64, U64, P, P]),
        "orc_bitmap_query": (None, [P, U64, U32, U32, P]),
        "orc_and_popcount_all": (None, [P, U64, U32, P, U64, P]),
        "orc_topk": (None, [P, P, U64, U64, P, P]),
        "orc_search_asymmetric": (None, [P, U64, U32, U32, P, U64, U64, P, P]),
        "orc_subset_batched": (None, [P, U64, U32, U32, P, U64, P, P, U64, P, P, C.c_int]),
        "orc_search_subset": (None, [P, U32, U32, P, P, U64, U64, P, P]),
        "orc_hamming_all": (None, [P, U64, U32, P, P]),
        "orc_sign_top_m": (None, [P, U64, U32, P, U64, U64, P, C.c_int]),
        "orc_sign_top_m_bruteforce": (None, [P, U64, U32, P, U64, U64, P]),
        "orc_fs_pack": (None, [P, U64, U32, P]),
        "orc_fs_query": (None, [P, U32, P, P, P]),
        "orc_fs_search": (None, [P, U64, U32, P, U64, U64, P, P]),
        "orc_num_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


# --- src/rank.rs -------------------------------------------------------------

def eval_search(corpus, queries, bits: int, k: int):
    """rankquant_eval_search (src/quant.rs:1541-1610)."""
    c = _f32(corpus)
    q = _f32(queries).reshape(-1, c.shape[1])
    n, d = c.shape
    nq = q.shape[0]
    kk = min(k, n)
    os_ = np.zeros((nq, kk), np.float32)
    oi = np.full((nq, kk), -1, np.int64)
    if kk and nq:
        lib().orc_eval_search(_p(c), n, _p(q), nq, d, bits, kk, _p(os_), _p(oi))
    return os_, oi


def simd(enable=None) -> bool:
    """AVX-512 forms of the stage-1 scan and the AVX-512 scoring tier
    (identical results): None queries, False forces the scalar restatements,
    True uses AVX-512 where the CPU has it.  Returns whether they are in use."""
    return bool(lib().orc_set_simd(-1 if enable is None else (1 if enable else 0)))


def simd_supported() -> bool:
    return bool(lib().orc_simd_supported())


def rank_transform(v) -> np.ndarray:
    """rank_transform (src/rank.rs:38-55)."""
    v = _f32(v)
    out = np.empty(v.shape[0], dtype=np.uint16)
    lib().orc_rank_transform(_p(v), v.shape[0], _p(out))
    return out


def rank_transform_qsort(v) -> np.ndarray:
    v = _f32(v)
    out = np.empty(v.shape[0], dtype=np.uint16)
    lib().orc_rank_transform_qsort(_p(v), v.shape[0], _p(out))
    return out


def rank_to_bucket(rank: int, d: int, bits: int) -> int:
    return int(lib().orc_rank_to_bucket(rank, d, bits))


def pack_buckets(buckets, bits: int) -> np.ndarray:
    b = np.ascontiguousarray(buckets, dtype=np.uint8)
    out = np.empty(b.shape[0] * bits // 8, dtype=np.uint8)
    lib().orc_pack_buckets(_p(b), b.shape[0], bits, _p(out))
    return out


def unpack_buckets(packed, d: int, bits: int) -> np.ndarray:
    p = np.ascontiguousarray(packed, dtype=np.uint8)
    out = np.empty(d, dtype=np.uint8)
    lib().orc_unpack_buckets(_p(p), d, bits, _p(out))
    return out


def bucket_centre(b: int, bits: int) -> float:
    return float(lib().orc_bucket_centre(b, bits))


def rankquant_norm(d: int, bits: int) -> np.float32:
    return np.float32(lib().orc_rankquant_norm(d, bits))


def rq_encode(x, bits: int) -> np.ndarray:
    """RankQuant::add (src/quant.rs:587-614): packed codes, n x d*bits/8."""
    x = _f32(x)
    n, d = x.shape
    out = np.empty((n, d * bits // 8), dtype=np.uint8)
    lib().orc_rq_encode(_p(x), n, d, bits, _p(out))
    return out


def sign_encode(x) -> np.ndarray:
    """SignBitmap::add (src/sign_bitmap.rs:180-200): n x d/64 u64 words."""
    x = _f32(x)
    n, d = x.shape
    out = np.empty((n, d // 64), dtype=np.uint64)
    lib().orc_sign_encode(_p(x), n, d, _p(out))
    return out


def l2_normalise(v) -> np.ndarray:
    v = _f32(v)
    out = np.empty_like(v)
    lib().orc_l2_normalise(_p(v), v.shape[0], _p(out))
    return out


def select_tier(d: int, bits: int) -> int:
    return int(lib().orc_select_tier(d, bits))


def rq_scores(codes, bits: int, q_unit, ids=None, tier: int = -1) -> np.ndarray:
    """Per-document score exactly as the reference's TopK sees it (kernel
    output plus score offset), for an already-normalised query."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    q_unit = _f32(q_unit)
    d = q_unit.shape[0]
    if ids is None:
        n = codes.shape[0]
        idp = None
    else:
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        n = ids.shape[0]
        idp = _p(ids)
    out = np.empty(n, dtype=np.float32)
    lib().orc_rq_scores(_p(codes), idp, n, d, bits, _p(q_unit), tier, _p(out))
    return out


def score_ids(codes, bits: int, queries, ids) -> np.ndarray:
    """Scores of rows ``ids`` [nq, k] (-1 = none -> -inf) for raw queries,
    normalised first (src/util.rs:84-115), as the reference's TopK sees them."""
    q = _f32(np.atleast_2d(queries))
    ids = np.asarray(ids, dtype=np.int64).reshape(q.shape[0], -1)
    out = np.full(ids.shape, -np.inf, dtype=np.float32)
    for b in range(q.shape[0]):
        ok = ids[b] >= 0
        out[b, ok] = rq_scores(codes, bits, l2_normalise(q[b]), ids[b, ok])
    return out


def score_ids_grad(codes, bits: int, queries, ids, grad_scores) -> np.ndarray:
    """f64 dL/dqueries of score_ids given dL/dscores (analytic; see the C
    function's comment)."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    q = _f32(np.atleast_2d(queries))
    nq, d = q.shape
    ids = np.ascontiguousarray(np.asarray(ids, dtype=np.int64).reshape(nq, -1))
    g = np.ascontiguousarray(np.asarray(grad_scores, dtype=np.float64).reshape(ids.shape))
    out = np.empty((nq, d), dtype=np.float64)
    lib().orc_score_ids_grad(_p(codes), d, bits, _p(q), nq, _p(ids), ids.shape[1], _p(g), _p(out))
    return out


def bitmap_encode(x, n_top: int) -> np.ndarray:
    """Bitmap::add (src/bitmap.rs:140-162): n x d/64 u64, row-major."""
    x = _f32(np.atleast_2d(x))
    n, d = x.shape
    out = np.empty((n, d // 64), dtype=np.uint64)
    lib().orc_bitmap_encode(_p(x), n, d, n_top, _p(out))
    return out


def bitmap_query(q, n_top: int) -> np.ndarray:
    """build_query_bitmap_fp32 (src/bitmap.rs:164-187): nq x d/64 u64."""
    q = _f32(np.atleast_2d(q))
    nq, d = q.shape
    out = np.empty((nq, d // 64), dtype=np.uint64)
    lib().orc_bitmap_query(_p(q), nq, d, n_top, _p(out))
    return out


def and_popcount_all(bm, qb) -> np.ndarray:
    """overlap[q, i] = popcount(doc_i AND query_q) (src/util.rs and_popcount)."""
    bm = np.ascontiguousarray(bm, dtype=np.uint64)
    qb = np.ascontiguousarray(np.atleast_2d(qb), dtype=np.uint64)
    out = np.empty((qb.shape[0], bm.shape[0]), dtype=np.uint32)
    lib().orc_and_popcount_all(_p(bm), bm.shape[0], bm.shape[1], _p(qb), qb.shape[0], _p(out))
    return out


def bitmap_search(bm, qb, k: int):
    """Bitmap::search (src/bitmap.rs:202-250): TopK over overlaps as f32 with
    the document id as tie key."""
    ov = and_popcount_all(bm, qb)
    kk = min(k, bm.shape[0])
    s = np.empty((qb.shape[0], kk), np.float32)
    i = np.empty((qb.shape[0], kk), np.int64)
    for q in range(qb.shape[0]):
        s[q], i[q] = topk(ov[q].astype(np.float32), kk)
    return s, i


def bitmap_top_m(bm, qb, m: int) -> np.ndarray:
    """top_m_candidates(_batched) (src/bitmap.rs:264-362): ids sorted by
    (overlap desc, id asc)."""
    ov = and_popcount_all(bm, qb)
    n = bm.shape[0]
    m_eff = min(m, n)
    ids = np.arange(n)
    return np.stack([np.lexsort((ids, -ov[q].astype(np.int64)))[:m_eff] for q in range(ov.shape[0])]).astype(np.uint32)


def bitmap_search_subset(bm, qb_row, rows, k: int):
    """Bitmap::search_subset (src/bitmap.rs:477-505): score every listed row,
    order (overlap desc, row asc), keep min(k, len)."""
    rows = np.asarray(rows, dtype=np.uint32)
    ov = and_popcount_all(bm[rows] if len(rows) else bm[:0], qb_row)[0].astype(np.int64)
    order = np.lexsort((rows.astype(np.int64), -ov))[:min(k, len(rows))]
    return ov[order].astype(np.float32), rows[order].astype(np.int64)


def decode_centres(codes, d: int, bits: int) -> np.ndarray:
    """Rows of packed codes -> centre(code) as f64 [n, d] (MSB-first packing,
    src/rank.rs:89-156; centre, src/rank.rs:220-228)."""
    c = np.asarray(codes, dtype=np.uint8)
    cpb = 8 // bits
    shifts = np.array([(cpb - 1 - j) * bits for j in range(cpb)], dtype=np.uint8)
    v = (c[..., :, None] >> shifts) & ((1 << bits) - 1)
    return v.reshape(*c.shape[:-1], d).astype(np.float64) - ((1 << bits) - 1) / 2.0


def topk(scores, k: int, tie_keys=None):
    """TopK (src/util.rs:386-600) over a dense score vector."""
    s = _f32(scores)
    tk = None if tie_keys is None else np.ascontiguousarray(tie_keys, dtype=np.int64)
    os_ = np.empty(k, dtype=np.float32)
    oi = np.empty(k, dtype=np.int64)
    lib().orc_topk(_p(s), None if tk is None else _p(tk), s.shape[0], k, _p(os_), _p(oi))
    return os_, oi


def search_asymmetric(codes, d: int, bits: int, queries, k: int):
    """RankQuant::search_asymmetric (src/quant.rs:713-847)."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    q = _f32(queries).reshape(-1, d)
    n = codes.shape[0]
    nq = q.shape[0]
    kk = min(k, n)
    os_ = np.zeros((nq, kk), dtype=np.float32)
    oi = np.full((nq, kk), -1, dtype=np.int64)
    if kk:
        lib().orc_search_asymmetric(_p(codes), n, d, bits, _p(q), nq, kk, _p(os_), _p(oi))
    return os_, oi


def subset_batched(codes, d: int, bits: int, queries, offsets, cands, k: int, threads: int = 1):
    """search_asymmetric_subset_batched_serial (src/quant.rs:1336-1451)."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    q = _f32(queries).reshape(-1, d)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    cd = np.ascontiguousarray(cands, dtype=np.uint32)
    n = codes.shape[0]
    nq = q.shape[0]
    out_k = min(k, n)
    os_ = np.full((nq, out_k), -np.inf, dtype=np.float32)
    oi = np.full((nq, out_k), -1, dtype=np.int64)
    if out_k and nq:
        lib().orc_subset_batched(_p(codes), n, d, bits, _p(q), nq, _p(off),
                                 _p(cd) if cd.size else None, k, _p(os_), _p(oi), threads)
    return os_, oi


def search_subset(codes, d: int, bits: int, query, rows, k: int):
    """search_asymmetric_subset (src/quant.rs:1048-1083)."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    q = _f32(query)
    r = np.ascontiguousarray(rows, dtype=np.uint32)
    out_k = min(k, r.shape[0])
    os_ = np.empty(out_k, dtype=np.float32)
    oi = np.empty(out_k, dtype=np.int64)
    if out_k:
        lib().orc_search_subset(_p(codes), d, bits, _p(q), _p(r), r.shape[0], k, _p(os_), _p(oi))
    return os_, oi


def hamming_all(bitmaps, qbits) -> np.ndarray:
    bm = np.ascontiguousarray(bitmaps, dtype=np.uint64)
    qb = np.ascontiguousarray(qbits, dtype=np.uint64)
    n, qpv = bm.shape
    out = np.empty(n, dtype=np.uint32)
    lib().orc_hamming_all(_p(bm), n, qpv, _p(qb), _p(out))
    return out


def sign_top_m(bitmaps, qbits, m: int, threads: int = 1) -> np.ndarray:
    """top_m_candidates_streamed (src/sign_bitmap.rs:231-328): nq x m_eff."""
    bm = np.ascontiguousarray(bitmaps, dtype=np.uint64)
    qb = np.ascontiguousarray(qbits, dtype=np.uint64).reshape(-1, bm.shape[1])
    n, qpv = bm.shape
    nq = qb.shape[0]
    m_eff = min(m, n)
    out = np.empty((nq, m_eff), dtype=np.uint32)
    if m_eff and nq:
        lib().orc_sign_top_m(_p(bm), n, qpv, _p(qb), nq, m, _p(out), threads)
    return out


def sign_top_m_bruteforce(bitmaps, qbits, m: int) -> np.ndarray:
    bm = np.ascontiguousarray(bitmaps, dtype=np.uint64)
    qb = np.ascontiguousarray(qbits, dtype=np.uint64).reshape(-1, bm.shape[1])
    n, qpv = bm.shape
    nq = qb.shape[0]
    m_eff = min(m, n)
    out = np.empty((nq, m_eff), dtype=np.uint32)
    if m_eff and nq:
        lib().orc_sign_top_m_bruteforce(_p(bm), n, qpv, _p(qb), nq, m, _p(out))
    return out


def fs_pack(buckets, d: int) -> np.ndarray:
    b = np.ascontiguousarray(buckets, dtype=np.uint8).reshape(-1, d)
    n = b.shape[0]
    out = np.empty(((n + 31) // 32) * (d // 2) * 32, dtype=np.uint8)
    lib().orc_fs_pack(_p(b), n, d, _p(out))
    return out


def fs_query(q_unit):
    q = _f32(q_unit)
    d = q.shape[0]
    lut = np.empty((d // 2) * 16, dtype=np.uint8)
    bias = C.c_float()
    invq = C.c_float()
    lib().orc_fs_query(_p(q), d, _p(lut), C.byref(bias), C.byref(invq))
    return lut, np.float32(bias.value), np.float32(invq.value)


def fs_search(packed, n: int, d: int, queries, k: int):
    """RankQuantFastscan::search (src/fastscan.rs:357-476, scalar kernel)."""
    p = np.ascontiguousarray(packed, dtype=np.uint8)
    q = _f32(queries).reshape(-1, d)
    nq = q.shape[0]
    kk = min(k, n)
    os_ = np.zeros((nq, kk), dtype=np.float32)
    oi = np.full((nq, kk), -1, dtype=np.int64)
    if kk:
        lib().orc_fs_search(_p(p), n, d, _p(q), nq, kk, _p(os_), _p(oi))
    return os_, oi


def fs4_query(q_unit):
    """b=4 FastScan query LUT (d x 16 u8), bias_sum, inv_q."""
    q = _f32(q_unit)
    d = q.shape[0]
    lut = np.empty(d * 16, dtype=np.uint8)
    bias = C.c_float()
    invq = C.c_float()
    lib().orc_fs4_query(_p(q), d, _p(lut), C.byref(bias), C.byref(invq))
    return lut, np.float32(bias.value), np.float32(invq.value)


def fs4_search(codes, d: int, queries, k: int):
    """b=4 FastScan over RankQuant b=4 codes (n x d/2 bytes)."""
    c = np.ascontiguousarray(codes, dtype=np.uint8).reshape(-1, d // 2)
    n = c.shape[0]
    q = _f32(queries).reshape(-1, d)
    nq = q.shape[0]
    kk = min(k, n)
    os_ = np.zeros((nq, kk), dtype=np.float32)
    oi = np.full((nq, kk), -1, dtype=np.int64)
    if kk:
        lib().orc_fs4_search(_p(c), n, d, _p(q), nq, kk, _p(os_), _p(oi))
    return os_, oi


def _search_out(nq, k):
    return np.zeros((nq, k), dtype=np.float32), np.full((nq, k), -1, dtype=np.int64)


def rq_search_symmetric(codes, d: int, bits: int, queries, k: int):
    """RankQuant::search, symmetric (src/quant.rs:629-695)."""
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    q = _f32(np.atleast_2d(queries))
    n = c.shape[0]
    kk = min(k, n)
    os_, oi = _search_out(q.shape[0], kk)
    if kk:
        lib().orc_rq_search_symmetric(_p(c), n, d, bits, _p(q), q.shape[0], kk, _p(os_), _p(oi))
    return os_, oi


def rank_encode(x) -> np.ndarray:
    """Rank::add (src/rank.rs:329-350): u16 ranks, row-major."""
    x = _f32(np.atleast_2d(x))
    return np.stack([rank_transform(r) for r in x]) if len(x) else np.zeros((0, x.shape[1]), np.uint16)


def rank_norm(d: int) -> np.float32:
    return np.float32(lib().orc_rank_norm(d))


def rank_search(ranks, queries, k: int):
    """Rank::search, symmetric (src/rank.rs:351-405)."""
    r = np.ascontiguousarray(ranks, dtype=np.uint16)
    q = _f32(np.atleast_2d(queries))
    n, d = r.shape
    kk = min(k, n)
    os_, oi = _search_out(q.shape[0], kk)
    if kk:
        lib().orc_rank_search(_p(r), n, d, _p(q), q.shape[0], kk, _p(os_), _p(oi))
    return os_, oi


def rank_search_asymmetric(ranks, queries, k: int):
    """Rank::search_asymmetric (src/rank.rs:420-471)."""
    r = np.ascontiguousarray(ranks, dtype=np.uint16)
    q = _f32(np.atleast_2d(queries))
    n, d = r.shape
    kk = min(k, n)
    os_, oi = _search_out(q.shape[0], kk)
    if kk:
        lib().orc_rank_search_asymmetric(_p(r), n, d, _p(q), q.shape[0], kk, _p(os_), _p(oi))
    return os_, oi


def num_threads() -> int:
    return int(lib().orc_num_threads())
