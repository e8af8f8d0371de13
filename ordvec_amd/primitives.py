"""Module-level rank-math primitives and the eval search
(``ordvec-python/src/lib.rs:1488-1760``, ``_ordvec.pyi:177-188``).

``rank_transform`` and ``rankquant_eval_search`` run on the GPU through the
C ABI (the same rank transform the encoders use).  The others are the
closed-form helpers of ``src/rank.rs`` (bucket of a rank, packing, centres,
norms): integer and scalar arithmetic on small host arrays, with the
reference binding's argument checks and ValueError messages.
"""
from __future__ import annotations

import math

import numpy as np

from . import _native as N
from .index import _bind_stream, _host_stream, _is_cuda_tensor, _out, _ptr, torch


def _check_bits_124(bits: int) -> None:
    if bits not in (1, 2, 4):
        raise ValueError(f"bits must be 1, 2, or 4; got {bits}")


def _check_bits_max7(bits: int) -> None:
    if not 1 <= bits <= 7:
        raise ValueError(f"bits must be in 1..=7; got {bits}")


def _f32_1d(v):
    if _is_cuda_tensor(v):
        return v.to(torch.float32).contiguous().reshape(-1), N.FLAG_DEVICE_PTRS
    a = np.ascontiguousarray(np.asarray(v.detach().cpu().numpy() if torch is not None and
                                        isinstance(v, torch.Tensor) else v))
    if a.ndim != 1:
        raise ValueError(f"expected a 1-D array, got {a.ndim}-D")
    if not np.issubdtype(a.dtype, np.floating):
        raise ValueError("expected a floating-point array")
    return np.ascontiguousarray(a, dtype=np.float32), 0


def rank_transform(v):
    """src/rank.rs:38-72: out[k] = rank of v[k] (stable ascending, ties by
    index, -0.0 == +0.0), as uint16; len(v) <= 65535, finite values."""
    x, flag = _f32_1d(v)
    d = int(x.shape[0])
    if d > 65535:
        raise ValueError(f"length {d} exceeds u16::MAX (65535)")
    out = _out((d,), np.uint16, flag) if not flag else torch.empty((d,), dtype=torch.int16, device="cuda")
    if d == 0:
        return out
    _bind_stream() if flag else _host_stream()
    N.check(N.lib().ordvec_rank_transform(_ptr(x), 1, d, _ptr(out), flag))
    return out


def rank_to_bucket(rank: int, d: int, bits: int) -> int:
    """src/rank.rs:89-101: floor(rank * 2^bits / d)."""
    _check_bits_max7(bits)
    if d == 0:
        raise ValueError("d must be > 0")
    if not 0 <= rank < d:
        raise ValueError(f"rank ({rank}) must be < d ({d})")
    return (int(rank) << bits) // int(d)


def bucket_ranks(ranks, bits: int):
    """src/rank.rs:103-112: rank_to_bucket of every entry of a rank vector."""
    _check_bits_max7(bits)
    r = np.ascontiguousarray(np.asarray(ranks, dtype=np.uint16))
    d = r.shape[0]
    if d == 0:
        return np.zeros(0, np.uint8)
    if int(r.max()) >= d:
        raise ValueError(f"rank ({int(r[r >= d][0])}) must be < d ({d})")
    return ((r.astype(np.uint64) << np.uint64(bits)) // np.uint64(d)).astype(np.uint8)


def pack_buckets(buckets, bits: int):
    """src/rank.rs:128-140: 8 / bits codes per byte, first code in the MSB."""
    _check_bits_124(bits)
    b = np.ascontiguousarray(np.asarray(buckets, dtype=np.uint8))
    cpb = 8 // bits
    if b.shape[0] % cpb:
        raise ValueError(f"len {b.shape[0]} must be a multiple of {cpb} for bits = {bits}")
    if b.size and int(b.max()) >= (1 << bits):
        raise ValueError(f"bucket value {int(b[b >= (1 << bits)][0])} out of range [0, {1 << bits}) for bits = {bits}")
    g = b.reshape(-1, cpb).astype(np.uint32)
    shifts = np.array([(cpb - 1 - j) * bits for j in range(cpb)], np.uint32)
    return (g << shifts).sum(axis=1).astype(np.uint8)


def unpack_buckets(packed, d: int, bits: int):
    """src/rank.rs:142-156: inverse of pack_buckets."""
    _check_bits_124(bits)
    p = np.ascontiguousarray(np.asarray(packed, dtype=np.uint8))
    cpb = 8 // bits
    if p.shape[0] * cpb != d:
        raise ValueError(f"packed length {p.shape[0]} * {cpb} codes/byte != d = {d}")
    shifts = np.array([(cpb - 1 - j) * bits for j in range(cpb)], np.uint8)
    return ((p[:, None] >> shifts) & ((1 << bits) - 1)).reshape(-1).astype(np.uint8)


def rankquant_bytes_per_vec(d: int, bits: int) -> int:
    """src/rank.rs:158-166: d * bits / 8."""
    _check_bits_124(bits)
    cpb = 8 // bits
    if d % cpb:
        raise ValueError(f"d {d} must be a multiple of {cpb} for bits = {bits}")
    return d // cpb


def bucket_centre(bucket: int, bits: int) -> float:
    """src/rank.rs:220-228: bucket - (2^bits - 1) / 2, as f32."""
    _check_bits_max7(bits)
    if not 0 <= bucket < (1 << bits):
        raise ValueError(f"bucket {bucket} out of range [0, {1 << bits}) for bits = {bits}")
    return float(np.float32(bucket) - np.float32(((1 << bits) - 1) / 2))


def rank_norm(d: int) -> float:
    """src/rank.rs:249-252: sqrt(d (d^2 - 1) / 12), computed in f64, as f32."""
    x = float(d)
    return float(np.float32(math.sqrt(x * (x * x - 1.0) / 12.0)))


def rankquant_norm(d: int, bits: int) -> float:
    """src/rank.rs:275-280: sqrt(d (4^bits - 1) / 12), f64 then f32."""
    _check_bits_124(bits)
    nb = float(1 << bits)
    return float(np.float32(math.sqrt(float(d) * (nb * nb - 1.0) / 12.0)))


def rankquant_eval_search(corpus, queries, bits: int, k: int):
    """rankquant_eval_search (src/quant.rs:1541-1610): symmetric search at
    any width 1..7, both sides rank-bucketed on the fly; (scores f32[nq, k'],
    ids i64[nq, k']) with k' = min(k, n)."""
    if not 1 <= bits <= 7:
        raise ValueError(f"bits must be in 1..=7; got {bits}")
    if _is_cuda_tensor(corpus):
        cx = corpus.to(torch.float32).contiguous()
        dim = int(cx.shape[1]) if cx.dim() == 2 else -1
    else:
        cx = np.ascontiguousarray(np.asarray(corpus), dtype=np.float32)
        dim = int(cx.shape[1]) if cx.ndim == 2 else -1
    if not 2 <= dim <= 65535:
        raise ValueError("corpus width must be in [2, 65535]")
    from .index import _float_rows

    c, n, cflag = _float_rows(cx, dim, "corpus")
    q, nq, qflag = _float_rows(queries, dim, "queries")
    if cflag != qflag:
        raise ValueError("corpus and queries must both be host arrays or both CUDA tensors")
    kk = min(int(k), n)
    s = _out((nq, kk), np.float32, qflag)
    i = _out((nq, kk), np.int64, qflag)
    if nq and kk:
        _bind_stream() if qflag else _host_stream()
        N.check(N.lib().ordvec_rankquant_eval_search(_ptr(c), n, _ptr(q), nq, dim, int(bits), int(k), _ptr(s),
                                                     _ptr(i), qflag))
    return s, i


__all__ = ["rank_transform", "rank_to_bucket", "bucket_ranks", "pack_buckets", "unpack_buckets",
           "rankquant_bytes_per_vec", "bucket_centre", "rank_norm", "rankquant_norm", "rankquant_eval_search"]
