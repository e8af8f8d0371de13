"""Python mirror of ordvec's index API (ordvec-python/src/lib.rs:584-1480,
``_ordvec.pyi:42-188``) over the MI355X engine.

Every method calls the C ABI in ``libordvec_amd.so``; indexes live in HBM.
Inputs may be NumPy arrays (any float dtype, coerced to float32 like the
reference binding) or CUDA ``torch`` tensors.  NumPy in -> NumPy out; a CUDA
tensor in -> CUDA tensors out, enqueued on torch's current stream without a
host round trip.
"""
from __future__ import annotations

import ctypes as C
from typing import Tuple

import numpy as np

from . import _native as N

try:  # torch is optional for the host-array API
    import torch
except Exception:  # pragma: no cover
    torch = None

MAX_DIM = 65535
MAX_SIGN_BITMAP_DIM = 1 << 24
MAX_VECTORS = 64 * 1024 * 1024


def _is_cuda_tensor(x) -> bool:
    return torch is not None and isinstance(x, torch.Tensor) and x.is_cuda


def _bind_stream():
    N.check(N.lib().ordvec_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream)))


def _host_stream():
    N.check(N.lib().ordvec_set_stream(None))


def _ptr(a) -> C.c_void_p:
    if _is_cuda_tensor(a):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def _float_rows(x, dim: int, name: str):
    """2-D float32 contiguous view; returns (array_or_tensor, rows, device_flag)."""
    if _is_cuda_tensor(x):
        if x.dim() != 2 or x.shape[1] != dim:
            raise ValueError(f"{name} must be 2-D with shape (n, {dim}); got {tuple(x.shape)}")
        if not x.is_floating_point():
            raise ValueError(f"{name} must be a floating-point tensor")
        x = x.to(torch.float32).contiguous()
        return x, int(x.shape[0]), N.FLAG_DEVICE_PTRS
    if torch is not None and isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    a = np.asarray(x)
    if a.ndim != 2 or a.shape[1] != dim:
        raise ValueError(f"{name} must be 2-D with shape (n, {dim}); got {a.shape}")
    if not np.issubdtype(a.dtype, np.floating):
        raise ValueError(f"{name} must be a floating-point array")
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, int(a.shape[0]), 0


def _float_vec(x, dim: int, name: str):
    if _is_cuda_tensor(x):
        if x.dim() != 1 or x.shape[0] != dim:
            raise ValueError(f"{name} must be 1-D of length {dim}")
        return x.to(torch.float32).contiguous(), N.FLAG_DEVICE_PTRS
    a = np.asarray(x.detach().cpu().numpy() if torch is not None and isinstance(x, torch.Tensor) else x)
    if a.ndim != 1 or a.shape[0] != dim:
        raise ValueError(f"{name} must be 1-D of length {dim}; got {a.shape}")
    if not np.issubdtype(a.dtype, np.floating):
        raise ValueError(f"{name} must be a floating-point array")
    return np.ascontiguousarray(a, dtype=np.float32), 0


def _ids(x, name: str):
    if _is_cuda_tensor(x):
        return x.to(torch.int32).contiguous(), N.FLAG_DEVICE_PTRS
    a = np.asarray(x.detach().cpu().numpy() if torch is not None and isinstance(x, torch.Tensor) else x)
    if a.ndim != 1:
        raise ValueError(f"{name} must be 1-D")
    if a.size and not np.issubdtype(a.dtype, np.integer):
        raise ValueError(f"{name} must be an integer array")
    if a.size and (a.min() < 0 or a.max() > 0xFFFFFFFF):
        raise ValueError(f"{name} ids must fit in u32")
    return np.ascontiguousarray(a, dtype=np.uint32), 0


def _out(shape, dtype, like_device):
    if like_device:
        tdt = {np.float32: torch.float32, np.int64: torch.int64, np.uint32: torch.int32}[dtype]
        return torch.empty(shape, dtype=tdt, device="cuda")
    return np.empty(shape, dtype=dtype)


def _swap_remove(fn, h, idx: int, n: int) -> int:
    """swap_remove: the last row moves into ``idx``; returns its old index."""
    idx = int(idx)
    if idx < 0 or idx >= n:
        raise IndexError(f"index {idx} out of range (index holds {n} vectors)")
    moved = C.c_uint64()
    _host_stream()
    N.check(fn(h, idx, C.byref(moved)))
    return moved.value


def _full_search(fn, h, queries, d: int, n: int, k: int):
    """Shared nq x min(k, n) search wrapper (host arrays or CUDA tensors)."""
    q, nq, flag = _float_rows(queries, d, "queries")
    if k < 0:
        raise ValueError("k must be non-negative")
    kk = min(int(k), n)
    s = _out((nq, kk), np.float32, flag)
    i = _out((nq, kk), np.int64, flag)
    if nq and kk:
        _bind_stream() if flag else _host_stream()
        N.check(fn(h, _ptr(q), nq, kk, _ptr(s), _ptr(i), flag))
    return s, i


class TwoStageCandidatePolicy:
    """TwoStageCandidatePolicy (src/quant.rs:235-262): stage-1 candidate count
    for a two-stage search, ``min(max(min_candidates, k * k_multiplier),
    max_candidates, search_space)`` (0 when k or the search space is 0).  The
    default is 256 / 32 / no maximum."""

    def __init__(self, min_candidates: int = 256, k_multiplier: int = 32, max_candidates=None):
        for name, v in (("min_candidates", min_candidates), ("k_multiplier", k_multiplier)):
            if int(v) < 0:
                raise ValueError(f"{name} must be non-negative")
        if max_candidates is not None and int(max_candidates) < 0:
            raise ValueError("max_candidates must be non-negative")
        self.min_candidates = int(min_candidates)
        self.k_multiplier = int(k_multiplier)
        self.max_candidates = None if max_candidates is None else int(max_candidates)

    def _c(self) -> N.TwoStagePolicy:
        p = N.TwoStagePolicy()
        p.min_candidates = self.min_candidates
        p.k_multiplier = self.k_multiplier
        p.max_candidates = self.max_candidates or 0
        p.has_max_candidates = 0 if self.max_candidates is None else 1
        return p

    def candidate_count(self, k: int, search_space: int) -> int:
        p = self._c()
        return int(N.lib().ordvec_two_stage_policy_candidate_count(C.byref(p), int(k), int(search_space)))

    def __eq__(self, other) -> bool:
        return isinstance(other, TwoStageCandidatePolicy) and (
            self.min_candidates, self.k_multiplier, self.max_candidates) == (
            other.min_candidates, other.k_multiplier, other.max_candidates)

    def __repr__(self) -> str:
        return (f"TwoStageCandidatePolicy(min_candidates={self.min_candidates}, "
                f"k_multiplier={self.k_multiplier}, max_candidates={self.max_candidates})")


class RankQuant:
    """B-bit bucketed-rank index resident in HBM (src/quant.rs:210-1017)."""

    def __init__(self, dim: int, bits: int):
        h = C.c_void_p()
        N.check(N.lib().ordvec_rankquant_new(int(dim), int(bits), C.byref(h)))
        self._h = h

    @classmethod
    def _wrap(cls, h):
        obj = cls.__new__(cls)
        obj._h = h
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and N._lib is not None:
            N.lib().ordvec_rankquant_free(h)
            self._h = None

    def _info(self):
        d, b, n = C.c_uint64(), C.c_uint32(), C.c_uint64()
        N.check(N.lib().ordvec_rankquant_info(self._h, C.byref(d), C.byref(b), C.byref(n)))
        return d.value, b.value, n.value

    @property
    def dim(self) -> int:
        return self._info()[0]

    @property
    def bits(self) -> int:
        return self._info()[1]

    def __len__(self) -> int:
        return self._info()[2]

    def is_empty(self) -> bool:
        return len(self) == 0

    @property
    def bytes_per_vec(self) -> int:
        d, b, _ = self._info()
        return d * b // 8

    @property
    def byte_size(self) -> int:
        d, b, n = self._info()
        return n * d * b // 8

    def __repr__(self) -> str:
        d, b, n = self._info()
        return f"RankQuant(dim={d}, bits={b}, n_vectors={n})"

    def add(self, vectors) -> None:
        """RankQuant::add (src/quant.rs:587-614); encode runs on the GPU."""
        x, n, flag = _float_rows(vectors, self.dim, "vectors")
        if flag:
            _bind_stream()
        else:
            _host_stream()
        N.check(N.lib().ordvec_rankquant_add(self._h, _ptr(x), n, flag))

    def search(self, queries, k: int):
        """search, symmetric (src/quant.rs:629-695): the query is rank-bucketed
        and scored against the codes; (scores f32[nq,k'], ids i64[nq,k'])."""
        d, _, n = self._info()
        return _full_search(N.lib().ordvec_rankquant_search, self._h, queries, d, n, k)

    def swap_remove(self, idx: int) -> int:
        """swap_remove (src/quant.rs:869-881); returns the moved row's old index."""
        return _swap_remove(N.lib().ordvec_rankquant_swap_remove, self._h, idx, len(self))

    def search_asymmetric(self, queries, k: int):
        """search_asymmetric (src/quant.rs:713-847): (scores f32[nq,k'], ids i64[nq,k'])."""
        d, _, n = self._info()
        q, nq, flag = _float_rows(queries, d, "queries")
        if k < 0:
            raise ValueError("k must be non-negative")
        kk = min(int(k), n)
        s = _out((nq, kk), np.float32, flag)
        i = _out((nq, kk), np.int64, flag)
        if nq and kk:
            _bind_stream() if flag else _host_stream()
            N.check(N.lib().ordvec_rankquant_search_asymmetric(self._h, _ptr(q), nq, kk, _ptr(s), _ptr(i), flag))
        elif not flag:
            s[:] = 0.0
            i[:] = -1
        return s, i

    def search_asymmetric_subset(self, query, candidates, k: int):
        """search_asymmetric_subset (src/quant.rs:1048-1083): 1-D arrays of
        length min(k, len(candidates)), ordered (score desc, global id asc)."""
        d, _, n = self._info()
        q, qflag = _float_vec(query, d, "query")
        c, cflag = _ids(candidates, "candidates")
        if qflag != cflag:
            raise ValueError("query and candidates must both be host arrays or both CUDA tensors")
        m = int(c.shape[0])
        out_k = min(int(k), m)
        s = _out((out_k,), np.float32, qflag)
        i = _out((out_k,), np.int64, qflag)
        ret = C.c_uint64()
        if qflag:
            _bind_stream()
        else:
            _host_stream()
        N.check(N.lib().ordvec_rankquant_search_subset(self._h, _ptr(q), _ptr(c) if m else None, m, int(k),
                                                       _ptr(s) if out_k else None, _ptr(i) if out_k else None,
                                                       C.byref(ret), qflag))
        return s, i

    def try_search_with_sign_probe(self, sign_probe, query, k: int, policy=None):
        """try_search_with_sign_probe[_with_policy] (src/quant.rs:1453-1497):
        the policy's candidate count of sign-bitmap candidates, then
        search_asymmetric_subset over them; 1-D arrays of length min(k, m).
        ValueError on a sign probe of another dim or length, a query of the
        wrong length or a non-finite query."""
        if not isinstance(sign_probe, SignBitmap):
            raise TypeError("sign_probe must be a SignBitmap")
        d, _, n = self._info()
        if k < 0:
            raise ValueError("k must be non-negative")
        if _is_cuda_tensor(query):
            q = query.to(torch.float32).contiguous().reshape(-1)
            flag = N.FLAG_DEVICE_PTRS
        else:
            q = np.ascontiguousarray(np.asarray(query, dtype=np.float32).reshape(-1))
            flag = 0
        qlen = int(q.shape[0])
        pol = (policy or TwoStageCandidatePolicy())._c()
        m = int(N.lib().ordvec_two_stage_policy_candidate_count(C.byref(pol), int(k), n))
        out_k = min(int(k), m)
        s = _out((out_k,), np.float32, flag)
        i = _out((out_k,), np.int64, flag)
        ret = C.c_uint64()
        _bind_stream() if flag else _host_stream()
        N.check(N.lib().ordvec_rankquant_search_with_sign_probe(
            self._h, sign_probe._h, _ptr(q), qlen, int(k), C.byref(pol), _ptr(s) if out_k else None,
            _ptr(i) if out_k else None, C.byref(ret), flag))
        return s, i

    def search_with_sign_probe(self, sign_probe, query, k: int):
        """search_with_sign_probe (src/quant.rs:1499-1507): the default policy."""
        return self.try_search_with_sign_probe(sign_probe, query, k)

    def search_asymmetric_subset_batched(self, queries, offsets, candidates, k: int):
        """search_asymmetric_subset_batched_serial (src/quant.rs:1417-1451):
        CSR candidates, rectangular nq x min(k, len) output, padded -inf/-1."""
        d, _, n = self._info()
        q, nq, flag = _float_rows(queries, d, "queries")
        if flag:
            off = offsets.to(torch.int64).contiguous()
            c = candidates.to(torch.int32).contiguous()
        else:
            off = np.ascontiguousarray(np.asarray(offsets), dtype=np.uint64)
            c, _ = _ids(candidates, "candidates")
            if off.shape[0] != nq + 1:
                raise ValueError(f"candidate_offsets length {off.shape[0]} must be nq+1 ({nq + 1})")
        nc = int(c.shape[0])
        out_k = min(int(k), n)
        s = _out((nq, out_k), np.float32, flag)
        i = _out((nq, out_k), np.int64, flag)
        if nq == 0 or out_k == 0:
            return s, i
        _bind_stream() if flag else _host_stream()
        N.check(N.lib().ordvec_rankquant_search_subset_batched(
            self._h, _ptr(q), nq, _ptr(off), _ptr(c) if nc else None, nc, int(k), _ptr(s), _ptr(i), flag))
        return s, i

    def _id_matrix(self, ids, nq: int, flag: int):
        if flag:
            if not _is_cuda_tensor(ids):
                raise ValueError("ids must be a CUDA tensor when queries are")
            t = ids.to(torch.int64).contiguous()
        else:
            t = np.ascontiguousarray(np.asarray(ids.detach().cpu().numpy() if torch is not None and
                                                isinstance(ids, torch.Tensor) else ids), dtype=np.int64)
        if t.ndim != 2 or t.shape[0] != nq:
            raise ValueError(f"ids must have shape ({nq}, k); got {tuple(t.shape)}")
        return t, int(t.shape[1])

    def score_ids(self, queries, ids, assume_finite: bool = False):
        """Scores of rows ``ids`` (nq, k) for each query, bit-identical to the
        scores the search methods return; id -1 gives -inf."""
        d, _, _ = self._info()
        q, nq, flag = _float_rows(queries, d, "queries")
        t, k = self._id_matrix(ids, nq, flag)
        out = _out((nq, k), np.float32, flag)
        if nq and k:
            _bind_stream() if flag else _host_stream()
            f = flag | (N.FLAG_ASSUME_FINITE if assume_finite else 0)
            N.check(N.lib().ordvec_rankquant_score_ids(self._h, _ptr(q), nq, _ptr(t), k, _ptr(out), f))
        return out

    def score_ids_backward(self, queries, ids, grad_scores):
        """dL/dqueries (nq, dim) of ``score_ids`` given dL/dscores (nq, k),
        from the rows' re-decoded codes."""
        d, _, _ = self._info()
        q, nq, flag = _float_rows(queries, d, "queries")
        t, k = self._id_matrix(ids, nq, flag)
        if flag:
            g = grad_scores.to(torch.float32).contiguous()
        else:
            g = np.ascontiguousarray(np.asarray(grad_scores), dtype=np.float32)
        if tuple(g.shape) != (nq, k):
            raise ValueError(f"grad_scores must have shape ({nq}, {k})")
        out = _out((nq, d), np.float32, flag)
        if nq:
            _bind_stream() if flag else _host_stream()
            N.check(N.lib().ordvec_rankquant_score_ids_backward(self._h, _ptr(q), nq, _ptr(t), k, _ptr(g),
                                                                _ptr(out), flag))
        return out

    def codes(self) -> np.ndarray:
        """Packed codes, n x bytes_per_vec (the .ovrq payload)."""
        d, b, n = self._info()
        bpv = d * b // 8
        out = np.empty((n, bpv), dtype=np.uint8)
        _host_stream()
        N.check(N.lib().ordvec_rankquant_codes(self._h, out.ctypes.data_as(C.c_void_p), out.size, 0))
        return out

    def write(self, path: str) -> None:
        _host_stream()
        N.check(N.lib().ordvec_rankquant_write(self._h, str(path).encode()))

    @classmethod
    def load(cls, path: str) -> "RankQuant":
        h = C.c_void_p()
        _host_stream()
        N.check(N.lib().ordvec_rankquant_load(str(path).encode(), C.byref(h)))
        return cls._wrap(h)


class Rank:
    """Full-precision u16 rank vectors resident in HBM (src/rank.rs:291-584)."""

    def __init__(self, dim: int):
        h = C.c_void_p()
        N.check(N.lib().ordvec_rank_new(int(dim), C.byref(h)))
        self._h = h

    @classmethod
    def _wrap(cls, h):
        obj = cls.__new__(cls)
        obj._h = h
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and N._lib is not None:
            N.lib().ordvec_rank_free(h)
            self._h = None

    def _info(self):
        d, n = C.c_uint64(), C.c_uint64()
        N.check(N.lib().ordvec_rank_info(self._h, C.byref(d), C.byref(n)))
        return d.value, n.value

    @property
    def dim(self) -> int:
        return self._info()[0]

    def __len__(self) -> int:
        return self._info()[1]

    def is_empty(self) -> bool:
        return len(self) == 0

    @property
    def bytes_per_vec(self) -> int:
        return self.dim * 2

    @property
    def byte_size(self) -> int:
        d, n = self._info()
        return n * d * 2

    def __repr__(self) -> str:
        d, n = self._info()
        return f"Rank(dim={d}, n={n})"

    def add(self, vectors) -> None:
        """add (src/rank.rs:337-356): rank transform on the GPU."""
        x, n, flag = _float_rows(vectors, self.dim, "vectors")
        _bind_stream() if flag else _host_stream()
        N.check(N.lib().ordvec_rank_add(self._h, _ptr(x), n, flag))

    def search(self, queries, k: int):
        """search (src/rank.rs:364-412): Spearman correlation of the query's
        ranks with every stored rank vector."""
        d, n = self._info()
        return _full_search(N.lib().ordvec_rank_search, self._h, queries, d, n, k)

    def search_asymmetric(self, queries, k: int):
        """search_asymmetric (src/rank.rs:429-480): L2-normalised float query
        against the stored ranks."""
        d, n = self._info()
        return _full_search(N.lib().ordvec_rank_search_asymmetric, self._h, queries, d, n, k)

    def swap_remove(self, idx: int) -> int:
        """swap_remove (src/rank.rs:506-517); returns the moved row's old index."""
        return _swap_remove(N.lib().ordvec_rank_swap_remove, self._h, idx, len(self))

    def ranks(self) -> np.ndarray:
        """The rank vectors, n x dim u16 (the .ovr payload)."""
        d, n = self._info()
        out = np.empty((n, d), dtype=np.uint16)
        _host_stream()
        N.check(N.lib().ordvec_rank_ranks(self._h, out.ctypes.data_as(C.c_void_p), out.size, 0))
        return out

    def write(self, path: str) -> None:
        _host_stream()
        N.check(N.lib().ordvec_rank_write(self._h, str(path).encode()))

    @classmethod
    def load(cls, path: str) -> "Rank":
        h = C.c_void_p()
        _host_stream()
        N.check(N.lib().ordvec_rank_load(str(path).encode(), C.byref(h)))
        return cls._wrap(h)


class SignBitmap:
    """1-bit sign code index resident in HBM (src/sign_bitmap.rs:94-660)."""

    def __init__(self, dim: int):
        h = C.c_void_p()
        N.check(N.lib().ordvec_sign_bitmap_new(int(dim), C.byref(h)))
        self._h = h

    @classmethod
    def _wrap(cls, h):
        obj = cls.__new__(cls)
        obj._h = h
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and N._lib is not None:
            N.lib().ordvec_sign_bitmap_free(h)
            self._h = None

    def _info(self):
        d, n = C.c_uint64(), C.c_uint64()
        N.check(N.lib().ordvec_sign_bitmap_info(self._h, C.byref(d), C.byref(n)))
        return d.value, n.value

    @property
    def dim(self) -> int:
        return self._info()[0]

    def __len__(self) -> int:
        return self._info()[1]

    def is_empty(self) -> bool:
        return len(self) == 0

    @property
    def bytes_per_vec(self) -> int:
        return self.dim // 8

    @property
    def byte_size(self) -> int:
        d, n = self._info()
        return n * d // 8

    def __repr__(self) -> str:
        d, n = self._info()
        return f"SignBitmap(dim={d}, n_vectors={n}, bytes_per_vector={d // 8})"

    def add(self, vectors) -> None:
        """SignBitmap::add (src/sign_bitmap.rs:180-200)."""
        x, n, flag = _float_rows(vectors, self.dim, "vectors")
        _bind_stream() if flag else _host_stream()
        N.check(N.lib().ordvec_sign_bitmap_add(self._h, _ptr(x), n, flag))

    def score_all_batched(self, queries):
        """score_all_batched (src/sign_bitmap.rs:507-560): u32 [nq, n] sign
        agreements (dim - Hamming distance), columns by document id."""
        d, n = self._info()
        q, nq, flag = _float_rows(queries, d, "queries")
        out = _out((nq, n), np.uint32, flag)
        if nq and n:
            _bind_stream() if flag else _host_stream()
            N.check(N.lib().ordvec_sign_bitmap_score_all(self._h, _ptr(q), nq, _ptr(out), flag))
        return out

    def score_all(self, query):
        """score_all (src/sign_bitmap.rs:492-505): u32 [n] sign agreements."""
        d, _ = self._info()
        q, _ = _float_vec(query, d, "query")
        return self.score_all_batched(q.reshape(1, d))[0]

    def build_query_bitmap(self, query):
        """build_query_bitmap (src/sign_bitmap.rs:204-214): u64 [dim / 64],
        bit j set iff q[j] > 0."""
        d, _ = self._info()
        q, flag = _float_vec(query, d, "query")
        out = _out((d // 64,), np.uint64, flag) if not flag else torch.empty((d // 64,), dtype=torch.int64,
                                                                              device="cuda")
        _bind_stream() if flag else _host_stream()
        N.check(N.lib().ordvec_sign_bitmap_query_bits(self._h, _ptr(q), 1, _ptr(out), flag))
        return out

    def top_m_candidates_batched(self, queries, m: int, return_hamming: bool = False):
        """top_m_candidates_batched (src/sign_bitmap.rs:369-415): nq x min(m, n)
        candidate ids ordered (hamming asc, id asc)."""
        d, n = self._info()
        q, nq, flag = _float_rows(queries, d, "queries")
        m_eff = min(int(m), n)
        c = _out((nq, m_eff), np.uint32, flag)
        h = _out((nq, m_eff), np.uint32, flag) if return_hamming else None
        if nq and m_eff:
            _bind_stream() if flag else _host_stream()
            N.check(N.lib().ordvec_sign_bitmap_top_m(self._h, _ptr(q), nq, int(m), _ptr(c),
                                                     _ptr(h) if h is not None else None, flag))
        return (c, h) if return_hamming else c

    def top_m_candidates(self, query, m: int):
        """top_m_candidates (src/sign_bitmap.rs:330-362): 1-D, length min(m, n)."""
        if _is_cuda_tensor(query):
            return self.top_m_candidates_batched(query.reshape(1, -1), m)[0]
        q = np.asarray(query)
        if q.ndim != 1:
            raise ValueError("query must be 1-D")
        return self.top_m_candidates_batched(q.reshape(1, -1), m)[0]

    def top_m_candidates_batched_serial_csr(self, queries, m: int):
        """top_m_candidates_batched_serial_csr (src/sign_bitmap.rs:453-483):
        returns (candidates u32[nq*m_eff], offsets u64[nq+1])."""
        c = self.top_m_candidates_batched(queries, m)
        nq, m_eff = c.shape
        if _is_cuda_tensor(c):
            offs = torch.arange(nq + 1, dtype=torch.int64, device=c.device) * m_eff
            return c.reshape(-1), offs
        return c.reshape(-1), np.arange(nq + 1, dtype=np.uint64) * np.uint64(m_eff)

    def swap_remove(self, idx: int) -> int:
        """swap_remove (src/sign_bitmap.rs:578-590); returns the moved row's old index."""
        return _swap_remove(N.lib().ordvec_sign_bitmap_swap_remove, self._h, idx, len(self))

    def words(self) -> np.ndarray:
        """Bitmaps row-major, n x dim/64 u64 (the .ovsb payload)."""
        d, n = self._info()
        out = np.empty((n, d // 64), dtype=np.uint64)
        _host_stream()
        N.check(N.lib().ordvec_sign_bitmap_words(self._h, out.ctypes.data_as(C.c_void_p), out.size, 0))
        return out

    def write(self, path: str) -> None:
        _host_stream()
        N.check(N.lib().ordvec_sign_bitmap_write(self._h, str(path).encode()))

    @classmethod
    def load(cls, path: str) -> "SignBitmap":
        h = C.c_void_p()
        _host_stream()
        N.check(N.lib().ordvec_sign_bitmap_load(str(path).encode(), C.byref(h)))
        return cls._wrap(h)


class Bitmap:
    """Top-bucket (constant-weight) bitmap index in HBM (src/bitmap.rs;
    ``_ordvec.pyi:109-144``): bit j of a document is set iff coordinate j is
    among its ``n_top`` highest ranks; score = popcount(Q AND D)."""

    def __init__(self, dim: int, n_top: int):
        h = C.c_void_p()
        N.check(N.lib().ordvec_bitmap_new(int(dim), int(n_top), C.byref(h)))
        self._h = h

    @classmethod
    def _wrap(cls, h):
        obj = cls.__new__(cls)
        obj._h = h
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and N._lib is not None:
            N.lib().ordvec_bitmap_free(h)
            self._h = None

    def _info(self):
        d, t, n = C.c_uint64(), C.c_uint64(), C.c_uint64()
        N.check(N.lib().ordvec_bitmap_info(self._h, C.byref(d), C.byref(t), C.byref(n)))
        return d.value, t.value, n.value

    @property
    def dim(self) -> int:
        return self._info()[0]

    @property
    def n_top(self) -> int:
        return self._info()[1]

    def __len__(self) -> int:
        return self._info()[2]

    def is_empty(self) -> bool:
        return len(self) == 0

    @property
    def bytes_per_vec(self) -> int:
        return self.dim // 8

    @property
    def byte_size(self) -> int:
        d, _, n = self._info()
        return n * d // 8

    def __repr__(self) -> str:
        d, t, n = self._info()
        return f"Bitmap(dim={d}, n_top={t}, n_vectors={n}, bytes_per_vector={d // 8})"

    def add(self, vectors) -> None:
        """Bitmap::add (src/bitmap.rs:140-162)."""
        x, n, flag = _float_rows(vectors, self.dim, "vectors")
        _bind_stream() if flag else _host_stream()
        N.check(N.lib().ordvec_bitmap_add(self._h, _ptr(x), n, flag))

    def search(self, queries, k: int):
        """Bitmap::search (src/bitmap.rs:202-250): (nq, min(k, n)) overlaps as
        f32 and i64 ids, ordered (overlap desc, id asc)."""
        d, _, n = self._info()
        q, nq, flag = _float_rows(queries, d, "queries")
        kk = min(int(k), n)
        s = _out((nq, kk), np.float32, flag)
        i = _out((nq, kk), np.int64, flag)
        if nq and kk:
            _bind_stream() if flag else _host_stream()
            N.check(N.lib().ordvec_bitmap_search(self._h, _ptr(q), nq, kk, _ptr(s), _ptr(i), flag))
        return s, i

    def top_m_candidates_batched(self, queries, m: int):
        """top_m_candidates_batched (src/bitmap.rs:295-362): nq x min(m, n) ids
        ordered (overlap desc, id asc)."""
        d, _, n = self._info()
        q, nq, flag = _float_rows(queries, d, "queries")
        m_eff = min(int(m), n)
        c = _out((nq, m_eff), np.uint32, flag)
        if nq and m_eff:
            _bind_stream() if flag else _host_stream()
            N.check(N.lib().ordvec_bitmap_top_m(self._h, _ptr(q), nq, int(m), _ptr(c), flag))
        return c

    def top_m_candidates(self, query, m: int):
        """top_m_candidates (src/bitmap.rs:264-293): 1-D, length min(m, n)."""
        if _is_cuda_tensor(query):
            return self.top_m_candidates_batched(query.reshape(1, -1), m)[0]
        q = np.asarray(query)
        if q.ndim != 1:
            raise ValueError("query must be 1-D")
        return self.top_m_candidates_batched(q.reshape(1, -1), m)[0]

    def top_m_candidates_batched_chunked(self, queries, m: int, batch_size: int):
        """top_m_candidates_batched_chunked (src/bitmap.rs:364-396): same rows as
        the batched form, computed batch_size queries at a time."""
        if int(batch_size) <= 0:
            raise ValueError("batch_size must be > 0")
        q, nq, _ = _float_rows(queries, self.dim, "queries")
        parts = [self.top_m_candidates_batched(q[i:i + int(batch_size)], m) for i in range(0, nq, int(batch_size))]
        if not parts:
            return self.top_m_candidates_batched(q, m)
        return torch.cat(parts) if _is_cuda_tensor(parts[0]) else np.concatenate(parts)

    def search_subset(self, query, doc_ids, k: int):
        """Bitmap::search_subset (src/bitmap.rs:477-505): 1-D arrays of length
        min(k, len(doc_ids)); duplicates are scored independently."""
        d, _, n = self._info()
        q, _ = _float_vec(query, d, "query")
        if _is_cuda_tensor(q):
            q = q.cpu().numpy()
        ids, _ = _ids(doc_ids.cpu() if _is_cuda_tensor(doc_ids) else doc_ids, "doc_ids")
        kk = min(int(k), ids.shape[0])
        s = np.empty(kk, np.float32)
        i = np.empty(kk, np.int64)
        if kk:
            offs = np.array([0, ids.shape[0]], dtype=np.uint64)
            _host_stream()
            N.check(N.lib().ordvec_bitmap_search_subset_batched(
                self._h, q.ctypes.data_as(C.c_void_p), 1, offs.ctypes.data_as(C.c_void_p),
                ids.ctypes.data_as(C.c_void_p), ids.shape[0], kk, s.ctypes.data_as(C.c_void_p),
                i.ctypes.data_as(C.c_void_p), None, 0))
        return s, i

    def build_query_bitmap_fp32(self, query) -> np.ndarray:
        """build_query_bitmap_fp32 (src/bitmap.rs:164-187): dim/64 u64 words."""
        d, _, _ = self._info()
        q, _ = _float_vec(query, d, "query")
        if _is_cuda_tensor(q):
            q = q.cpu().numpy()
        out = np.empty(d // 64, np.uint64)
        _host_stream()
        N.check(N.lib().ordvec_bitmap_query_bits(self._h, q.ctypes.data_as(C.c_void_p), 1,
                                                 out.ctypes.data_as(C.c_void_p), 0))
        return out

    def body_overlap_scores_subset(self, q_bitmap, doc_ids) -> np.ndarray:
        """body_overlap_scores_subset (src/bitmap.rs:425-460): u32 overlaps in
        doc_ids order."""
        d, _, _ = self._info()
        qb = np.ascontiguousarray(np.asarray(q_bitmap), dtype=np.uint64)
        if qb.shape != (d // 64,):
            raise ValueError(f"q_bitmap must have {d // 64} u64 words")
        ids, _ = _ids(doc_ids.cpu() if _is_cuda_tensor(doc_ids) else doc_ids, "doc_ids")
        out = np.empty(ids.shape[0], np.uint32)
        if ids.shape[0]:
            _host_stream()
            N.check(N.lib().ordvec_bitmap_overlap_subset(self._h, qb.ctypes.data_as(C.c_void_p),
                                                         ids.ctypes.data_as(C.c_void_p), ids.shape[0],
                                                         out.ctypes.data_as(C.c_void_p)))
        return out

    def swap_remove(self, idx: int) -> int:
        """swap_remove (src/bitmap.rs:515-527); returns the moved row's old index."""
        return _swap_remove(N.lib().ordvec_bitmap_swap_remove, self._h, idx, len(self))

    def words(self) -> np.ndarray:
        """Bitmaps row-major, n x dim/64 u64 (the .ovbm payload)."""
        d, _, n = self._info()
        out = np.empty((n, d // 64), dtype=np.uint64)
        _host_stream()
        N.check(N.lib().ordvec_bitmap_words(self._h, out.ctypes.data_as(C.c_void_p), out.size))
        return out

    def write(self, path: str) -> None:
        _host_stream()
        N.check(N.lib().ordvec_bitmap_write(self._h, str(path).encode()))

    @classmethod
    def load(cls, path: str) -> "Bitmap":
        h = C.c_void_p()
        _host_stream()
        N.check(N.lib().ordvec_bitmap_load(str(path).encode(), C.byref(h)))
        return cls._wrap(h)


class RankQuantFastscan:
    """FastScan nibble-LUT index in HBM: b=2 is the reference's
    RankQuantFastscan (src/fastscan.rs:532-705); ``bits=4`` is the b=4
    extension (one 4-bit rank code per coordinate, same LUT recipe)."""

    def __init__(self, dim: int, bits: int = 2):
        h = C.c_void_p()
        N.check(N.lib().ordvec_fastscan_new_bits(int(dim), int(bits), C.byref(h)))
        self._h = h

    @property
    def bits(self) -> int:
        b = C.c_uint32()
        N.check(N.lib().ordvec_fastscan_bits(self._h, C.byref(b)))
        return b.value

    @classmethod
    def _wrap(cls, h):
        obj = cls.__new__(cls)
        obj._h = h
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and N._lib is not None:
            N.lib().ordvec_fastscan_free(h)
            self._h = None

    def _info(self):
        d, n = C.c_uint64(), C.c_uint64()
        N.check(N.lib().ordvec_fastscan_info(self._h, C.byref(d), C.byref(n)))
        return d.value, n.value

    @property
    def dim(self) -> int:
        return self._info()[0]

    def __len__(self) -> int:
        return self._info()[1]

    def add(self, vectors) -> None:
        x, n, flag = _float_rows(vectors, self.dim, "vectors")
        _bind_stream() if flag else _host_stream()
        N.check(N.lib().ordvec_fastscan_add(self._h, _ptr(x), n, flag))

    def search(self, queries, k: int):
        d, n = self._info()
        q, nq, flag = _float_rows(queries, d, "queries")
        kk = min(int(k), n)
        s = _out((nq, kk), np.float32, flag)
        i = _out((nq, kk), np.int64, flag)
        if nq and kk:
            _bind_stream() if flag else _host_stream()
            N.check(N.lib().ordvec_fastscan_search(self._h, _ptr(q), nq, kk, _ptr(s), _ptr(i), flag))
        return s, i

    def write(self, path: str) -> None:
        _host_stream()
        N.check(N.lib().ordvec_fastscan_write(self._h, str(path).encode()))

    @classmethod
    def load(cls, path: str) -> "RankQuantFastscan":
        h = C.c_void_p()
        _host_stream()
        N.check(N.lib().ordvec_fastscan_load(str(path).encode(), C.byref(h)))
        return cls._wrap(h)


class FastscanBatch:
    """Many b=4 FastScan indexes searched as one batch: query group i (rows
    ``queries_per_index[i]``, concatenated in index order) against index i with
    its own k, in one persistent launch sequence captured as a HIP graph.

    ``run`` returns (scores, ids) of shape (total_queries, kout), row q valid
    up to its index's min(k, n) (padding -inf / -1); ``split`` cuts them into
    per-index arrays.  ``autotune`` times tile shapes on real queries and keeps
    the fastest."""

    def __init__(self, indexes, queries_per_index, k_per_index):
        self._keep = list(indexes)
        n = len(self._keep)
        arr = (C.c_void_p * n)(*[ix._h.value for ix in self._keep])
        nq = np.ascontiguousarray(queries_per_index, dtype=np.uint32)
        ks = np.ascontiguousarray(k_per_index, dtype=np.uint32)
        if nq.shape != (n,) or ks.shape != (n,):
            raise ValueError("queries_per_index and k_per_index need one entry per index")
        h = C.c_void_p()
        N.check(N.lib().ordvec_fastscan_batch_new(C.cast(arr, C.c_void_p), n, nq.ctypes.data_as(C.c_void_p),
                                                  ks.ctypes.data_as(C.c_void_p), C.byref(h)))
        self._h = h
        self.queries_per_index = nq
        self.k_per_index = np.array([min(int(k), len(ix)) for k, ix in zip(ks, self._keep)], dtype=np.int64)
        self.dim = self._keep[0].dim

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and N._lib is not None:
            N.lib().ordvec_fastscan_batch_free(h)
            self._h = None

    def config(self) -> dict:
        tq, nt, ntu = C.c_uint64(), C.c_uint64(), C.c_uint64()
        ko, qpt, bpt, grid = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        N.check(N.lib().ordvec_fastscan_batch_config(self._h, C.byref(tq), C.byref(ko), C.byref(qpt), C.byref(bpt),
                                                     C.byref(grid), C.byref(nt), C.byref(ntu)))
        return {"total_queries": tq.value, "kout": ko.value, "queries_per_tile": qpt.value,
                "blocks64_per_tile": bpt.value, "grid": grid.value, "tiles": nt.value, "tuned_configs": ntu.value}

    def set_config(self, queries_per_tile: int, blocks64_per_tile: int, grid: int) -> dict:
        """Pin the tile shape (4 or 8 queries per tile: LDS LUT lookups; 16:
        matrix-core tiles, dim % 64 == 0; 20 / 24: whole 16-query sets on the
        matrix cores and each index's remaining queries on LUT tiles of 4 / 8)
        and re-capture the graph."""
        _host_stream()
        N.check(N.lib().ordvec_fastscan_batch_set_config(self._h, int(queries_per_tile), int(blocks64_per_tile),
                                                         int(grid)))
        return self.config()

    def tuning(self) -> list:
        rows = []
        for i in range(self.config()["tuned_configs"]):
            qpt, bpt, grid, ms = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_double()
            N.check(N.lib().ordvec_fastscan_batch_tuning(self._h, i, C.byref(qpt), C.byref(bpt), C.byref(grid),
                                                         C.byref(ms)))
            rows.append({"queries_per_tile": qpt.value, "blocks64_per_tile": bpt.value, "grid": grid.value,
                         "ms": ms.value})
        return rows

    def _queries(self, queries):
        q, nq, flag = _float_rows(queries, self.dim, "queries")
        if nq != int(self.queries_per_index.sum()):
            raise ValueError(f"expected {int(self.queries_per_index.sum())} query rows, got {nq}")
        return q, nq, flag

    def autotune(self, queries, reps: int = 2) -> list:
        q, _, flag = self._queries(queries)
        _bind_stream() if flag else _host_stream()
        N.check(N.lib().ordvec_fastscan_batch_autotune(self._h, _ptr(q), int(reps), flag))
        return self.tuning()

    def run(self, queries, use_graph: bool = True):
        q, nq, flag = self._queries(queries)
        kout = self.config()["kout"]
        s = _out((nq, kout), np.float32, flag)
        i = _out((nq, kout), np.int64, flag)
        _bind_stream() if flag else _host_stream()
        f = flag | (0 if use_graph else N.FLAG_NO_GRAPH)
        N.check(N.lib().ordvec_fastscan_batch_run(self._h, _ptr(q), _ptr(s), _ptr(i), f))
        return s, i

    def split(self, scores, ids):
        """Per-index (scores, ids) arrays of shape (queries_i, min(k_i, n_i))."""
        out, r = [], 0
        for nq, k in zip(self.queries_per_index, self.k_per_index):
            out.append((scores[r:r + nq, :k], ids[r:r + nq, :k]))
            r += int(nq)
        return out


class TwoStageBatch:
    """A batched two-stage search plan over one (SignBitmap, RankQuant, m, k):
    ``run(queries)`` equals ``two_stage_search(sign, rq, queries, m, k)``, with
    the launch sequence replayed from HIP graphs captured per (query count,
    buffers) and stage-1 launch shapes picked by ``autotune``.  While the plan
    lives, ``add`` / ``swap_remove`` on the two indexes raise."""

    _SHAPE_KEYS = ("scan_tiles", "scan_pairs", "scan_flat", "sample_wgs", "side_prep")

    def __init__(self, sign: SignBitmap, rq: RankQuant, max_queries: int, m: int, k: int):
        self._keep = (sign, rq)
        h = C.c_void_p()
        N.check(N.lib().ordvec_two_stage_batch_new(sign._h, rq._h, int(max_queries), int(m), int(k), C.byref(h)))
        self._h = h
        self.dim, _, n = rq._info()
        self.max_queries = int(max_queries)
        self.out_k = min(int(k), n)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and N._lib is not None:
            N.lib().ordvec_two_stage_batch_free(h)
            self._h = None

    def close(self) -> None:
        """Frees the plan now (its indexes accept add / swap_remove again)."""
        self.__del__()

    @classmethod
    def _shape_dict(cls, sh) -> dict:
        return {key: int(getattr(sh, key)) for key in cls._SHAPE_KEYS}

    def shape(self) -> dict:
        sh = N.TwoStageShape()
        N.check(N.lib().ordvec_two_stage_batch_shape(self._h, C.byref(sh)))
        return self._shape_dict(sh)

    def set_shape(self, **shape) -> dict:
        """Pins launch shapes (keys of ``shape()``; missing keys keep their
        current value) and drops the captured graphs."""
        cur = self.shape()
        unknown = set(shape) - set(self._SHAPE_KEYS)
        if unknown:
            raise ValueError(f"unknown shape keys {sorted(unknown)}")
        cur.update({k: int(v) for k, v in shape.items()})
        sh = N.TwoStageShape(**cur)
        _bind_stream()
        N.check(N.lib().ordvec_two_stage_batch_set_shape(self._h, C.byref(sh)))
        return self.shape()

    def stats(self) -> dict:
        v = [C.c_uint64() for _ in range(5)]
        N.check(N.lib().ordvec_two_stage_batch_stats(self._h, *[C.byref(x) for x in v]))
        keys = ("graph_captures", "graph_calls", "eager_calls", "exact_queries", "tuned_shapes")
        return {key: x.value for key, x in zip(keys, v)}

    def tuning(self) -> list:
        rows = []
        for i in range(self.stats()["tuned_shapes"]):
            sh, ms = N.TwoStageShape(), C.c_double()
            N.check(N.lib().ordvec_two_stage_batch_tuning(self._h, i, C.byref(sh), C.byref(ms)))
            rows.append(dict(self._shape_dict(sh), ms=ms.value))
        return rows

    def autotune(self, queries, reps: int = 3) -> list:
        q, nq, flag = _float_rows(queries, self.dim, "queries")
        _bind_stream() if flag else _host_stream()
        N.check(N.lib().ordvec_two_stage_batch_autotune(self._h, _ptr(q), nq, int(reps), flag))
        return self.tuning()

    def run(self, queries, use_graph: bool = True, assume_finite: bool = False, out=None):
        """(scores f32[nq, min(k, n)], ids i64[nq, min(k, n)]); ``out`` may
        pass preallocated (scores, ids) of that shape on the queries' device."""
        q, nq, flag = _float_rows(queries, self.dim, "queries")
        if out is None:
            s = _out((nq, self.out_k), np.float32, flag)
            i = _out((nq, self.out_k), np.int64, flag)
        else:
            s, i = out
            for a, dt in ((s, "float32"), (i, "int64")):
                ok = (_is_cuda_tensor(a) if flag else isinstance(a, np.ndarray) and a.flags.c_contiguous)
                if not ok or tuple(a.shape) != (nq, self.out_k) or str(a.dtype).split(".")[-1] != dt:
                    raise ValueError(f"out arrays must be contiguous ({nq}, {self.out_k}) float32 / int64 "
                                     "on the queries' device")
                if flag and not a.is_contiguous():
                    raise ValueError("out tensors must be contiguous")
        if nq:
            _bind_stream() if flag else _host_stream()
            f = flag | (0 if use_graph else N.FLAG_NO_GRAPH) | (N.FLAG_ASSUME_FINITE if assume_finite else 0)
            N.check(N.lib().ordvec_two_stage_batch_run(self._h, _ptr(q), nq, _ptr(s), _ptr(i), f))
        return s, i


def two_stage_search(sign: SignBitmap, rq: RankQuant, queries, m: int, k: int,
                     assume_finite: bool = False):
    """Sign-bitmap top-m candidates then RankQuant asymmetric rerank to k
    (src/sign_bitmap.rs:453-483 + src/quant.rs:1336-1394) in one GPU call."""
    d, _, n = rq._info()
    q, nq, flag = _float_rows(queries, d, "queries")
    out_k = min(int(k), n)
    s = _out((nq, out_k), np.float32, flag)
    i = _out((nq, out_k), np.int64, flag)
    if nq and out_k:
        _bind_stream() if flag else _host_stream()
        f = flag | (N.FLAG_ASSUME_FINITE if assume_finite else 0)
        N.check(N.lib().ordvec_two_stage_search(sign._h, rq._h, _ptr(q), nq, int(m), int(k), _ptr(s), _ptr(i), f))
    return s, i


def two_stage_add(sign: SignBitmap, rq: RankQuant, vectors) -> None:
    """SignBitmap.add and RankQuant.add of the same vectors in one pass over
    them (same results as the two adds; a non-finite value leaves both
    indexes unchanged)."""
    d, _, _ = rq._info()
    x, n, flag = _float_rows(vectors, d, "vectors")
    if n == 0:
        return
    _bind_stream() if flag else _host_stream()
    N.check(N.lib().ordvec_two_stage_add(sign._h, rq._h, _ptr(x), n, flag))


def two_stage_search_policy(sign: SignBitmap, rq: RankQuant, queries, k: int, policy=None,
                            assume_finite: bool = False):
    """two_stage_search with m = policy.candidate_count(k, n) (default policy
    256 / 32 / none, src/quant.rs:235-262); rows nq x min(k, n), padded
    -inf / -1 past min(k, m)."""
    d, _, n = rq._info()
    q, nq, flag = _float_rows(queries, d, "queries")
    out_k = min(int(k), n)
    s = _out((nq, out_k), np.float32, flag)
    i = _out((nq, out_k), np.int64, flag)
    if nq and out_k:
        pol = (policy or TwoStageCandidatePolicy())._c()
        _bind_stream() if flag else _host_stream()
        f = flag | (N.FLAG_ASSUME_FINITE if assume_finite else 0)
        N.check(N.lib().ordvec_two_stage_search_policy(sign._h, rq._h, _ptr(q), nq, int(k), C.byref(pol), _ptr(s),
                                                       _ptr(i), f))
    return s, i


def stage1_exact_queries() -> int:
    """Queries whose stage-1 candidates went through the exact fallback so far
    (default context): a monitoring counter."""
    c = C.c_uint64()
    N.check(N.lib().ordvec_stage1_exact_queries(C.byref(c)))
    return c.value


def synchronize() -> None:
    N.check(N.lib().ordvec_synchronize())
