"""Continuous-batching search server over resident SignBitmap + RankQuant
indexes.

The batching, slots and load generator are native (``ordvec_server_*`` in
``libordvec_amd.so``); this module is a thin handle around them.  A request's
result equals what ``two_stage_search`` returns for the same queries alone:
rows of a coalesced batch are cut to each request's k, and the engine's
(score desc, id asc) order makes a smaller k a prefix of a larger one.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .index import RankQuant, SignBitmap


class Server:
    """``slots`` worker contexts (own HIP stream + workspace each) serving
    two-stage searches with candidate count ``m`` and k up to ``max_k``.

    Batch capacity is the largest query count whose workspace fits
    ``memory_budget_bytes / slots`` of device memory, at most
    ``max_batch_queries``.
    """

    def __init__(self, sign: SignBitmap, rq: RankQuant, m: int = 1000, max_k: int = 100, slots: int = 4,
                 max_batch_queries: int = 1024, memory_budget_bytes: int = 8 << 30):
        h = C.c_void_p()
        N.check(N.lib().ordvec_server_start(sign._h, rq._h, int(m), int(max_k), int(slots),
                                            int(max_batch_queries), int(memory_budget_bytes), C.byref(h)))
        self._h = h
        self._keep = (sign, rq)
        self.dim = rq.dim
        self.n = len(rq)
        self._live = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    def stop(self) -> None:
        if self._h:
            N.lib().ordvec_server_stop(self._h)
            self._h = None
            self._live.clear()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass

    def info(self) -> dict:
        cap, b, q = C.c_uint64(), C.c_uint64(), C.c_uint64()
        N.check(N.lib().ordvec_server_info(self._h, C.byref(cap), C.byref(b), C.byref(q)))
        return {"max_batch_queries": cap.value, "batches": b.value, "batched_queries": q.value}

    def submit(self, queries, k: int) -> int:
        """Enqueue ``queries`` (nq, dim) f32 for top-``k``; returns a ticket."""
        q = np.ascontiguousarray(np.asarray(queries, dtype=np.float32))
        if q.ndim == 1:
            q = q[None, :]
        if q.ndim != 2 or q.shape[1] != self.dim:
            raise ValueError(f"queries must have shape (nq, {self.dim})")
        rk = min(int(k), self.n)
        s = np.empty((q.shape[0], rk), np.float32)
        i = np.empty((q.shape[0], rk), np.int64)
        t = C.c_uint64()
        N.check(N.lib().ordvec_server_submit(self._h, q.ctypes.data_as(C.c_void_p), q.shape[0], int(k),
                                             s.ctypes.data_as(C.c_void_p), i.ctypes.data_as(C.c_void_p),
                                             C.byref(t)))
        self._live[t.value] = (q, s, i)  # buffers must outlive the request
        return t.value

    def wait(self, ticket: int):
        """Block for a ticket; returns (scores, ids, latency_ms, batch_queries)."""
        lat, bq = C.c_uint64(), C.c_uint64()
        try:
            N.check(N.lib().ordvec_server_wait(self._h, int(ticket), C.byref(lat), C.byref(bq)))
        finally:
            bufs = self._live.pop(int(ticket), None)
        _, s, i = bufs
        return s, i, lat.value * 1e-6, bq.value

    def search(self, queries, k: int):
        s, i, _, _ = self.wait(self.submit(queries, k))
        return s, i

    def open_loop(self, query_pool, offered_qps: float, duration_s: float, nq_range=(1, 64),
                  k_range=(10, 100), seed: int = 0) -> dict:
        """Native open-loop load: Poisson request arrivals offering
        ``offered_qps`` queries/s for ``duration_s``; request sizes and k
        uniform over the given inclusive ranges.  Latency counts from each
        request's scheduled arrival to its completion."""
        pool = np.ascontiguousarray(np.asarray(query_pool, dtype=np.float32))
        if pool.ndim != 2 or pool.shape[1] != self.dim:
            raise ValueError(f"query_pool must have shape (n, {self.dim})")
        rep = N.ServeReport()
        rep.struct_size = C.sizeof(N.ServeReport)
        N.check(N.lib().ordvec_server_open_loop(self._h, pool.ctypes.data_as(C.c_void_p), pool.shape[0],
                                                float(offered_qps), float(duration_s), int(nq_range[0]),
                                                int(nq_range[1]), int(k_range[0]), int(k_range[1]), int(seed),
                                                C.byref(rep)))
        return {f: (getattr(rep, f) if not isinstance(getattr(rep, f), float) else float(getattr(rep, f)))
                for f, _ in N.ServeReport._fields_ if f not in ("struct_size", "reserved")}
