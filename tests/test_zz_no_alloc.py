"""The allocation-free two-stage entry point (ordvec_two_stage_search_noalloc)
against ordvec_two_stage_search on the same device inputs: ids and score bits
identical, on two calls in a row (the second smaller), on a corpus whose
queries take the exact stage-1 recomputation, and the size limits rejected on
the host before any work is enqueued.  Calls go through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _noalloc(engine, sb, rq, q, m, k):
    import torch

    from ordvec_amd import _native as N

    nq = q.shape[0]
    kk = min(k, len(rq))
    s = torch.empty((nq, kk), dtype=torch.float32, device="cuda")
    i = torch.empty((nq, kk), dtype=torch.int64, device="cuda")
    N.check(N.lib().ordvec_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    st = N.lib().ordvec_two_stage_search_noalloc(sb._h, rq._h, C.c_void_p(q.data_ptr()), nq, m, k,
                                                 C.c_void_p(s.data_ptr()), C.c_void_p(i.data_ptr()),
                                                 N.FLAG_DEVICE_PTRS)
    return st, s, i


def _reference(engine, sb, rq, q, m, k):
    import torch

    s, i = engine.two_stage_search(sb, rq, q, m, k)
    torch.cuda.synchronize()
    return s, i


def _same(a, b):
    sa, ia = a
    sb, ib = b
    assert np.array_equal(ia.cpu().numpy(), ib.cpu().numpy())
    assert np.array_equal(sa.cpu().numpy().view(np.uint32), sb.cpu().numpy().view(np.uint32))


def test_noalloc_equals_two_stage_on_consecutive_calls(engine):
    import torch

    from ordvec_amd import _native as N

    N.check(N.lib().ordvec_init())
    d, n = 1024, 60_000
    gen = S.MixtureNP(d, seed=31)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 300)
    sb = engine.SignBitmap(d)
    sb.add(docs)
    rq = engine.RankQuant(d, 2)
    rq.add(docs)
    tq = torch.from_numpy(qs).cuda()
    # a matrix-core batch, then a smaller one (VALU scan, split rerank), then one query
    for nq, m, k in ((300, 1000, 10), (7, 500, 25), (1, 1000, 10)):
        q = tq[:nq].contiguous()
        st, s, i = _noalloc(engine, sb, rq, q, m, k)
        assert st == N.OK, N.last_error()
        _same((s, i), _reference(engine, sb, rq, q, m, k))


def test_noalloc_exact_recomputation_path(engine):
    """Queries near 20k identical documents overflow the survivor buffer and
    are recomputed by the exact pipeline inside the fixed scratch."""
    import torch

    from ordvec_amd import _native as N

    rng = np.random.default_rng(17)
    dim = 256
    gen = S.MixtureNP(dim, seed=19)
    docs = gen.docs(40_000)
    dup = docs[123].copy()
    docs[1000:21000] = dup
    qs, _ = gen.queries(docs, 40)
    qs[::7] = dup + 0.001 * rng.standard_normal((len(qs[::7]), dim)).astype(np.float32)
    sb = engine.SignBitmap(dim)
    sb.add(docs)
    rq = engine.RankQuant(dim, 2)
    rq.add(docs)
    tq = torch.from_numpy(qs).cuda()
    for nq in (40, 9):
        q = tq[:nq].contiguous()
        e0 = engine.stage1_exact_queries()
        st, s, i = _noalloc(engine, sb, rq, q, 700, 10)
        assert st == N.OK, N.last_error()
        assert engine.stage1_exact_queries() > e0
        _same((s, i), _reference(engine, sb, rq, q, 700, 10))


def test_noalloc_limits_are_rejected_before_any_work(engine):
    import torch

    from ordvec_amd import _native as N

    d = 1024
    x = S.iid_unit(70_000, d, seed=3)
    sb = engine.SignBitmap(d)
    sb.add(x)
    rq = engine.RankQuant(d, 2)
    rq.add(x)
    q = torch.from_numpy(np.ascontiguousarray(x[:1025])).cuda()
    # more queries than the entry point takes
    st, _, _ = _noalloc(engine, sb, rq, q, 1000, 10)
    assert st == N.UNSUPPORTED_OPERATION and "at most 1024" in N.last_error()
    # 1024 queries x 64 KiB of survivor keys exceed the 64 MiB scratch
    st, _, _ = _noalloc(engine, sb, rq, q[:1024].contiguous(), 1000, 10)
    assert st == N.UNSUPPORTED_OPERATION and "bytes of scratch" in N.last_error()
    # a call that fits still runs after the rejections
    st, s, i = _noalloc(engine, sb, rq, q[:64].contiguous(), 1000, 10)
    assert st == N.OK, N.last_error()
    _same((s, i), _reference(engine, sb, rq, q[:64].contiguous(), 1000, 10))
