"""Serving: continuous batching returns, for every request, exactly what the
request gets when searched alone, and the open-loop driver completes every
request it offers."""
import threading

import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def corpus(engine):
    dim, n = 1024, 50_000
    gen = S.MixtureNP(dim, seed=5)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, 600)
    sb = engine.SignBitmap(dim)
    sb.add(docs)
    rq = engine.RankQuant(dim, 2)
    rq.add(docs)
    return docs, qs, sb, rq


def _requests(qs, rng, count):
    out = []
    for _ in range(count):
        nq = int(rng.integers(1, 65))
        start = int(rng.integers(0, len(qs) - nq + 1))
        out.append((qs[start:start + nq], int(rng.integers(10, 101))))
    return out


def test_each_result_same_alone_and_in_mix(engine, corpus, oracle):
    docs, qs, sb, rq = corpus
    rng = np.random.default_rng(21)
    reqs = _requests(qs, rng, 60)
    with engine.Server(sb, rq, m=1000, max_k=100, slots=3, max_batch_queries=256) as srv:
        # All submitted before any completes: the slots coalesce them.
        tickets = [srv.submit(q, k) for q, k in reqs]
        got = [srv.wait(t) for t in tickets]
        info = srv.info()
    assert info["batches"] < len(reqs), info  # some requests really were batched together
    assert max(g[3] for g in got) > 64
    for (q, k), (s, i, lat, bq) in zip(reqs, got):
        ws, wi = engine.two_stage_search(sb, rq, q, 1000, k)
        assert s.shape == (len(q), k)
        assert np.array_equal(i, wi)
        assert np.array_equal(s.view(np.uint32), ws.view(np.uint32))
        assert lat > 0
    # One request against the oracle as well (two-stage pipeline on the CPU).
    q, k = reqs[0]
    cand = oracle.sign_top_m(oracle.sign_encode(docs), oracle.sign_encode(q), 1000)
    offs = np.arange(len(q) + 1, dtype=np.uint64) * 1000
    ws, wi = oracle.subset_batched(oracle.rq_encode(docs, 2), 1024, 2, q, offs, cand.reshape(-1), k)
    assert np.array_equal(got[0][1], wi)
    assert np.array_equal(got[0][0].view(np.uint32), ws.view(np.uint32))


def test_concurrent_submitters(engine, corpus):
    _, qs, sb, rq = corpus
    rng = np.random.default_rng(22)
    reqs = _requests(qs, rng, 48)
    results = [None] * len(reqs)
    with engine.Server(sb, rq, m=1000, max_k=100, slots=2, max_batch_queries=512) as srv:
        def client(lo):
            for j in range(lo, len(reqs), 4):
                results[j] = srv.search(*reqs[j])

        threads = [threading.Thread(target=client, args=(c,)) for c in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    for (q, k), (s, i) in zip(reqs, results):
        ws, wi = engine.two_stage_search(sb, rq, q, 1000, k)
        assert np.array_equal(i, wi) and np.array_equal(s.view(np.uint32), ws.view(np.uint32))


def test_admission_and_request_validation(engine, corpus):
    _, qs, sb, rq = corpus
    # A tight memory budget lowers the batch capacity instead of failing.
    with engine.Server(sb, rq, m=1000, max_k=100, slots=2, max_batch_queries=1024,
                       memory_budget_bytes=2 * 40 << 20) as srv:
        cap = srv.info()["max_batch_queries"]
        assert 1 <= cap < 1024
        with pytest.raises(ValueError):
            srv.submit(np.repeat(qs[:1], cap + 1, axis=0), 10)
        with pytest.raises(ValueError):
            srv.submit(qs[:2], 101)
        bad = qs[:2].copy()
        bad[1, 3] = np.nan
        with pytest.raises(ValueError):
            srv.submit(bad, 10)
        s, i = srv.search(qs[:cap], 10)
        ws, wi = engine.two_stage_search(sb, rq, qs[:cap], 1000, 10)
        assert np.array_equal(i, wi)
    with pytest.raises(ValueError):
        engine.Server(sb, rq, m=1000, max_k=100, slots=1, memory_budget_bytes=1 << 16)


def test_open_loop_completes_every_request(engine, corpus):
    _, qs, sb, rq = corpus
    with engine.Server(sb, rq, m=1000, max_k=100, slots=3, max_batch_queries=512) as srv:
        rep = srv.open_loop(qs, offered_qps=5000, duration_s=1.0, nq_range=(1, 64), k_range=(10, 100), seed=3)
    assert rep["errors"] == 0
    assert rep["requests"] > 50
    assert rep["queries"] >= rep["requests"]
    assert 0 < rep["p50_ms"] <= rep["p99_ms"] <= rep["max_ms"]
    assert rep["batches"] >= 1 and rep["mean_batch_queries"] >= 1
    # every batch ran inside its share of the pool
    assert rep["workspace_overflows"] == 0
    assert 0 < rep["peak_pool_bytes"] <= rep["pool_bytes"]
    assert rep["max_batches_in_flight"] >= 1


def test_indexes_are_pinned_while_a_server_runs(engine, corpus):
    """add / swap_remove of a served index fail while the server runs, and
    freeing it is deferred to the server's stop (in-flight batches never see
    its rows move or its memory go)."""
    import ctypes as C

    from ordvec_amd import _native as N

    docs, qs, sb, rq = corpus
    with engine.Server(sb, rq, m=1000, max_k=10, slots=1) as srv:
        with pytest.raises(ValueError, match="running server"):
            sb.add(docs[:3])
        with pytest.raises(ValueError, match="running server"):
            rq.add(docs[:3])
        with pytest.raises(ValueError, match="running server"):
            rq.swap_remove(5)
        with pytest.raises(ValueError, match="running server"):
            sb.swap_remove(5)
        s, i = srv.search(qs[:4], 10)
    assert len(rq) == len(docs) and len(sb) == len(docs)
    ws, wi = engine.two_stage_search(sb, rq, qs[:4], 1000, 10)
    assert np.array_equal(i, wi)
    ref_rq = engine.RankQuant(1024, 2)
    ref_rq.add(docs[:3000])
    ref_sb = engine.SignBitmap(1024)
    ref_sb.add(docs[:3000])
    # a raw handle freed while served: the server keeps answering, stop frees it
    lib = N.lib()
    h_rq, h_sb = C.c_void_p(), C.c_void_p()
    N.check(lib.ordvec_rankquant_new(1024, 2, C.byref(h_rq)))
    N.check(lib.ordvec_sign_bitmap_new(1024, C.byref(h_sb)))
    x = np.ascontiguousarray(docs[:3000])
    N.check(lib.ordvec_rankquant_add(h_rq, x.ctypes.data_as(C.c_void_p), 3000, 0))
    N.check(lib.ordvec_sign_bitmap_add(h_sb, x.ctypes.data_as(C.c_void_p), 3000, 0))
    h_srv = C.c_void_p()
    N.check(lib.ordvec_server_start(h_sb, h_rq, 500, 10, 1, 64, 1 << 30, C.byref(h_srv)))
    lib.ordvec_rankquant_free(h_rq)
    lib.ordvec_sign_bitmap_free(h_sb)
    q = np.ascontiguousarray(qs[:2])
    s = np.empty((2, 10), np.float32)
    ids = np.empty((2, 10), np.int64)
    t = C.c_uint64()
    N.check(lib.ordvec_server_submit(h_srv, q.ctypes.data_as(C.c_void_p), 2, 10, s.ctypes.data_as(C.c_void_p),
                                     ids.ctypes.data_as(C.c_void_p), C.byref(t)))
    lat, bq = C.c_uint64(), C.c_uint64()
    N.check(lib.ordvec_server_wait(h_srv, t.value, C.byref(lat), C.byref(bq)))
    lib.ordvec_server_stop(h_srv)
    ws, wi = engine.two_stage_search(ref_sb, ref_rq, qs[:2], 500, 10)
    assert np.array_equal(ids, wi)
    assert np.array_equal(s.view(np.uint32), ws.view(np.uint32))
