"""The one-query two-stage path (three launches: stage-1 head with the
query's sign words, sample and threshold; the vector-ALU scan with the
rerank's query preparation; the select + rerank + merge tail) against the
oracle (src/sign_bitmap.rs:231-328 top-m, src/quant.rs:1126-1281 subset
rerank) and against the general batched pipeline on the same queries: ids
and scores bit-identical.  Shapes cover every tier and width, the unsampled
case (n at most the survivor capacity), m = 1 and m = 4096, k up to 1024,
cut bins split by heavy ties, and repeated calls (the head and tail leave
their histogram and counters at zero for the next call)."""
import numpy as np
import pytest

from ordvec_amd import _native as N
from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _oracle(oracle, codes, words, qs, bits, m, k):
    cand = oracle.sign_top_m(words, oracle.sign_encode(qs), m, 8)
    offs = np.arange(len(qs) + 1, dtype=np.uint64) * cand.shape[1]
    return oracle.subset_batched(codes, qs.shape[1], bits, qs, offs, cand.reshape(-1), k, 8)


def _general(engine, sb, rq, q, m, k):
    prev = N.engine_option("one_query_path", 0)
    try:
        return engine.two_stage_search(sb, rq, q, m, k)
    finally:
        N.engine_option("one_query_path", prev)


def _same(a, b):
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(np.asarray(a[0]).view(np.uint32), np.asarray(b[0]).view(np.uint32))


CASES = [
    # (n, dim, bits, m, k, data)
    (120_000, 1024, 2, 1000, 10, "mixture"),
    (60_000, 768, 4, 300, 100, "iid"),
    (40_000, 256, 1, 64, 5, "iid"),
    (30_000, 512, 2, 4096, 1024, "mixture"),
    (50_000, 1024, 2, 1, 1, "mixture"),
    (5_000, 1024, 2, 1000, 10, "iid"),       # n <= survivor capacity: no sample
    (20_000, 64, 2, 200, 20, "iid"),         # one word per document
    (20_000, 128, 4, 200, 20, "iid"),
    (24_000, 1024, 2, 700, 10, "dups"),      # every document 40 times: split cut bins
    (8_000, 8192, 2, 500, 10, "iid"),
]


@pytest.mark.parametrize("n,dim,bits,m,k,data", CASES)
def test_one_query_path_matches_oracle_and_general(engine, oracle, n, dim, bits, m, k, data):
    import torch

    if data == "mixture":
        gen = S.MixtureNP(dim, seed=n % 97)
        docs = gen.docs(n)
        qs = gen.queries(docs, 6)[0]
    elif data == "dups":
        base = S.iid_unit(n // 40, dim, seed=9)
        docs = np.repeat(base, 40, axis=0)
        qs = S.iid_unit(6, dim, seed=10)
    else:
        docs = S.iid_unit(n, dim, seed=n)
        qs = S.iid_unit(6, dim, seed=n + 1)
    sb = engine.SignBitmap(dim)
    rq = engine.RankQuant(dim, bits)
    engine.two_stage_add(sb, rq, docs)
    codes, words = rq.codes(), sb.words()
    ws, wi = _oracle(oracle, codes, words, qs, bits, m, k)
    e0 = engine.stage1_exact_queries()
    for j in range(len(qs)):
        print(f"case n={n} dim={dim} bits={bits} m={m} k={k} {data} query {j}", flush=True)
        got = engine.two_stage_search(sb, rq, qs[j:j + 1], m, k)
        assert np.array_equal(got[1][0], wi[j]), j
        assert np.array_equal(got[0][0].view(np.uint32), ws[j].view(np.uint32)), j
        tq = torch.from_numpy(qs[j:j + 1].copy()).cuda()
        dev = engine.two_stage_search(sb, rq, tq, m, k)
        _same((dev[0].cpu().numpy(), dev[1].cpu().numpy()), got)
        _same(_general(engine, sb, rq, qs[j:j + 1], m, k), got)
    if data != "dups" and m <= 1000:
        assert engine.stage1_exact_queries() == e0


def test_one_query_path_nonfinite_and_option(engine, oracle):
    import torch

    dim = 1024
    docs = S.iid_unit(30_000, dim, seed=3)
    sb = engine.SignBitmap(dim)
    rq = engine.RankQuant(dim, 2)
    engine.two_stage_add(sb, rq, docs)
    q = S.iid_unit(2, dim, seed=4)
    ref = engine.two_stage_search(sb, rq, q[:1], 500, 10)
    bad = torch.from_numpy(q[:1].copy()).cuda()
    bad[0, 17] = float("nan")
    with pytest.raises(ValueError):
        engine.two_stage_search(sb, rq, bad, 500, 10)
    # the failed call leaves nothing behind: the next call is unchanged
    _same(engine.two_stage_search(sb, rq, q[:1], 500, 10), ref)
    assert N.engine_option("one_query_path") is True
    assert N.engine_option("mfma_scan") is True
    with pytest.raises(ValueError):
        N.check(N.lib().ordvec_engine_option(99, 1, None))


BATCH_CASES = [
    # (n, dim, bits, m, k, nq, data)
    (120_000, 1024, 2, 1000, 10, 100, "mixture"),   # rows split over 5 workgroups, wave top-k
    (60_000, 768, 4, 300, 100, 40, "iid"),          # split rows, block top-k (k > 32)
    (40_000, 256, 1, 64, 5, 7, "iid"),
    (30_000, 512, 2, 4096, 1024, 3, "mixture"),
    (24_000, 1024, 2, 700, 10, 16, "dups"),
    (8_000, 8192, 2, 500, 10, 5, "iid"),
]


@pytest.mark.parametrize("n,dim,bits,m,k,nq,data", BATCH_CASES)
def test_batch_tail_matches_oracle_and_separate_launches(engine, oracle, n, dim, bits, m, k, nq, data):
    """Batches whose stage-1 select and rerank run in one launch (the tail)
    against the oracle and against the same batch with separate select and
    rerank launches."""
    import torch

    if data == "mixture":
        gen = S.MixtureNP(dim, seed=n % 89)
        docs = gen.docs(n)
        qs = gen.queries(docs, nq)[0]
    elif data == "dups":
        docs = np.repeat(S.iid_unit(n // 40, dim, seed=19), 40, axis=0)
        qs = S.iid_unit(nq, dim, seed=20)
    else:
        docs = S.iid_unit(n, dim, seed=n + 3)
        qs = S.iid_unit(nq, dim, seed=n + 4)
    sb = engine.SignBitmap(dim)
    rq = engine.RankQuant(dim, bits)
    engine.two_stage_add(sb, rq, docs)
    print(f"batch case n={n} dim={dim} bits={bits} m={m} k={k} nq={nq} {data}", flush=True)
    ws, wi = _oracle(oracle, rq.codes(), sb.words(), qs, bits, m, k)
    got = engine.two_stage_search(sb, rq, qs, m, k)
    assert np.array_equal(got[1], wi)
    assert np.array_equal(got[0].view(np.uint32), ws.view(np.uint32))
    dev = engine.two_stage_search(sb, rq, torch.from_numpy(qs).cuda(), m, k)
    _same((dev[0].cpu().numpy(), dev[1].cpu().numpy()), got)
    prev = N.engine_option("batch_tail", 0)
    try:
        _same(engine.two_stage_search(sb, rq, qs, m, k), got)
    finally:
        N.engine_option("batch_tail", prev)
