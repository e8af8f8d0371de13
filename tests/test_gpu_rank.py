"""Rank index, symmetric RankQuant search and swap_remove on the GPU against
the oracle: rank vectors, ids and scores bit-exact; .ovr round trips."""
import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _data(n, d, seed, nq=9, ties=False):
    gen = S.MixtureNP(d, seed=seed)
    docs = gen.docs(n)
    qs, _ = gen.queries(docs, nq)
    if ties:
        docs = np.round(docs * 6).astype(np.float32)
        docs[: n // 5] = docs[0]
        qs = np.round(qs * 6).astype(np.float32)
        qs[0, :10] = -0.0
        qs[1, :10] = 0.0
    return docs, qs


def _same(got, want, what):
    s, i = got
    ws, wi = want
    assert s.shape == ws.shape, what
    assert np.array_equal(i, wi), what
    assert np.array_equal(s.view(np.uint32), ws.view(np.uint32)), what


@pytest.mark.parametrize("d,n,ties", [(2, 50, False), (64, 3000, True), (100, 2000, False), (768, 20_000, False),
                                      (1024, 6000, True), (2048, 3000, False), (4096, 700, False)])
def test_rank_add_and_search_match_oracle(engine, oracle, d, n, ties):
    docs, qs = _data(n, d, seed=d + n, ties=ties)
    ix = engine.Rank(d)
    ix.add(docs[: n // 3])
    ix.add(docs[n // 3:])
    assert len(ix) == n and ix.dim == d and ix.byte_size == n * d * 2
    ranks = oracle.rank_encode(docs)
    assert np.array_equal(ix.ranks(), ranks)
    for k in (1, 10, 100, 1024):
        _same(ix.search(qs, k), oracle.rank_search(ranks, qs, k), ("sym", k))
        _same(ix.search_asymmetric(qs, k), oracle.rank_search_asymmetric(ranks, qs, k), ("asym", k))


def test_rank_zero_query_self_query_and_device_tensors(engine, oracle):
    import torch

    d, n = 256, 5000
    docs, qs = _data(n, d, seed=3)
    ix = engine.Rank(d)
    ix.add(docs)
    ranks = oracle.rank_encode(docs)
    z = np.zeros((2, d), np.float32)
    _same(ix.search_asymmetric(z, 5), oracle.rank_search_asymmetric(ranks, z, 5), "zero")
    _same(ix.search(z, 5), oracle.rank_search(ranks, z, 5), "zero sym")
    s, i = ix.search(docs[17:18], 3)
    assert i[0, 0] == 17 and abs(float(s[0, 0]) - 1.0) < 1e-6
    ts, ti = ix.search(torch.tensor(qs, device="cuda"), 10)
    assert ts.is_cuda and ti.is_cuda
    _same((ts.cpu().numpy(), ti.cpu().numpy()), oracle.rank_search(ranks, qs, 10), "device")
    ts, ti = ix.search_asymmetric(torch.tensor(qs, device="cuda"), 10)
    _same((ts.cpu().numpy(), ti.cpu().numpy()), oracle.rank_search_asymmetric(ranks, qs, 10), "device asym")
    bad = qs.copy()
    bad[2, 5] = np.inf
    for fn in (ix.search, ix.search_asymmetric):
        with pytest.raises(ValueError):
            fn(bad, 5)
        with pytest.raises(ValueError):
            fn(torch.tensor(bad, device="cuda"), 5)
    s, i = ix.search(qs, 0)
    assert s.shape == (len(qs), 0)
    empty = engine.Rank(d)
    s, i = empty.search(qs, 10)
    assert s.shape == (len(qs), 0) and i.shape == (len(qs), 0)


def test_rank_many_queries_cross_the_query_chunk(engine, oracle):
    d, n = 64, 300
    docs, _ = _data(n, d, seed=5)
    qs = S.iid_unit(20_000, d, seed=6)
    ix = engine.Rank(d)
    ix.add(docs)
    ranks = oracle.rank_encode(docs)
    _same(ix.search(qs, 10), oracle.rank_search(ranks, qs, 10), "sym")
    _same(ix.search_asymmetric(qs, 10), oracle.rank_search_asymmetric(ranks, qs, 10), "asym")


def test_rank_ovr_round_trip_and_swap_remove(engine, oracle, tmp_path):
    d, n = 128, 1000
    docs, qs = _data(n, d, seed=8)
    ix = engine.Rank(d)
    ix.add(docs)
    p = tmp_path / "a.ovr"
    ix.write(str(p))
    raw = p.read_bytes()
    assert raw[:5] == b"OVR1\x01" and len(raw) == 13 + n * d * 2
    assert np.array_equal(np.frombuffer(raw[13:], "<u2").reshape(n, d), oracle.rank_encode(docs))
    back = engine.Rank.load(str(p))
    assert len(back) == n and back.dim == d
    _same(back.search(qs, 10), ix.search(qs, 10), "loaded")
    legacy = tmp_path / "b.tvr"
    legacy.write_bytes(b"TVR1" + raw[4:])
    assert np.array_equal(engine.Rank.load(str(legacy)).ranks(), back.ranks())
    ranks = oracle.rank_encode(docs)
    for idx in (5, None, 0):  # None: the current last row
        idx = len(ranks) - 1 if idx is None else idx
        moved = back.swap_remove(idx)
        assert moved == len(ranks) - 1
        ranks[idx] = ranks[-1]
        ranks = ranks[:-1]
    assert np.array_equal(back.ranks(), ranks)
    _same(back.search(qs, 10), oracle.rank_search(ranks, qs, 10), "after swap_remove")
    with pytest.raises(IndexError):
        back.swap_remove(len(back))


@pytest.mark.parametrize("bits", [1, 2, 4])
@pytest.mark.parametrize("d,n", [(64, 4000), (768, 20_000), (1024, 3000), (4096, 500), (40, 900)])
def test_rankquant_symmetric_search_matches_oracle(engine, oracle, bits, d, n):
    if d % (1 << bits) or (d * bits) % 8:
        pytest.skip("dim not valid for this bit width")
    docs, qs = _data(n, d, seed=d * bits + n, ties=(d == 64))
    ix = engine.RankQuant(d, bits)
    ix.add(docs)
    codes = oracle.rq_encode(docs, bits)
    for k in (1, 10, 100, 1000):
        _same(ix.search(qs, k), oracle.rq_search_symmetric(codes, d, bits, qs, k), (bits, d, k))


def test_swap_remove_rankquant_sign_bitmap_and_bitmap(engine, oracle):
    d, n = 256, 3000
    docs, qs = _data(n, d, seed=12)
    rq = engine.RankQuant(d, 2)
    sb = engine.SignBitmap(d)
    bm = engine.Bitmap(d, 64)
    for ix in (rq, sb, bm):
        ix.add(docs)
    codes = oracle.rq_encode(docs, 2)
    words = oracle.sign_encode(docs)
    bwords = oracle.bitmap_encode(docs, 64)
    for idx in (17, None, 0, 64, 2000):  # None: the current last row
        last = len(codes) - 1
        idx = last if idx is None else idx
        for ix in (rq, sb, bm):
            assert ix.swap_remove(idx) == last
        for arr in (codes, words, bwords):
            arr[idx] = arr[last]
        codes, words, bwords = codes[:-1], words[:-1], bwords[:-1]
    assert np.array_equal(rq.codes(), codes)
    assert np.array_equal(sb.words(), words)
    assert np.array_equal(bm.words(), bwords)
    _same(rq.search_asymmetric(qs, 10), oracle.search_asymmetric(codes, d, 2, qs, 10), "rq")
    got = sb.top_m_candidates_batched(qs, 100)
    want = oracle.sign_top_m(words, oracle.sign_encode(qs), 100)
    assert np.array_equal(got, want)
    got = bm.top_m_candidates_batched(qs, 100)
    assert np.array_equal(got, oracle.bitmap_top_m(bwords, oracle.bitmap_query(qs, 64), 100))
    # adding after removals keeps the layout consistent
    more, _ = _data(100, d, seed=13)
    sb.add(more)
    bm.add(more)
    assert np.array_equal(sb.words(), np.concatenate([words, oracle.sign_encode(more)]))
    assert np.array_equal(bm.words(), np.concatenate([bwords, oracle.bitmap_encode(more, 64)]))
    for ix in (rq, sb, bm):
        with pytest.raises(IndexError):
            ix.swap_remove(10**6)
    one = engine.SignBitmap(64)
    one.add(docs[:1, :64])
    assert one.swap_remove(0) == 0 and len(one) == 0
