"""Loaded-file validation: files of 64 MiB and more are checked on the GPU
after the upload; the error names the same first invalid row, with the same
message, as the host check that smaller files get (src/rank_io.rs:600-640,
764-816, 938-951).  Also the device-side interleave of .ovsb/.ovbm rows."""
import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu

BIG = 64 << 20  # payload bytes from which the GPU validates


def _u32(v):
    return int(v).to_bytes(4, "little")


def _ovrq(path, rows, dim, bits):
    path.write_bytes(b"OVRQ\x01" + bytes([bits]) + _u32(dim) + _u32(rows.shape[0]) + rows.tobytes())


def _ovr(path, ranks):
    path.write_bytes(b"OVR1\x01" + _u32(ranks.shape[1]) + _u32(ranks.shape[0]) + ranks.astype("<u2").tobytes())


def _ovbm(path, words, dim, n_top):
    path.write_bytes(b"OVBM\x01" + _u32(dim) + _u32(n_top) + _u32(words.shape[0]) + words.astype("<u8").tobytes())


def _ovsb(path, words, dim):
    path.write_bytes(b"OVSB\x01" + _u32(dim) + _u32(words.shape[0]) + words.astype("<u8").tobytes())


def _message(fn, path):
    with pytest.raises(OSError) as e:
        fn(str(path))
    return str(e.value)


def test_ovrq_rows_checked_on_device_match_host(engine, tmp_path):
    dim, bits = 1024, 2
    n_big = BIG // 256
    rows = np.full((n_big, 256), 0x1B, np.uint8)  # codes 0,1,2,3 repeating: every bucket 256 times
    ok = tmp_path / "ok.ovrq"
    _ovrq(ok, rows, dim, bits)
    ix = engine.RankQuant.load(str(ok))
    assert len(ix) == n_big and np.array_equal(ix.codes()[[0, 7, n_big - 1]], rows[[0, 7, n_big - 1]])
    bad_rows = rows.copy()
    bad_rows[200_000, 17] = 0x1F  # codes 0,1,3,3: bucket 2 short by one
    bad_rows[250_000, 3] = 0x00
    big = tmp_path / "big.ovrq"
    _ovrq(big, bad_rows, dim, bits)
    small = tmp_path / "small.ovrq"
    _ovrq(small, np.concatenate([rows[:200_000 % 4096], bad_rows[200_000:200_001]]), dim, bits)
    msg_big = _message(engine.RankQuant.load, big)
    assert "OVRQ row 200000 violates constant composition: bucket 2 has 255 codes, expected 256" in msg_big
    msg_small = _message(engine.RankQuant.load, small)
    row_small = 200_000 % 4096
    assert msg_small.replace(f"row {row_small} ", "row 200000 ") == msg_big
    for b, byte in ((1, 0x55), (4, 0x01)):
        d = 1024
        bpv = d * b // 8
        if b == 1:
            good = np.full((BIG // bpv, bpv), 0x55, np.uint8)  # half the bits set
        else:  # nibbles 0..15 once per 8 bytes: every bucket d / 16 times
            good = np.tile(np.array([0x01, 0x23, 0x45, 0x67, 0x89, 0xAB, 0xCD, 0xEF], np.uint8),
                           (BIG // bpv, bpv // 8))
        p = tmp_path / f"ok{b}.ovrq"
        _ovrq(p, good, d, b)
        assert len(engine.RankQuant.load(str(p))) == good.shape[0]
        good[123_456 % good.shape[0], 5] ^= byte if b == 1 else 0x10
        _ovrq(p, good, d, b)
        assert f"OVRQ row {123_456 % good.shape[0]} violates constant composition" in _message(engine.RankQuant.load, p)


def test_ovr_rows_checked_on_device_match_host(engine, tmp_path):
    dim = 64
    n_big = BIG // (2 * dim)
    ranks = np.tile(np.arange(dim, dtype=np.uint16), (n_big, 1))
    ok = tmp_path / "ok.ovr"
    _ovr(ok, ranks)
    rk = engine.Rank.load(str(ok))
    assert len(rk) == n_big
    q = S.iid_unit(3, dim, seed=1)
    s, i = rk.search(q, 5)
    assert i.shape == (3, 5) and np.all(i[:, 0] == 0)  # identical rows: ids in order
    rep = ranks.copy()
    rep[300_000, 5] = 6
    rep[400_000, 0] = 64
    big = tmp_path / "rep.ovr"
    _ovr(big, rep)
    assert "OVR1 row 300000 is not a permutation of [0, dim): value 6 repeats" in _message(engine.Rank.load, big)
    small = tmp_path / "rep_small.ovr"
    _ovr(small, rep[299_990:300_001])
    assert "OVR1 row 10 is not a permutation of [0, dim): value 6 repeats" in _message(engine.Rank.load, small)
    over = ranks.copy()
    over[400_000, 3] = 64
    p = tmp_path / "over.ovr"
    _ovr(p, over)
    assert "OVR1 rank value 64 >= dim (64)" in _message(engine.Rank.load, p)


def test_ovbm_rows_checked_on_device_and_interleaved(engine, oracle, tmp_path):
    dim, n_top = 1024, 256
    n_big = BIG // 128
    words = np.zeros((n_big, 16), np.uint64)
    words[:, :4] = np.uint64(0xFFFFFFFFFFFFFFFF)
    bad = words.copy()
    bad[333_333, 2] = np.uint64(0x7FFFFFFFFFFFFFFF)
    p = tmp_path / "bad.ovbm"
    _ovbm(p, bad, dim, n_top)
    assert "OVBM row 333333 has 255 bits set, expected n_top = 256" in _message(engine.Bitmap.load, p)
    # a valid large file round-trips through the device interleave
    docs = S.MixtureNP(dim, seed=5).docs(4000)
    real = oracle.bitmap_encode(docs, n_top)
    words[: real.shape[0]] = real
    ok = tmp_path / "ok.ovbm"
    _ovbm(ok, words, dim, n_top)
    bm = engine.Bitmap.load(str(ok))
    assert len(bm) == n_big
    got = bm.words()
    assert np.array_equal(got[:4000], real) and np.array_equal(got[-5:], words[-5:])
    qs = S.MixtureNP(dim, seed=5).queries(docs, 3)[0]
    top = bm.top_m_candidates_batched(qs, 50)
    assert np.array_equal(top, oracle.bitmap_top_m(words, oracle.bitmap_query(qs, n_top), 50))


def test_ovsb_large_file_interleaved_on_device(engine, oracle, tmp_path):
    dim = 1024
    n = BIG // 128 + 777  # a ragged last block
    rng = np.random.default_rng(9)
    words = rng.integers(0, 2**63, size=(n, 16), dtype=np.uint64)
    docs = S.MixtureNP(dim, seed=6).docs(2000)
    words[:2000] = oracle.sign_encode(docs)
    p = tmp_path / "big.ovsb"
    _ovsb(p, words, dim)
    sb = engine.SignBitmap.load(str(p))
    assert len(sb) == n
    assert np.array_equal(sb.words(), words)
    qs = S.MixtureNP(dim, seed=6).queries(docs, 4)[0]
    got = sb.top_m_candidates_batched(qs, 100)
    assert np.array_equal(got, oracle.sign_top_m(words, oracle.sign_encode(qs), 100, 8))
