"""Module-level rank-math helpers (ordvec-python/src/lib.rs:1488-1700) on
the host: the reference's KATs (src/rank.rs:591-820) and the oracle's
restatement on random inputs; argument checks raise ValueError."""
import numpy as np
import pytest

import ordvec_amd as ov


def test_rank_to_bucket_and_bucket_ranks(oracle):
    assert [ov.rank_to_bucket(r, 1024, 2) for r in (0, 255, 256, 511, 512, 1023)] == [0, 0, 1, 1, 2, 3]
    rng = np.random.default_rng(0)
    for d, bits in ((1024, 2), (768, 4), (100, 3), (7, 1)):
        r = rng.permutation(d).astype(np.uint16)
        assert ov.bucket_ranks(r, bits).tolist() == [oracle.rank_to_bucket(int(x), d, bits) for x in r]
        assert all(ov.rank_to_bucket(int(x), d, bits) == oracle.rank_to_bucket(int(x), d, bits) for x in r[:20])
    assert ov.bucket_ranks(np.zeros(0, np.uint16), 2).shape == (0,)
    with pytest.raises(ValueError):
        ov.rank_to_bucket(8, 8, 2)
    with pytest.raises(ValueError):
        ov.rank_to_bucket(0, 0, 2)
    with pytest.raises(ValueError):
        ov.bucket_ranks(np.array([0, 5], np.uint16), 1)
    with pytest.raises(ValueError):
        ov.rank_to_bucket(0, 8, 8)


def test_pack_unpack_and_sizes(oracle):
    b = np.array([0, 0, 1, 1, 2, 2, 3, 3], np.uint8)
    assert ov.pack_buckets(b, 2).tolist() == [0x05, 0xAF]  # tests/persistence_compat.rs:130-157
    rng = np.random.default_rng(1)
    for bits, d in ((1, 64), (2, 1024), (4, 768)):
        x = rng.integers(0, 1 << bits, size=d).astype(np.uint8)
        p = ov.pack_buckets(x, bits)
        assert np.array_equal(p, oracle.pack_buckets(x, bits))
        assert np.array_equal(ov.unpack_buckets(p, d, bits), x)
        assert ov.rankquant_bytes_per_vec(d, bits) == d * bits // 8
    with pytest.raises(ValueError):
        ov.pack_buckets(np.array([4, 0, 0, 0], np.uint8), 2)
    with pytest.raises(ValueError):
        ov.pack_buckets(np.array([0, 0, 0], np.uint8), 2)
    with pytest.raises(ValueError):
        ov.unpack_buckets(np.zeros(3, np.uint8), 10, 2)
    with pytest.raises(ValueError):
        ov.rankquant_bytes_per_vec(6, 2)
    with pytest.raises(ValueError):
        ov.pack_buckets(b, 3)


def test_centres_and_norms(oracle):
    assert [ov.bucket_centre(i, 2) for i in range(4)] == [-1.5, -0.5, 0.5, 1.5]
    assert [ov.bucket_centre(i, 1) for i in range(2)] == [-0.5, 0.5]
    assert ov.bucket_centre(7, 3) == 3.5
    with pytest.raises(ValueError):
        ov.bucket_centre(4, 2)
    for d in (2, 3, 64, 1024, 65535):
        assert ov.rank_norm(d) == float(np.float32(np.sqrt(d * (d * d - 1) / 12.0)))
    for d, bits in ((1024, 2), (768, 4), (128, 1)):
        assert ov.rankquant_norm(d, bits) == float(oracle.rankquant_norm(d, bits))
    with pytest.raises(ValueError):
        ov.rankquant_norm(64, 3)
