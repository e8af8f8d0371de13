"""One-pass encode (k_encode_rows: rank codes by histogram with exact
in-bin ordering, sign words and the finite check in one read of each row)
against the oracle's rank transform + bucket + pack (src/rank.rs:38-156) and
sign bits (src/sign_bitmap.rs:180-200): codes and words bit-exact, on
ordinary rows and on rows built to hit every path of the kernel (all-equal
rows, heavy ties, signed zeros, outliers and ranges that force the
whole-row sort, subnormals, near-equal values)."""
import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _hard_rows(dim, rng):
    rows = []
    rows.append(np.full(dim, 0.5, np.float32))                       # all equal
    rows.append(np.zeros(dim, np.float32))                           # all zero
    z = np.zeros(dim, np.float32)
    z[rng.integers(0, dim, size=dim // 2)] = -0.0
    z[::7] = -0.0
    rows.append(z)                                                   # +0 / -0 only
    rows.append(rng.integers(-3, 4, size=dim).astype(np.float32))    # heavy ties
    r = rng.standard_normal(dim).astype(np.float32)
    r[5] = 1e30
    rows.append(r)                                                   # one huge outlier
    r = rng.standard_normal(dim).astype(np.float32)
    r[rng.random(dim) < 0.6] = 0.0
    rows.append(r)                                                   # sparse: many zeros
    r = rng.standard_normal(dim).astype(np.float32)
    r[0], r[1] = -3e38, 3e38
    rows.append(r)                                                   # range overflows f32
    rows.append((rng.standard_normal(dim) * 1e-40).astype(np.float32))  # subnormals
    rows.append((1.0 + np.arange(dim) % 5 * np.finfo(np.float32).eps).astype(np.float32)[rng.permutation(dim)])
    r = rng.standard_normal(dim).astype(np.float32)
    r[: dim // 4] = r[dim // 4: dim // 2]                            # duplicated values
    rows.append(r)
    rows.append(np.sort(rng.standard_normal(dim)).astype(np.float32))   # already sorted
    rows.append(np.sort(rng.standard_normal(dim)).astype(np.float32)[::-1].copy())
    return np.stack(rows)


@pytest.mark.parametrize("dim,bits", [(512, 1), (512, 2), (512, 4), (768, 2), (768, 4), (1024, 1), (1024, 2),
                                      (1024, 4), (2048, 2), (2048, 4)])
def test_one_pass_encode_matches_oracle(engine, oracle, dim, bits):
    import torch

    rng = np.random.default_rng(dim + bits)
    gen = S.MixtureNP(dim, seed=81)
    rows = np.concatenate([gen.docs(3000), _hard_rows(dim, rng),
                           rng.standard_normal((500, dim)).astype(np.float32)])
    rows = rows[rng.permutation(len(rows))]
    want_codes = oracle.rq_encode(rows, bits)
    rq = engine.RankQuant(dim, bits)
    rq.add(rows[:1000])
    rq.add(torch.from_numpy(rows[1000:]).cuda())
    assert np.array_equal(rq.codes(), want_codes)
    want_words = oracle.sign_encode(rows)
    sb = engine.SignBitmap(dim)
    sb.add(rows)
    assert np.array_equal(sb.words(), want_words)
    sb2 = engine.SignBitmap(dim)
    rq2 = engine.RankQuant(dim, bits)
    engine.two_stage_add(sb2, rq2, rows[:777])
    engine.two_stage_add(sb2, rq2, torch.from_numpy(rows[777:]).cuda())
    assert len(sb2) == len(rq2) == len(rows)
    assert np.array_equal(rq2.codes(), want_codes)
    assert np.array_equal(sb2.words(), want_words)


def test_one_pass_encode_rejects_nonfinite_and_leaves_indexes_unchanged(engine, oracle):
    import torch

    dim = 1024
    x = S.iid_unit(300, dim, seed=5)
    rq = engine.RankQuant(dim, 2)
    sb = engine.SignBitmap(dim)
    engine.two_stage_add(sb, rq, x[:100])
    before_c, before_w = rq.codes(), sb.words()
    bad = x[100:200].copy()
    bad[37, 900] = np.nan
    for v in (bad, torch.from_numpy(bad).cuda()):
        with pytest.raises(ValueError):
            rq.add(v)
        with pytest.raises(ValueError):
            sb.add(v)
        with pytest.raises(ValueError):
            engine.two_stage_add(sb, rq, v)
        assert len(rq) == 100 and len(sb) == 100
        assert np.array_equal(rq.codes(), before_c) and np.array_equal(sb.words(), before_w)
    inf = torch.from_numpy(x[100:200].copy()).cuda()
    inf[3, 3] = float("inf")
    with pytest.raises(ValueError):
        rq.add(inf)
    engine.two_stage_add(sb, rq, x[100:])
    assert np.array_equal(rq.codes(), oracle.rq_encode(x, 2))
    assert np.array_equal(sb.words(), oracle.sign_encode(x))
