"""Differentiable rerank: forward scores bit-identical to the search kernels and
the oracle; the backward kernel against the oracle's float64 analytic
gradient and against float64 central finite differences (h = 1e-3) of a
softmax cross-entropy through a query-side linear adapter."""
import numpy as np
import pytest

from ordvec_amd import synthetic as S

pytestmark = pytest.mark.gpu

H = 1e-3           # central-difference step
GRAD_RTOL = 1e-4   # f32 kernel gradient vs float64 analytic gradient (normwise)
FD_RTOL = 1e-4     # f32 kernel directional derivative vs float64 finite differences


@pytest.mark.parametrize("dim,bits", [(1024, 2), (768, 4), (80, 4), (48, 2), (64, 1)])
def test_score_ids_forward_bit_identical(engine, oracle, dim, bits):
    gen = S.MixtureNP(dim, seed=31)
    docs = gen.docs(3000)
    qs, _ = gen.queries(docs, 40)
    rq = engine.RankQuant(dim, bits)
    rq.add(docs)
    codes = oracle.rq_encode(docs, bits)
    rng = np.random.default_rng(dim)
    ids = rng.integers(0, 3000, size=(40, 130)).astype(np.int64)
    ids[5, :7] = -1
    s = rq.score_ids(qs, ids)
    want = oracle.score_ids(codes, bits, qs, ids)
    assert np.array_equal(s.view(np.uint32), want.view(np.uint32))
    # equals what the search returns for its own top-k
    ss, si = rq.search_asymmetric(qs, 50)
    s2 = rq.score_ids(qs, si)
    assert np.array_equal(s2.view(np.uint32), ss.view(np.uint32))
    with pytest.raises(ValueError):
        bad = ids.copy()
        bad[0, 0] = 3000
        rq.score_ids(qs, bad)


@pytest.mark.parametrize("dim,bits", [(1024, 2), (768, 4), (48, 2), (64, 1)])
def test_backward_matches_cpu_reference_gradient(engine, oracle, dim, bits):
    gen = S.MixtureNP(dim, seed=32)
    docs = gen.docs(2000)
    qs, _ = gen.queries(docs, 24)
    qs = qs * np.float32(3.0)  # raw, unnormalised queries
    qs[7] = 0.0                # zero query: zero gradient
    rq = engine.RankQuant(dim, bits)
    rq.add(docs)
    codes = oracle.rq_encode(docs, bits)
    rng = np.random.default_rng(dim + 1)
    ids = rng.integers(0, 2000, size=(24, 100)).astype(np.int64)
    ids[3, 10:20] = -1
    G = rng.standard_normal((24, 100)).astype(np.float32)
    got = rq.score_ids_backward(qs, ids, G)
    ref = oracle.score_ids_grad(codes, bits, qs, ids, G)
    assert np.all(got[7] == 0.0)
    for b in range(24):
        scale = max(np.abs(ref[b]).max(), 1e-30)
        np.testing.assert_allclose(got[b], ref[b], rtol=0, atol=GRAD_RTOL * scale)


def _ce_loss_f64(q_raw, W, cen, mask, R):
    """mean_b CE(logits_b, target = last column), logits = C_b qhat_b / R."""
    qa = q_raw @ W.T
    qh = qa / np.linalg.norm(qa, axis=1, keepdims=True)
    logits = np.einsum("bkd,bd->bk", cen, qh) / R
    logits[mask] = -np.inf
    m = logits.max(axis=1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(logits - m).sum(axis=1))
    return float(np.mean(lse - logits[:, -1]))


def test_adapter_softmax_ce_gradients(engine, oracle):
    """Small instance of the adapter workload: q = W q_raw (W = I + 0.01 N),
    top-100 by two-stage search, logits = scores of the top-100 plus the
    planted positive (masked where it is already in the top-100), CE loss."""
    import torch

    dim, n, nq, k = 1024, 20_000, 128, 100
    gen = S.MixtureNP(dim, seed=2)
    docs = gen.docs(n)
    q_raw_np, pos = gen.queries(docs, nq)
    sb = engine.SignBitmap(dim)
    sb.add(docs)
    rq = engine.RankQuant(dim, 2)
    rq.add(docs)
    codes = oracle.rq_encode(docs, 2)
    rng = np.random.default_rng(4)
    W_np = (np.eye(dim) + 0.01 * rng.standard_normal((dim, dim))).astype(np.float32)

    dev = torch.device("cuda")
    W = torch.tensor(W_np, device=dev, requires_grad=True)
    q_raw = torch.tensor(q_raw_np, device=dev, requires_grad=True)
    qa = q_raw @ W.T
    top_s, top_i = engine.two_stage_search(sb, rq, qa.detach(), 1000, k)
    ids = torch.cat([top_i, torch.tensor(pos, device=dev)[:, None]], dim=1)
    scores = engine.rerank_scores(rq, qa, ids)
    # forward: the top-k part is exactly what the search returned
    assert torch.equal(scores[:, :k], top_s)
    mask = torch.zeros_like(ids, dtype=torch.bool)
    mask[:, :k] = top_i == ids[:, k:]
    logits = scores.masked_fill(mask, float("-inf"))
    loss = torch.nn.functional.cross_entropy(logits, torch.full((nq,), k, device=dev, dtype=torch.long))
    loss.backward()
    dW = W.grad.double().cpu().numpy()
    dq_raw = q_raw.grad.double().cpu().numpy()

    # float64 CPU reference: same ids, same mask
    ids_np = ids.cpu().numpy()
    mask_np = mask.cpu().numpy()
    R = float(np.sqrt(dim * 15 / 12.0))
    cen = oracle.decode_centres(codes[ids_np], dim, 2)  # (nq, k+1, dim)
    W64, q64 = W_np.astype(np.float64), q_raw_np.astype(np.float64)
    qa64 = q64 @ W64.T
    qh = qa64 / np.linalg.norm(qa64, axis=1, keepdims=True)
    lg = np.einsum("bkd,bd->bk", cen, qh) / R
    lg[mask_np] = -np.inf
    p = np.exp(lg - lg.max(axis=1, keepdims=True))
    p /= p.sum(axis=1, keepdims=True)
    G = p.copy()
    G[:, -1] -= 1.0
    G /= nq
    G[mask_np] = 0.0
    assert abs(loss.item() - _ce_loss_f64(q64, W64, cen, mask_np, R)) < 1e-5
    dqa_ref = oracle.score_ids_grad(codes, 2, qa64.astype(np.float32), ids_np, G)
    dW_ref = dqa_ref.T @ q64
    dq_ref = dqa_ref @ W64
    assert np.abs(dW - dW_ref).max() <= GRAD_RTOL * np.abs(dW_ref).max()
    assert np.abs(dq_raw - dq_ref).max() <= GRAD_RTOL * np.abs(dq_ref).max()

    # float64 central finite differences of the loss (ids and mask fixed)
    L = lambda Wm, qm: _ce_loss_f64(qm, Wm, cen, mask_np, R)  # noqa: E731
    for t in range(3):  # random directions in W
        V = rng.standard_normal((dim, dim))
        V /= np.linalg.norm(V)
        fd = (L(W64 + H * V, q64) - L(W64 - H * V, q64)) / (2 * H)
        an = float(np.sum(dW * V))
        assert abs(fd - an) <= FD_RTOL * np.linalg.norm(dW_ref) + 1e-9, (t, fd, an)
    big = np.argsort(-np.abs(dW_ref).ravel())[:4]
    for flat in list(big) + list(rng.integers(0, dim * dim, size=4)):
        r, c = divmod(int(flat), dim)
        E = np.zeros_like(W64)
        E[r, c] = H
        fd = (L(W64 + E, q64) - L(W64 - E, q64)) / (2 * H)
        assert abs(fd - dW[r, c]) <= FD_RTOL * np.abs(dW_ref).max() + 1e-9, (r, c, fd, dW[r, c])
    for b, j in [(0, 0), (5, 17), (77, 1023), (127, 500)]:
        E = np.zeros_like(q64)
        E[b, j] = H
        fd = (L(W64, q64 + E) - L(W64, q64 - E)) / (2 * H)
        assert abs(fd - dq_raw[b, j]) <= FD_RTOL * np.abs(dq_ref).max() + 1e-9, (b, j, fd, dq_raw[b, j])
