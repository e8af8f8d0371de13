"""Differentiable rerank: asymmetric scores of selected rows as a
``torch.autograd.Function``.

The forward scores rows ``ids`` (e.g. a search's top-k) for raw queries on the
GPU, exactly as the search kernels do; the backward returns dL/dqueries from a
hand-written kernel that re-decodes those rows' rank codes, so no (queries x
docs) score matrix is stored.  Codes and the candidate selection (``ids``)
receive no gradient.
"""
from __future__ import annotations

import torch

from .index import RankQuant


class RerankScores(torch.autograd.Function):
    @staticmethod
    def forward(ctx, queries: torch.Tensor, ids: torch.Tensor, rq: RankQuant):
        if not (queries.is_cuda and ids.is_cuda):
            raise ValueError("rerank_scores needs CUDA tensors")
        q = queries.detach().to(torch.float32).contiguous()
        i = ids.to(torch.int64).contiguous()
        ctx.rq = rq
        ctx.save_for_backward(q, i)
        return rq.score_ids(q, i)

    @staticmethod
    def backward(ctx, grad_scores: torch.Tensor):
        q, i = ctx.saved_tensors
        dq = ctx.rq.score_ids_backward(q, i, grad_scores.to(torch.float32).contiguous())
        return dq, None, None


def rerank_scores(rq: RankQuant, queries: torch.Tensor, ids: torch.Tensor) -> torch.Tensor:
    """(nq, k) scores of rows ``ids`` for ``queries`` (nq, dim), differentiable
    with respect to ``queries``; id -1 scores -inf and gets no gradient."""
    return RerankScores.apply(queries, ids, rq)
