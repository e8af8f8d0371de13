"""Seeded synthetic corpora for the benchmark workloads.

* ``iid_unit``: rows drawn iid N(0, 1), then L2-normalised (the CPU-runnable
  128-d workload: 20,000 docs with seed 0, 256 queries with seed 1).
* ``mixture``: 4096 centroids ~ N(0, I) normalised; doc = centroid +
  0.3 * N(0, I), times a per-dimension scale ~ lognormal(0, 0.5) fixed per
  corpus, then L2-normalised.  Queries = a random doc + 0.2 * N(0, I),
  renormalised.  The query's source doc id is returned as the planted
  positive.

Host (NumPy) and device (torch, any device) generators draw from different
bit streams; a workload is identified by (generator, seed, shape) and every
parity check compares engine and oracle on the *same* materialised arrays.
"""
from __future__ import annotations

import numpy as np


def _normalise_np(x: np.ndarray) -> np.ndarray:
    n = np.linalg.norm(x, axis=1, keepdims=True)
    n[n == 0] = 1.0
    return (x / n).astype(np.float32)


def iid_unit(n: int, dim: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return _normalise_np(rng.standard_normal((n, dim), dtype=np.float32))


class MixtureNP:
    """Host mixture generator (small/medium corpora, tests)."""

    def __init__(self, dim: int, seed: int, n_centroids: int = 4096):
        self.dim = dim
        self.rng = np.random.default_rng(seed)
        self.centroids = _normalise_np(self.rng.standard_normal((n_centroids, dim), dtype=np.float32))
        self.scale = np.exp(0.5 * self.rng.standard_normal(dim)).astype(np.float32)

    def docs(self, n: int) -> np.ndarray:
        c = self.rng.integers(0, self.centroids.shape[0], size=n)
        x = self.centroids[c] + 0.3 * self.rng.standard_normal((n, self.dim), dtype=np.float32)
        return _normalise_np(x * self.scale)

    def queries(self, docs: np.ndarray, nq: int):
        src = self.rng.integers(0, docs.shape[0], size=nq)
        q = docs[src] + 0.2 * self.rng.standard_normal((nq, self.dim), dtype=np.float32)
        return _normalise_np(q), src.astype(np.int64)


class MixtureTorch:
    """Device mixture generator (1M- and 20M-doc corpora, generated in chunks
    straight into HBM)."""

    def __init__(self, dim: int, seed: int, device="cuda", n_centroids: int = 4096):
        import torch

        self.torch = torch
        self.dim = dim
        self.device = device
        self.gen = torch.Generator(device=device)
        self.gen.manual_seed(seed)
        c = torch.randn((n_centroids, dim), generator=self.gen, device=device)
        self.centroids = torch.nn.functional.normalize(c, dim=1)
        self.scale = torch.exp(0.5 * torch.randn((dim,), generator=self.gen, device=device))

    def docs(self, n: int):
        t = self.torch
        idx = t.randint(0, self.centroids.shape[0], (n,), generator=self.gen, device=self.device)
        x = self.centroids[idx] + 0.3 * t.randn((n, self.dim), generator=self.gen, device=self.device)
        return t.nn.functional.normalize(x * self.scale, dim=1).contiguous()

    def queries(self, docs, nq: int):
        t = self.torch
        src = t.randint(0, docs.shape[0], (nq,), generator=self.gen, device=self.device)
        q = docs[src] + 0.2 * t.randn((nq, self.dim), generator=self.gen, device=self.device)
        return t.nn.functional.normalize(q, dim=1).contiguous(), src
