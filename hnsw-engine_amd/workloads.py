"""Synthetic workloads of the benchmark configurations (seeded, numpy PCG64).

Every generator is deterministic for its seed, so the host that builds an
index and the host that checks results see the same data without shipping
files.

  uniform_128     20,000 x 128 U[0,1) (seed 0), 1,000 queries (seed 1), L2
  mixture_128     Gaussian mixture of 256 centres U[0,255]^128, per-point
                  sigma 12, clipped to [0,255] and rounded (integer-valued
                  floats), L2; 200,000 base / 10,000 queries, or 1,000,000
                  base for the 1M x 128 workload
  cosine_100      1,000 centres N(0,1)^100, x = c + 0.3 N(0,1), unit-normalised
                  then scaled by a LogNormal(0, 0.5) norm factor; cosine
  sparse_960      |N(0,1)| * 0.1 with 96 of 960 dims zeroed per vector; L2
"""

from __future__ import annotations

import numpy as np


def uniform(n: int, dim: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).random((n, dim), dtype=np.float32)


def _chunks(n: int, step: int = 65536):
    for s in range(0, n, step):
        yield s, min(n, s + step)


def mixture_128(n: int, seed: int, centers_seed: int = 2, dim: int = 128, n_centers: int = 256,
                sigma: float = 12.0) -> np.ndarray:
    """Integer-valued Gaussian mixture (base and queries share centers_seed)."""
    centers = np.random.default_rng(centers_seed).uniform(0.0, 255.0, (n_centers, dim)).astype(np.float32)
    rng = np.random.default_rng(seed)
    out = np.empty((n, dim), dtype=np.float32)
    for s, e in _chunks(n):
        c = rng.integers(0, n_centers, e - s)
        x = centers[c] + rng.standard_normal((e - s, dim), dtype=np.float32) * np.float32(sigma)
        out[s:e] = np.rint(np.clip(x, 0.0, 255.0))
    return out


def cosine_100(n: int, seed: int, centers_seed: int = 3, dim: int = 100, n_centers: int = 1000) -> np.ndarray:
    centers = np.random.default_rng(centers_seed).standard_normal((n_centers, dim), dtype=np.float32)
    rng = np.random.default_rng(seed)
    out = np.empty((n, dim), dtype=np.float32)
    for s, e in _chunks(n):
        c = rng.integers(0, n_centers, e - s)
        x = centers[c] + np.float32(0.3) * rng.standard_normal((e - s, dim), dtype=np.float32)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        x *= rng.lognormal(0.0, 0.5, (e - s, 1)).astype(np.float32)
        out[s:e] = x
    return out


def sparse_960(n: int, seed: int, dim: int = 960, zeroed: int = 96) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = np.empty((n, dim), dtype=np.float32)
    for s, e in _chunks(n, 8192):
        x = np.abs(rng.standard_normal((e - s, dim), dtype=np.float32)) * np.float32(0.1)
        idx = np.argsort(rng.random((e - s, dim)), axis=1)[:, :zeroed]
        np.put_along_axis(x, idx, 0.0, axis=1)
        out[s:e] = x
    return out


def recall(got_labels: np.ndarray, truth_labels: np.ndarray, k: int) -> float:
    """Mean fraction of the k true labels found among the k returned."""
    g = np.asarray(got_labels)[:, :k]
    t = np.asarray(truth_labels)[:, :k]
    hits = sum(len(set(a.tolist()) & set(b.tolist())) for a, b in zip(g, t))
    return hits / float(t.shape[0] * k)


def qps_at_recall(points, target: float) -> float:
    """QPS at `target` recall from (ef, recall, qps) points ordered by ef: log(QPS)
    linear in recall between the two points that bracket the target; the first
    point's QPS if it already reaches the target; NaN if none does (the method
    of the reference's bench/python/results_common.py:37-59)."""
    pts = sorted(points)
    for i, (_, rec, qps) in enumerate(pts):
        if rec >= target:
            if i == 0:
                return float(qps)
            _, r0, q0 = pts[i - 1]
            if rec <= r0:
                return float(qps)
            w = (target - r0) / (rec - r0)
            return float(np.exp(np.log(q0) + w * (np.log(qps) - np.log(q0))))
    return float("nan")
