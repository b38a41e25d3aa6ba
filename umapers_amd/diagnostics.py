"""Embedding quality metrics used to judge the fit and transform.

Definitions follow the reference's evaluation code:
  trustworthiness    sklearn.manifold.trustworthiness on a fixed sub-sample
                     (umap_rs/python/umapers/diagnostics.py:1-53,
                     benchmarks/compare_real_impls_fair.py:487-541)
  knn_preservation   mean overlap of the k nearest neighbours in the input and
                     in the embedding, self excluded
                     (benchmarks/compare_real_impls_fair.py:476-485)
  recall_at_k        per-row |A ∩ B| / k, self included
"""
from __future__ import annotations

import numpy as np


def _subsample(n: int, cap: int | None, seed: int) -> np.ndarray:
    if cap is None or n <= cap:
        return np.arange(n)
    return np.sort(np.random.default_rng(seed).choice(n, cap, replace=False))


def trustworthiness(x, y, n_neighbors: int = 15, cap: int | None = 10000, seed: int = 0) -> float:
    from sklearn.manifold import trustworthiness as sk_trust

    sel = _subsample(x.shape[0], cap, seed)
    return float(sk_trust(np.asarray(x)[sel], np.asarray(y)[sel], n_neighbors=n_neighbors))


def _knn_indices(z: np.ndarray, k: int) -> np.ndarray:
    from sklearn.neighbors import NearestNeighbors

    nn = NearestNeighbors(n_neighbors=k + 1, algorithm="brute").fit(z)
    return nn.kneighbors(z, return_distance=False)[:, 1:]


def knn_preservation(x, y, n_neighbors: int = 15, cap: int | None = 10000, seed: int = 0) -> float:
    sel = _subsample(x.shape[0], cap, seed)
    xs = np.asarray(x, dtype=np.float64)[sel]
    ys = np.asarray(y, dtype=np.float64)[sel]
    a = _knn_indices(xs, n_neighbors)
    b = _knn_indices(ys, n_neighbors)
    hits = sum(len(set(r1.tolist()) & set(r2.tolist())) for r1, r2 in zip(a, b))
    return hits / (n_neighbors * a.shape[0])


def recall_at_k(idx_test, idx_ref) -> float:
    k = idx_ref.shape[1]
    hits = sum(len(set(a.tolist()) & set(b.tolist())) for a, b in zip(idx_test, idx_ref))
    return hits / (k * idx_ref.shape[0])
