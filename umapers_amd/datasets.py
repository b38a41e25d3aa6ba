"""Seeded synthetic workloads (the five benchmark configurations).

Each generator draws with ``numpy.random.default_rng(seed)`` in a fixed call
order and returns float32 rows, so the CPU baseline and the GPU runs consume
identical bytes (``sha256`` below fingerprints them).

  c1  1000 x 32 i.i.d. N(0, 1), seed 42 (the reference README quickstart)
  c2  20000 x 192, 10-cluster mixture: centres N(0, 4^2 I), points N(c, I), seed 7
  c3  200000 x 128, 32-cluster mixture: centres N(0, 3^2 I), cluster std
      U(0.5, 1.5), seed 11 (cosine variant: rows L2-normalised)
  c4  1000000 x 64, 50-cluster anisotropic mixture: centres N(0, 5^2 I),
      A_c with N(0, 0.3^2) entries, x = c + A_c z, seed 23
  c5  100000 x 192 fit rows (20 clusters, centres N(0, 4^2 I), unit variance,
      seed 31) and query rows from the same centres (seed 32)
"""
from __future__ import annotations

import hashlib

import numpy as np


def sha256(x: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(x).view(np.uint8)).hexdigest()


def c1(n: int = 1000, d: int = 32, seed: int = 42) -> np.ndarray:
    return np.random.default_rng(seed).normal(size=(n, d)).astype(np.float32)


def gaussian_mixture(n: int, d: int, n_clusters: int, center_std: float, seed: int,
                     point_std=1.0, return_labels: bool = False):
    rng = np.random.default_rng(seed)
    centers = rng.normal(0.0, center_std, size=(n_clusters, d))
    if isinstance(point_std, tuple):
        stds = rng.uniform(point_std[0], point_std[1], size=n_clusters)
    else:
        stds = np.full(n_clusters, float(point_std))
    labels = rng.integers(0, n_clusters, size=n)
    x = np.empty((n, d), np.float32)
    chunk = 65536
    for s in range(0, n, chunk):
        lab = labels[s:s + chunk]
        z = rng.normal(0.0, 1.0, size=(lab.shape[0], d))
        x[s:s + chunk] = (centers[lab] + stds[lab, None] * z).astype(np.float32)
    return (x, labels) if return_labels else x


def mixture_queries(m: int, d: int, n_clusters: int, center_std: float, seed: int, point_std,
                    query_seed: int) -> np.ndarray:
    """Fresh rows from the mixture that gaussian_mixture(.., seed, point_std)
    draws: its centres and cluster scales, new labels and noise from
    query_seed."""
    rng = np.random.default_rng(seed)
    centers = rng.normal(0.0, center_std, size=(n_clusters, d))
    if isinstance(point_std, tuple):
        stds = rng.uniform(point_std[0], point_std[1], size=n_clusters)
    else:
        stds = np.full(n_clusters, float(point_std))
    rq = np.random.default_rng(query_seed)
    labels = rq.integers(0, n_clusters, size=m)
    x = np.empty((m, d), np.float32)
    chunk = 65536
    for s in range(0, m, chunk):
        lab = labels[s:s + chunk]
        z = rq.normal(0.0, 1.0, size=(lab.shape[0], d))
        x[s:s + chunk] = (centers[lab] + stds[lab, None] * z).astype(np.float32)
    return x


def c3_queries(m: int = 20000, d: int = 128, query_seed: int = 111) -> np.ndarray:
    """Transform queries for the c3 model: same centres, new draws."""
    return mixture_queries(m, d, 32, 3.0, 11, (0.5, 1.5), query_seed)


def c2(n: int = 20000, d: int = 192, return_labels: bool = False):
    return gaussian_mixture(n, d, 10, 4.0, 7, 1.0, return_labels)


def c3(n: int = 200000, d: int = 128, cosine: bool = False, return_labels: bool = False):
    out = gaussian_mixture(n, d, 32, 3.0, 11, (0.5, 1.5), return_labels)
    if not cosine:
        return out
    x, lab = out if return_labels else (out, None)
    nrm = np.linalg.norm(x.astype(np.float64), axis=1, keepdims=True)
    nrm[nrm == 0] = 1.0
    x = (x / nrm).astype(np.float32)
    return (x, lab) if return_labels else x


def c4(n: int = 1_000_000, d: int = 64, return_labels: bool = False):
    rng = np.random.default_rng(23)
    k = 50
    centers = rng.normal(0.0, 5.0, size=(k, d))
    a = rng.normal(0.0, 0.3, size=(k, d, d))
    labels = rng.integers(0, k, size=n)
    x = np.empty((n, d), np.float32)
    chunk = 65536
    for s in range(0, n, chunk):
        lab = labels[s:s + chunk]
        z = rng.normal(0.0, 1.0, size=(lab.shape[0], d))
        x[s:s + chunk] = (centers[lab] + np.einsum("nij,nj->ni", a[lab], z)).astype(np.float32)
    return (x, labels) if return_labels else x


def c5(n_fit: int = 100_000, n_query: int = 1_000_000, d: int = 192, return_labels: bool = False):
    rng = np.random.default_rng(31)
    centers = rng.normal(0.0, 4.0, size=(20, d))
    lab_fit = rng.integers(0, 20, size=n_fit)
    x_fit = (centers[lab_fit] + rng.normal(size=(n_fit, d))).astype(np.float32)
    rq = np.random.default_rng(32)
    lab_q = rq.integers(0, 20, size=n_query)
    x_q = np.empty((n_query, d), np.float32)
    chunk = 65536
    for s in range(0, n_query, chunk):
        lab = lab_q[s:s + chunk]
        x_q[s:s + chunk] = (centers[lab] + rq.normal(size=(lab.shape[0], d))).astype(np.float32)
    if return_labels:
        return x_fit, lab_fit, x_q, lab_q
    return x_fit, x_q
