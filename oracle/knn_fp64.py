"""TEST INFRASTRUCTURE ONLY — FP64 brute-force kNN and the kNN metrics.

Exact neighbours in float64 (direct difference for euclidean, the cosine
formula of rust_umap/src/lib.rs:2740-2765 for cosine), ordered by
(distance, index) like neighbor_cmp (rust_umap/src/lib.rs:2768-2772).
"""
from __future__ import annotations

import numpy as np


def knn_fp64(q, r=None, k: int = 15, metric: str = "euclidean", chunk: int = 1024):
    """Top-k of every query row against the reference rows, in float64.

    Returns (idx int64 [m,k], dist float64 [m,k]).  With r=None the queries
    are their own references (self included, as the reference does)."""
    q64 = np.asarray(q, dtype=np.float64)
    r64 = q64 if r is None else np.asarray(r, dtype=np.float64)
    m = q64.shape[0]
    idx = np.empty((m, k), np.int64)
    dist = np.empty((m, k), np.float64)
    if metric == "cosine":
        rn = np.sqrt((r64 * r64).sum(1))
    else:
        rsq = (r64 * r64).sum(1)
    for s in range(0, m, chunk):
        qb = q64[s:s + chunk]
        if metric == "cosine":
            qn = np.sqrt((qb * qb).sum(1))
            dots = qb @ r64.T
            denom = qn[:, None] * rn[None, :]
            with np.errstate(invalid="ignore", divide="ignore"):
                c = np.clip(dots / denom, -1.0, 1.0)
            d = 1.0 - c
            both0 = (qn[:, None] == 0) & (rn[None, :] == 0)
            one0 = (qn[:, None] == 0) ^ (rn[None, :] == 0)
            d = np.where(both0, 0.0, np.where(one0, 1.0, d))
        else:
            # expanded form for candidate screening, then exact direct
            # difference on a generous candidate set
            d = rsq[None, :] - 2.0 * (qb @ r64.T) + (qb * qb).sum(1)[:, None]
        kk = min(r64.shape[0], k + 16)
        cand = np.argpartition(d, kk - 1, axis=1)[:, :kk]
        for row in range(qb.shape[0]):
            cs = cand[row]
            if metric == "cosine":
                dv = d[row, cs]
            else:
                diff = r64[cs] - qb[row]
                dv = np.sqrt((diff * diff).sum(1))
            order = np.lexsort((cs, dv))
            # guard: the (k+16)-candidate screen must hold the true top-k
            idx[s + row] = cs[order[:k]]
            dist[s + row] = dv[order[:k]]
    return idx, dist


def recall_at_k(idx_a, idx_b) -> float:
    """Mean over rows of |A_i ∩ B_i| / k (self included)."""
    k = idx_a.shape[1]
    hits = 0
    for a, b in zip(idx_a, idx_b):
        hits += len(set(a.tolist()) & set(b.tolist()))
    return hits / (k * idx_a.shape[0])


def index_mismatches_beyond_ties(idx_test, dist_oracle64, idx_oracle, q, r=None,
                                 metric="euclidean", rel_tol=1e-6):
    """Count positions where idx_test differs from the FP64 oracle by more
    than a tie: a differing neighbour is allowed only when its FP64 distance
    is within rel_tol of the oracle's k-th distance.  Returns the number of
    rows with a non-tie difference."""
    q64 = np.asarray(q, dtype=np.float64)
    r64 = q64 if r is None else np.asarray(r, dtype=np.float64)
    bad = 0
    for i in range(idx_test.shape[0]):
        a = idx_test[i]
        b = idx_oracle[i]
        if np.array_equal(a, b):
            continue
        kth = dist_oracle64[i, -1]
        extra = set(a.tolist()) - set(b.tolist())
        ok = True
        for j in extra:
            if metric == "cosine":
                x, y = q64[i], r64[j]
                nx, ny = np.linalg.norm(x), np.linalg.norm(y)
                dj = 1.0 - np.clip(x @ y / (nx * ny), -1, 1) if nx > 0 and ny > 0 else 1.0
            else:
                dj = np.sqrt(((q64[i] - r64[j]) ** 2).sum())
            if abs(dj - kth) > rel_tol * max(kth, 1e-30):
                ok = False
                break
        if not ok:
            bad += 1
    return bad
