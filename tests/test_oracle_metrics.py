"""The oracle's non-inner-product metrics (lib.rs:2494-2698) against their
closed forms in float64, and its kNN under them against an FP64 brute force
(ties within 1e-6 relative aside).  Manhattan is checked both below 64
features (scalar sum) and above (eight lane sums, lib.rs:2518-2541)."""
import numpy as np
import pytest

from oracle import oracle as O


def _f64_dist(metric, q, r, p=3.0):
    a, b = q.astype(np.float64)[:, None, :], r.astype(np.float64)[None, :, :]
    if metric == "manhattan":
        return np.abs(a - b).sum(-1)
    if metric == "chebyshev":
        return np.abs(a - b).max(-1)
    if metric == "minkowski":
        return (np.abs(a - b) ** p).sum(-1) ** (1.0 / p)
    if metric == "canberra":
        den = np.abs(a) + np.abs(b)
        return np.where(den > 0, np.abs(a - b) / np.where(den > 0, den, 1), 0).sum(-1)
    if metric == "braycurtis":
        return np.abs(a - b).sum(-1) / np.abs(a + b).sum(-1)
    qc = q - q.mean(1, keepdims=True)
    rc = r - r.mean(1, keepdims=True)
    qn, rn = np.linalg.norm(qc, axis=1), np.linalg.norm(rc, axis=1)
    return 1.0 - np.clip(qc @ rc.T / np.outer(qn, rn), -1, 1)


@pytest.mark.parametrize("metric", ["manhattan", "chebyshev", "minkowski", "correlation",
                                    "canberra", "braycurtis"])
@pytest.mark.parametrize("d", [40, 70])
def test_metric_knn_against_fp64(metric, d):
    rng = np.random.default_rng(d)
    x = rng.normal(size=(400, d)).astype(np.float32)
    k = 12
    idx, dist = O.knn_metric(x, x, k, metric, 3.0)
    ref = _f64_dist(metric, x, x)
    got = np.take_along_axis(ref, idx, 1)
    assert np.allclose(dist, got, rtol=2e-6, atol=2e-6)
    kth = np.sort(ref, axis=1)[:, k - 1]
    for i in range(x.shape[0]):
        assert set(np.flatnonzero(ref[i] < kth[i] * (1 - 1e-6)).tolist()) <= set(idx[i].tolist())
    assert np.all(np.diff(dist, axis=1) >= 0)


def test_metric_special_cases():
    # zero-variance and zero rows: correlation 0 / 1, braycurtis 0 / 0,
    # canberra skips 0 / 0 terms (lib.rs:2565-2594, 2678-2698)
    x = np.array([[1, 1, 1, 1], [2, 2, 2, 2], [1, 2, 3, 4], [0, 0, 0, 0]], np.float32)
    _, dc = O.knn_metric(x[:2], x[:2], 2, "correlation")
    assert np.all(dc == 0.0)
    i, dc = O.knn_metric(x[2:3], x[:3], 3, "correlation")
    assert dc[0, 0] == 0.0 and np.all(dc[0, 1:] == 1.0)
    _, db = O.knn_metric(x[3:], x[3:], 1, "braycurtis")
    assert db[0, 0] == 0.0
    _, dca = O.knn_metric(x[3:], x[3:], 1, "canberra")
    assert dca[0, 0] == 0.0
