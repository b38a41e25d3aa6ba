"""The headline workload at full size: 200000 x 128, 32-cluster mixture
(seed 11), k = 15, 200 epochs, euclidean and cosine (rows L2-normalised),
FP32 and BF16+rerank candidates.

  - kNN: recall@15 against an FP64 brute force on 1000 sampled rows (self
    included, as the reference's kNN is) >= 0.99, and the largest relative
    distance error;
  - fit: mean trustworthiness@15 and kNN preservation@15 (10000-row
    sub-sample) over the oracle's recorded seeds within 0.01 of the oracle's
    mean, both ways (tests/golden/quality.json, c3_*_200)."""
import json
import os

import numpy as np
import pytest

from oracle.knn_fp64 import index_mismatches_beyond_ties, knn_fp64
from umapers_amd import Umap, datasets
from umapers_amd.diagnostics import knn_preservation, trustworthiness

from . import capi

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_X = {}


def _x(metric):
    if metric not in _X:
        _X[metric] = datasets.c3(cosine=(metric == "cosine"))
    return _X[metric]


def _fp64_pair_dist(x, rows, idx, metric):
    """FP64 distance of each row to each of its returned neighbours."""
    q = x[rows].astype(np.float64)
    nb = x[idx[rows]].astype(np.float64)  # [r, k, d]
    if metric == "cosine":
        dots = np.einsum("rd,rkd->rk", q, nb)
        nq = np.linalg.norm(q, axis=1)[:, None]
        nn = np.linalg.norm(nb, axis=2)
        return 1.0 - np.clip(dots / (nq * nn), -1.0, 1.0)
    return np.sqrt(((nb - q[:, None, :]) ** 2).sum(2))


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
@pytest.mark.parametrize("precision", [0, 1])
def test_c3_knn_recall_against_fp64(gpu, metric, precision):
    """recall@15 >= 0.99 against FP64 (both precisions), FP32 euclidean
    indices equal to FP64 except at ties within 1e-6 relative, and each
    returned distance equal to its pair's FP64 distance to the f32
    arithmetic's rounding (the reference's own: relative 1e-5 for euclidean,
    absolute 5e-6 for cosine's 1 - clamp(x.y / (|x||y|)))."""
    x = _x(metric)
    print(f"c3 kNN metric={metric} precision={precision}", flush=True)
    idx, dist = capi.knn(x, 15, 0 if metric == "euclidean" else 1, precision, 2)
    rows = np.sort(np.random.default_rng(5).choice(x.shape[0], 1000, replace=False))
    i64, d64 = knn_fp64(x[rows], x, 15, metric)
    recall = np.mean([len(set(idx[r]) & set(i64[j])) / 15 for j, r in enumerate(rows)])
    order_err = np.max(np.abs(dist[rows].astype(np.float64) - d64) / np.maximum(d64, 1e-12))
    pair = _fp64_pair_dist(x, rows, idx, metric)
    dev = np.abs(dist[rows].astype(np.float64) - pair)
    print(f"  recall@15 {recall:.5f}; k-th order distance vs FP64: max rel {order_err:.3g}; "
          f"own pairs vs FP64: max abs {dev.max():.3g}", flush=True)
    assert recall >= 0.99
    if metric == "euclidean":
        assert np.max(dev / np.maximum(pair, 1e-12)) <= 1e-5
        if precision == 0:
            assert index_mismatches_beyond_ties(idx[rows], d64, i64, x[rows], x) == 0
    else:
        assert dev.max() <= 5e-6


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
@pytest.mark.parametrize("precision", ["fp32", "bf16_rerank"])
def test_c3_fit_parity_with_oracle(gpu, metric, precision):
    with open(os.path.join(HERE, "quality.json")) as f:
        q = json.load(f)[f"c3_{metric}_200"]
    x = _x(metric)
    t, k = [], []
    for s in q["seeds"]:
        y = Umap(random_seed=s, ann_mode="exact", n_epochs=200, metric=metric,
                 knn_precision=precision).fit_transform(x)
        assert np.isfinite(y).all()
        t.append(trustworthiness(x, y))
        k.append(knn_preservation(x, y))
        print(f"c3 fit {metric} {precision} seed {s}: trust {t[-1]:.5f} knn-pres {k[-1]:.5f}",
              flush=True)
    assert abs(np.mean(t) - np.mean(q["trust"])) <= 0.01, (t, q["trust"])
    assert abs(np.mean(k) - np.mean(q["knn_preservation"])) <= 0.01, (k, q["knn_preservation"])
