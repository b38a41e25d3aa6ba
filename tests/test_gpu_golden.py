"""The HIP stages against the golden fixtures (tests/golden/), and the fit's
statistical parity at the densMAP configuration and at k = 30 against the oracle's
recorded trustworthiness."""
import json
import os

import numpy as np
import pytest

from umapers_amd import Umap, datasets
from umapers_amd.diagnostics import trustworthiness

from . import capi

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def c1():
    return dict(np.load(os.path.join(HERE, "c1_stages.npz")))


def test_fp32_knn_equals_fixture(gpu, c1):
    idx, dist = capi.knn(datasets.c1(), 15, 0, 0, 2)
    assert np.array_equal(idx, c1["knn_idx"])
    assert np.array_equal(dist, c1["knn_dist"])


def test_smooth_knn_equals_fixture(gpu, c1):
    s, r = capi.smooth(c1["knn_dist"], 15.0, 1.0, True, fp64=False)
    assert np.array_equal(r, c1["rho"])
    rel = np.abs(s - c1["sigma"]) / c1["sigma"]
    assert np.mean(rel == 0) > 0.95 and rel.max() < 1e-5


def test_graph_and_schedule_equal_fixture(gpu, c1):
    h, t, w = capi.fuzzy_graph(c1["knn_idx"], c1["knn_dist"], c1["sigma"], c1["rho"], 1.0, 200)
    assert np.array_equal(h, c1["heads"]) and np.array_equal(t, c1["tails"])
    assert np.allclose(w, c1["weights"], rtol=1e-5, atol=0)
    y0 = np.zeros((1000, 2), np.float32)
    _, pos, neg = capi.optimize(y0, c1["heads"], c1["tails"], c1["weights"], 200, 1.58, 0.9,
                                counts=True)
    assert np.array_equal(pos, c1["pos_counts"]) and np.array_equal(neg, c1["neg_counts"])


def test_densmap_fit_parity_config2(gpu):
    # 20000 x 192, 10 clusters, densMAP on, 500 epochs: mean trustworthiness
    # against the oracle's mean over 16 recorded seeds.  One seed's value moves
    # by ~0.01 between seeds and between Hogwild runs, so the GPU side is
    # averaged over 24 seeds (spread of the mean ~0.002; measured GPU means
    # 0.844-0.847 against the oracle's 0.850).
    with open(os.path.join(HERE, "quality.json")) as f:
        q = json.load(f)["c2_densmap_500"]
    x = datasets.c2()
    seeds = list(q["seeds"]) + [59, 61, 67, 71, 73, 79, 83, 89]
    t = [trustworthiness(x, Umap(random_seed=s, ann_mode="exact", n_epochs=500, densmap=True,
                                 dens_lambda=2.0, dens_frac=0.3).fit_transform(x))
         for s in seeds]
    assert np.mean(t) >= np.mean(q["trust"]) - 0.01, (t, q["trust"])


@pytest.mark.parametrize("precision", ["fp32", "bf16_rerank"])
def test_k30_fit_parity_config4_draw(gpu, precision):
    # 50000 x 64 draw of the config-4 mixture (50 anisotropic clusters), k = 30,
    # 200 epochs: the kNN path with a 32-entry per-query list.  Mean
    # trustworthiness over 4 seeds against the oracle's mean over its 4
    # recorded seeds (one seed's value moves by ~0.001 here).
    with open(os.path.join(HERE, "quality.json")) as f:
        q = json.load(f)["c4_sub50k_k30_200"]
    x = datasets.c4(n=50_000)
    t = [trustworthiness(x, Umap(random_seed=s, ann_mode="exact", n_epochs=200, n_neighbors=30,
                                 knn_precision=precision).fit_transform(x))
         for s in q["seeds"]]
    assert np.mean(t) >= np.mean(q["trust"]) - 0.01, (t, q["trust"])
