"""The HIP stages against the golden fixtures (tests/golden/), and the fit's
statistical parity at the densMAP configuration and at k = 30 against the oracle's
recorded trustworthiness."""
import json
import os

import numpy as np
import pytest

from umapers_amd import Umap, datasets
from umapers_amd.diagnostics import knn_preservation, trustworthiness

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


def _two_sided(gpu_vals, oracle_vals, bar=0.01):
    return abs(np.mean(gpu_vals) - np.mean(oracle_vals)) <= bar


def test_densmap_fit_parity_config2(gpu):
    # 20000 x 192, 10 clusters, densMAP on, 500 epochs: mean trustworthiness
    # and kNN preservation against the oracle's means over 32 recorded seeds,
    # both ways.  One seed's trustworthiness moves by ~0.01 between seeds and
    # between Hogwild runs, so both sides are averaged over 32 seeds.
    with open(os.path.join(HERE, "quality.json")) as f:
        q = json.load(f)["c2_densmap_500"]
    x = datasets.c2()
    seeds = list(q["seeds"])
    t, k = [], []
    for s in seeds:
        y = Umap(random_seed=s, ann_mode="exact", n_epochs=500, densmap=True, dens_lambda=2.0,
                 dens_frac=0.3).fit_transform(x)
        t.append(trustworthiness(x, y))
        k.append(knn_preservation(x, y))
    print(f"c2 densMAP: gpu trust {np.mean(t):.4f} knn-pres {np.mean(k):.5f}; oracle "
          f"{np.mean(q['trust']):.4f} / {np.mean(q['knn_preservation']):.5f}", flush=True)
    assert _two_sided(t, q["trust"]), (t, q["trust"])
    assert _two_sided(k, q["knn_preservation"]), (k, q["knn_preservation"])


@pytest.mark.parametrize("precision", ["fp32", "bf16_rerank"])
def test_k30_fit_parity_config4_draw(gpu, precision):
    # 50000 x 64 draw of the config-4 mixture (50 anisotropic clusters), k = 30,
    # 200 epochs: the kNN path with a 32-entry per-query list and a graph of
    # more than 2^20 edges (the edge-parallel epoch kernel).  Mean
    # trustworthiness and kNN preservation over 4 seeds against the oracle's
    # means over its 4 recorded seeds, both ways.
    with open(os.path.join(HERE, "quality.json")) as f:
        q = json.load(f)["c4_sub50k_k30_200"]
    x = datasets.c4(n=50_000)
    t, k = [], []
    for s in q["seeds"]:
        y = Umap(random_seed=s, ann_mode="exact", n_epochs=200, n_neighbors=30,
                 knn_precision=precision).fit_transform(x)
        t.append(trustworthiness(x, y))
        k.append(knn_preservation(x, y))
    assert _two_sided(t, q["trust"]), (t, q["trust"])
    assert _two_sided(k, q["knn_preservation"]), (k, q["knn_preservation"])


def test_hub_graph_fit_parity_and_extent_config5(gpu):
    # 100000 x 192 rows in 20 clusters: the kNN graph has hub rows of about a
    # thousand entries, which the epoch kernel walks one block per row in the
    # reference's order.  Quality within 0.01 of the oracle's both ways; no
    # point thrown out of the body of the embedding (summed simultaneous pulls
    # once threw such hubs to radius ~350 in an embedding whose points all lie
    # within ~17 of its centre); and the extent (largest distance from the
    # median) within 20% of the oracle's.  The epoch reads other rows'
    # positions as they stood at the start of its pass where the reference's
    # single loop sees every earlier move; on this data that leaves the
    # embedding about 15% smaller than the oracle's at equal quality.
    with open(os.path.join(HERE, "quality.json")) as f:
        q = json.load(f)["c5_fit_200"]
    x = datasets.c5(n_fit=100_000, n_query=1)[0]
    t, k, r, r50 = [], [], [], []
    for s in q["seeds"]:
        y = Umap(random_seed=s, ann_mode="exact", n_epochs=200).fit_transform(x)
        assert np.isfinite(y).all()
        rad = np.linalg.norm(y - np.median(y, 0), axis=1)
        assert rad.max() <= 1.2 * np.percentile(rad, 99.9), (s, rad.max(), np.percentile(rad, 99.9))
        t.append(trustworthiness(x, y))
        k.append(knn_preservation(x, y))
        r.append(float(rad.max()))
        r50.append(float(np.percentile(rad, 50)))
    print(f"c5: gpu trust {np.mean(t):.5f} knn-pres {np.mean(k):.5f} extent {np.mean(r):.2f} "
          f"median radius {np.mean(r50):.2f}; oracle {np.mean(q['trust']):.5f} / "
          f"{np.mean(q['knn_preservation']):.5f} / {np.mean(q['radius_max']):.2f} / "
          f"{np.mean(q['radius_p50']):.2f}", flush=True)
    assert _two_sided(t, q["trust"]), (t, q["trust"])
    assert _two_sided(k, q["knn_preservation"]), (k, q["knn_preservation"])
    assert abs(np.mean(r) / np.mean(q["radius_max"]) - 1.0) <= 0.20, (r, q["radius_max"])


def test_plain_fit_parity_config1(gpu):
    # 1000 x 32 i.i.d. N(0, 1), 200 epochs: mean trustworthiness and kNN
    # preservation over 20 seeds against the oracle's means over its 40
    # recorded seeds, both ways (one fit moves by ~0.006 between seeds)
    with open(os.path.join(HERE, "quality.json")) as f:
        q = json.load(f)["c1_plain_200"]
    x = datasets.c1()
    t, k = [], []
    for s in range(20):
        y = Umap(random_seed=s, ann_mode="exact", n_epochs=200).fit_transform(x)
        t.append(trustworthiness(x, y))
        k.append(knn_preservation(x, y))
    print(f"c1: gpu trust {np.mean(t):.4f} knn-pres {np.mean(k):.4f}; oracle "
          f"{np.mean(q['trust']):.4f} / {np.mean(q['knn_preservation']):.4f}", flush=True)
    assert _two_sided(t, q["trust"]), (t, q["trust"])
    assert _two_sided(k, q["knn_preservation"]), (k, q["knn_preservation"])
