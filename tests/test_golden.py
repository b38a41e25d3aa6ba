"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle).

Each fixture is pinned here independently of the code that wrote it: the kNN
against an FP64 brute force, sigma/rho against the log2(k) target, the graph
against the membership/union closed forms recomputed in numpy from the
fixture's own kNN and sigma/rho, the schedule counts against the firing
closed form, and the oracle must reproduce every array bit for bit.
"""
import json
import os

import numpy as np
import pytest

from umapers_amd import datasets

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def c1():
    return dict(np.load(os.path.join(HERE, "c1_stages.npz")))


def test_inputs_are_the_recorded_bytes():
    with open(os.path.join(HERE, "inputs.json")) as f:
        want = json.load(f)
    assert datasets.sha256(datasets.c1()) == want["c1"]
    assert datasets.sha256(datasets.c2()) == want["c2"]
    assert datasets.sha256(datasets.c3()) == want["c3"]
    assert datasets.sha256(datasets.c5(n_fit=100_000, n_query=1)[0]) == want["c5_fit"]


def test_knn_fixture_against_fp64(c1):
    x = datasets.c1().astype(np.float64)
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    ref = np.sqrt(d2)
    idx, dist = c1["knn_idx"], c1["knn_dist"]
    assert idx.shape == (1000, 15) and np.all(idx[:, 0] == np.arange(1000))
    got = np.take_along_axis(ref, idx.astype(np.int64), 1)
    assert np.allclose(dist, got, rtol=1e-6, atol=1e-6)
    # the k nearest by FP64, up to ties within 1e-6 relative at the k-th
    kth = np.sort(ref, axis=1)[:, 14]
    for i in range(1000):
        want = set(np.flatnonzero(ref[i] < kth[i] * (1 - 1e-6)).tolist())
        assert want <= set(idx[i].tolist())
        assert np.all(ref[i, idx[i]] <= kth[i] * (1 + 1e-6))


def test_sigma_rho_fixture(c1):
    dist, sig, rho = c1["knn_dist"], c1["sigma"], c1["rho"]
    assert np.array_equal(rho, dist[:, 1])  # local_connectivity 1: nearest non-self
    act = dist[:, 1:].astype(np.float64)
    psum = np.where(act - rho[:, None] > 0, np.exp(-(act - rho[:, None]) / sig[:, None]), 1.0).sum(1)
    assert np.all(np.abs(psum - np.log2(15.0)) < 2e-5)


def test_graph_fixture_closed_form(c1):
    idx, dist, sig, rho = c1["knn_idx"], c1["knn_dist"], c1["sigma"], c1["rho"]
    h, t, w = c1["heads"], c1["tails"], c1["weights"]
    m = {}
    for i in range(idx.shape[0]):
        for c in range(1, idx.shape[1]):
            d = float(dist[i, c]) - float(rho[i])
            m[(i, int(idx[i, c]))] = 1.0 if d <= 0 else np.exp(-d / float(sig[i]))
    union = {}
    for (i, j), a in m.items():
        b = m.get((j, i), 0.0)
        union[(i, j)] = a + b - a * b
        union[(j, i)] = union[(i, j)]
    wmax = max(union.values())
    kept = {e: v for e, v in union.items() if v >= wmax / 200.0}
    got = {(int(a), int(b)): float(c) for a, b, c in zip(h, t, w)}
    assert set(got) == set(kept)
    for e, v in got.items():
        assert abs(v - kept[e]) <= 1e-5 * kept[e]
        assert got[(e[1], e[0])] == v
    assert np.all(np.diff(h) >= 0)


def test_schedule_fixture_closed_form(c1):
    w, pos, neg = c1["weights"], c1["pos_counts"], c1["neg_counts"]
    eps = w.max() / w.astype(np.float64)
    # an edge first fires at epoch ceil(eps) and then every eps epochs
    want = np.floor(199.0 / eps)
    assert np.all(np.abs(pos - want) <= 1)
    # negatives accrue at eps / 5 per epoch from eps / 5 and are drawn when the
    # edge fires, which is up to one epoch after its nominal time: 5 per firing
    # less the partial first interval, plus at most 5 / eps for that last epoch
    assert np.all(neg >= 5 * pos - 5) and np.all(neg <= 5 * pos + np.ceil(5.0 / eps))


def test_oracle_reproduces_fixtures(oracle, c1):
    x = datasets.c1()
    idx, dist = oracle.knn(x, 15)
    assert np.array_equal(idx, c1["knn_idx"]) and np.array_equal(dist, c1["knn_dist"])
    sig, rho = oracle.smooth_knn_dist(dist, 15.0)
    assert np.array_equal(sig, c1["sigma"]) and np.array_equal(rho, c1["rho"])
    h, t, w = oracle.fuzzy_graph(idx, dist, sig, rho, 1.0, 200)
    assert np.array_equal(h, c1["heads"]) and np.array_equal(t, c1["tails"])
    assert np.array_equal(w, c1["weights"])
    pos, neg, _, _ = oracle.schedule_counts(w, 200, 5)
    assert np.array_equal(pos, c1["pos_counts"]) and np.array_equal(neg, c1["neg_counts"])


def test_quality_targets_recorded():
    with open(os.path.join(HERE, "quality.json")) as f:
        q = json.load(f)
    assert len(q["c1_plain_200"]["trust"]) == 40
    assert len(q["c2_densmap_500"]["trust"]) == 32
    assert 0.6 < np.mean(q["c1_plain_200"]["trust"]) < 0.7
    assert 0.8 < np.mean(q["c2_densmap_500"]["trust"]) < 0.9
    assert len(q["c4_sub50k_k30_200"]["trust"]) == 4
    assert 0.98 < np.mean(q["c4_sub50k_k30_200"]["trust"]) < 1.0
