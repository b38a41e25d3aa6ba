"""The reference's approximate kNN on the GPU (knn_approx.hip) against the
oracle's restatement (lib.rs:2932-3168): the same RNG stream, candidate
lists and f32 distances, so the tables must be equal entry for entry, for
euclidean and cosine, below and above 64 features, after 0, 2 and 14
rounds.  Then a fit in ann_mode="approximate" against the oracle fitted on
the same approximate graph."""
import numpy as np
import pytest

from umapers_amd import Umap
from umapers_amd.diagnostics import trustworthiness

from . import capi

pytestmark = pytest.mark.gpu


def _x(n, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 3, size=(8, d))
    return (c[rng.integers(0, 8, n)] + rng.normal(size=(n, d))).astype(np.float32)


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
@pytest.mark.parametrize("n,d,k,iters", [(1500, 16, 10, 0), (1500, 16, 10, 2), (2500, 70, 15, 14),
                                         (900, 128, 30, 3)])
def test_table_equals_oracle(gpu, oracle, metric, n, d, k, iters):
    x = _x(n, d, n + d)
    seed = 1234 ^ 0xC0FEFEED1234ABCD
    gi, gd = capi.knn_approx(x, k, metric, 50, iters, seed)
    oi, od = oracle.knn_approx(x, k, metric, 50, iters, seed)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd, od)


def test_pool_and_seed_reach_the_kernel(gpu, oracle):
    x = _x(200, 12, 4)
    gi, gd = capi.knn_approx(x, 8, "euclidean", candidate_pool=199, n_iters=0, seed=9)
    ei, ed = oracle.knn(x, 8)
    assert np.array_equal(gi, ei) and np.array_equal(gd, ed)  # a full pool is exact
    a = capi.knn_approx(x, 8, seed=1, n_iters=1)[0]
    b = capi.knn_approx(x, 8, seed=2, n_iters=1)[0]
    assert not np.array_equal(a, b)


def test_approximate_fit_parity(gpu, oracle):
    x = _x(4000, 32, 11)
    seeds = [3, 4, 5]
    g, o = [], []
    for s in seeds:
        m = Umap(ann_mode="approximate", n_epochs=200, random_seed=s)
        g.append(trustworthiness(x, m.fit_transform(x)))
        idx, dist = oracle.knn_approx(x, 15, seed=s ^ 0xC0FEFEED1234ABCD)
        fit = oracle.fit(x, oracle.OracleParams(n_epochs=200, random_seed=s), knn_idx=idx,
                         knn_dist=dist)
        o.append(trustworthiness(x, fit.embedding))
    assert abs(np.mean(g) - np.mean(o)) <= 0.01, (g, o)
