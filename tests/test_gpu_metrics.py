"""Exact kNN under the reference's other metrics on the GPU
(knn_metric.hip) against the oracle (lib.rs:2494-2698, 2891-2928,
3346-3399): indices and distances bit for bit, except minkowski, whose
per-feature powf differs from the host's in the last bit (distances within
1e-5 relative, index changes only between near-equal distances).  Then a
whole fit under manhattan and correlation against the oracle's
trustworthiness."""
import numpy as np
import pytest

from umapers_amd import Umap
from umapers_amd.diagnostics import trustworthiness

from . import capi

pytestmark = pytest.mark.gpu

EXACT = ["manhattan", "chebyshev", "correlation", "canberra", "braycurtis"]


def _data(n, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 3, size=(6, d))
    return (c[rng.integers(0, 6, n)] + rng.normal(size=(n, d))).astype(np.float32)


@pytest.mark.parametrize("metric", EXACT)
@pytest.mark.parametrize("n,d,k", [(700, 40, 10), (1000, 70, 15), (613, 128, 20), (300, 7, 5),
                                   (400, 300, 10), (330, 778, 15)])
def test_self_knn_bit_exact(gpu, oracle, metric, n, d, k):
    x = _data(n, d, n + d)
    gi, gd = capi.knn_metric(x, x, k, metric)
    oi, od = oracle.knn_metric(x, x, k, metric)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd, od)


@pytest.mark.parametrize("metric", EXACT + ["minkowski"])
def test_to_reference_knn(gpu, oracle, metric):
    r = _data(900, 64, 3)
    q = _data(130, 64, 4)
    gi, gd = capi.knn_metric(q, r, 15, metric, 3.0)
    oi, od = oracle.knn_metric(q, r, 15, metric, 3.0)
    if metric != "minkowski":
        assert np.array_equal(gi, oi) and np.array_equal(gd, od)
        return
    assert np.allclose(gd, od, rtol=1e-5, atol=0)
    diff = gi != oi
    # an index may differ only where its two candidates' distances are near-equal
    for i, c in zip(*np.nonzero(diff)):
        assert abs(float(gd[i, c]) - float(od[i, c])) <= 1e-5 * float(od[i, c])
    assert diff.mean() < 0.01


@pytest.mark.parametrize("metric", ["manhattan", "correlation"])
def test_fit_parity_under_metric(gpu, oracle, metric):
    x = _data(3000, 32, 17)
    seeds = [1, 2, 3]
    g = [trustworthiness(x, Umap(metric=metric, n_epochs=200, random_seed=s,
                                 ann_mode="exact").fit_transform(x)) for s in seeds]
    o = [trustworthiness(x, oracle.fit(x, oracle.OracleParams(metric=metric, n_epochs=200,
                                                                 random_seed=s)).embedding)
         for s in seeds]
    assert abs(np.mean(g) - np.mean(o)) <= 0.01, (g, o)


def test_minkowski_p_reaches_the_kernel(gpu):
    x = _data(400, 16, 9)
    i2, d2 = capi.knn_metric(x, x, 8, "minkowski", 2.0)
    ie, de = capi.knn_metric(x, x, 8, "euclidean")
    assert np.allclose(d2, de, rtol=1e-5, atol=1e-5)
    _, d1 = capi.knn_metric(x, x, 8, "minkowski", 1.0)
    _, dm = capi.knn_metric(x, x, 8, "manhattan")
    assert np.allclose(d1, dm, rtol=1e-5, atol=1e-5)


def test_too_wide_for_the_metric_kernel_is_an_error(gpu):
    x = np.zeros((40, 2561), np.float32)
    with pytest.raises(Exception, match="at most 2560 features"):
        capi.knn_metric(x, x, 5, "manhattan")
