"""smooth_knn_dist, membership, fuzzy union, prune and radii on the GPU vs the
oracle."""
import numpy as np
import pytest

from umapers_amd import datasets

from . import capi

pytestmark = pytest.mark.gpu


def _knn(oracle, x, k=15):
    return oracle.knn(x, k)


@pytest.fixture(scope="module")
def mixture():
    return datasets.gaussian_mixture(4000, 64, 6, 3.0, 21, (0.5, 1.5))


def test_include_self_predicate(gpu, oracle, mixture):
    idx, dist = _knn(oracle, mixture)
    assert capi.include_self(idx, dist) == oracle.rows_include_self(idx, dist) is True
    idx2 = idx.copy()
    idx2[5, 0], idx2[5, 1] = idx[5, 1], idx[5, 0]
    assert capi.include_self(idx2, dist) is False


def test_smooth_knn_f32_mode_matches_oracle(gpu, oracle, mixture):
    idx, dist = _knn(oracle, mixture)
    s_o, r_o = oracle.smooth_knn_dist(dist, 15.0, 1.0, include_self=True)
    s_g, r_g = capi.smooth(dist, 15.0, 1.0, True, fp64=False)
    assert np.array_equal(r_g, r_o)
    # exp differs by at most an ulp between libm and the GPU
    rel = np.abs(s_g - s_o) / s_o
    assert np.mean(rel == 0) > 0.95 and rel.max() < 1e-5


@pytest.mark.parametrize("which", ["mixture", "c2"])
def test_smooth_knn_fp64_default_path_matches_oracle(gpu, oracle, mixture, which):
    """The default (f64-psum) bisection against the oracle's f64-psum form:
    rho equal, sigma within 1e-6 relative, identical edge lists and weights
    within 1e-5 relative after membership, union and prune."""
    x = mixture if which == "mixture" else datasets.c2()
    idx, dist = _knn(oracle, x)
    s_o, r_o = oracle.smooth_knn_dist(dist, 15.0, 1.0, include_self=True, fp64_sums=True)
    s_g, r_g = capi.smooth(dist, 15.0, 1.0, True, fp64=True)
    assert np.array_equal(r_g, r_o)
    rel = np.abs(s_g - s_o) / s_o
    print(f"{which}: sigma max rel {rel.max():.3g}, identical {np.mean(rel == 0):.4f}", flush=True)
    assert rel.max() <= 1e-6
    h_o, t_o, w_o = oracle.fuzzy_graph(idx, dist, s_o, r_o, 1.0, 500)
    h_g, t_g, w_g = capi.fuzzy_graph(idx, dist, s_g, r_g, 1.0, 500)
    assert np.array_equal(h_g, h_o) and np.array_equal(t_g, t_o)
    assert np.max(np.abs(w_g - w_o) / w_o) <= 1e-5


def test_smooth_knn_local_connectivity_variants(gpu, oracle, rng):
    d = np.sort(rng.uniform(0.0, 2.0, size=(300, 10)).astype(np.float32), axis=1)
    d[::7, :4] = 0.0  # rows with several zero distances
    for lc in (0.0, 0.5, 1.0, 1.5, 3.0):
        s_o, r_o = oracle.smooth_knn_dist(d, 10.0, lc, include_self=False)
        s_g, r_g = capi.smooth(d, 10.0, lc, False, fp64=False)
        assert np.array_equal(r_g, r_o), lc
        assert np.allclose(s_g, s_o, rtol=1e-5, atol=0), lc


def test_fuzzy_graph_matches_oracle(gpu, oracle, mixture):
    idx, dist = _knn(oracle, mixture)
    sig, rho = oracle.smooth_knn_dist(dist, 15.0, 1.0, include_self=True)
    for prune in (0, 200, 5):
        h_o, t_o, w_o = oracle.fuzzy_graph(idx, dist, sig, rho, 1.0, prune)
        h_g, t_g, w_g = capi.fuzzy_graph(idx, dist, sig, rho, 1.0, prune)
        # identical edge list in identical (head, tail) order
        assert np.array_equal(h_g, h_o) and np.array_equal(t_g, t_o)
        assert np.all(np.abs(w_g - w_o) <= 1e-5 * w_o)
        # symmetric weights, head-sorted
        key = dict(zip(zip(h_g.tolist(), t_g.tolist()), w_g.tolist()))
        assert all(key[(b, a)] == v for (a, b), v in key.items())


@pytest.mark.parametrize("mix", [0.0, 0.3, 1.0])
def test_fuzzy_union_mix_ratio(gpu, oracle, mixture, mix):
    idx, dist = _knn(oracle, mixture[:1500])
    sig, rho = oracle.smooth_knn_dist(dist, 15.0, 1.0, include_self=True)
    h_o, t_o, w_o = oracle.fuzzy_graph(idx, dist, sig, rho, mix, 0)
    h_g, t_g, w_g = capi.fuzzy_graph(idx, dist, sig, rho, mix, 0)
    assert np.array_equal(h_g, h_o) and np.array_equal(t_g, t_o)
    assert np.all(np.abs(w_g - w_o) <= 1e-5 * w_o + 1e-7)


def test_radii_match_oracle(gpu, oracle, mixture):
    idx, dist = _knn(oracle, mixture)
    assert np.allclose(capi.local_radii(idx, dist), oracle.local_radii(idx, dist), rtol=1e-6, atol=1e-6)
    y = np.random.default_rng(2).normal(size=(mixture.shape[0], 2)).astype(np.float32)
    assert np.allclose(capi.embedding_radii(y, idx), oracle.embedding_radii(y, idx), rtol=1e-6,
                       atol=1e-6)


def test_components_match_oracle(gpu, oracle, mixture):
    idx, dist = _knn(oracle, mixture)
    sig, rho = oracle.smooth_knn_dist(dist, 15.0)
    h, t, w = oracle.fuzzy_graph(idx, dist, sig, rho, 1.0, 200)
    nc_o, lab_o = oracle.connected_components(mixture.shape[0], h, t)
    nc_g, lab_g = capi.components(mixture.shape[0], h, t, w)
    assert nc_g == nc_o and np.array_equal(lab_g, lab_o)


def test_fuzzy_union_hub_rows(gpu, oracle):
    # Plant two hubs in a real kNN graph: every row also points at row 0
    # (in-degree ~20000, the serial path) and 2500 rows point at row 1
    # (in-degree ~2500, the block path of up to 16384 entries); the remaining
    # rows with more than 64 union entries take the block path of up to 1024.
    x = datasets.gaussian_mixture(20000, 16, 8, 3.0, 41, 1.0)
    idx, dist = oracle.knn(x, 15)
    idx = idx.copy()
    for r in range(2, idx.shape[0]):
        if 0 not in idx[r]:
            idx[r, 14] = 0
        if r < 2502 and 1 not in idx[r]:
            idx[r, 13] = 1
    sig, rho = oracle.smooth_knn_dist(dist, 15.0, 1.0, include_self=True)
    h_o, t_o, w_o = oracle.fuzzy_graph(idx, dist, sig, rho, 1.0, 0)
    h_g, t_g, w_g = capi.fuzzy_graph(idx, dist, sig, rho, 1.0, 0)
    indeg = np.bincount(idx[:, 1:].ravel(), minlength=idx.shape[0])
    assert indeg[0] > 16384 and 1024 < indeg[1] <= 16384
    assert np.array_equal(h_g, h_o) and np.array_equal(t_g, t_o)
    assert np.all(np.abs(w_g - w_o) <= 1e-5 * w_o)
