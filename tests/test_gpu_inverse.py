"""Inverse transform on the GPU (inverse_transform_rows, lib.rs:1800-1915)
against the oracle: the kNN + local linear initialisation must equal the
oracle's to f32 rounding (its only difference is the device powf), training
embedding points must return their training rows exactly, and the whole
inverse (with its RNG-fed SGD) must reconstruct held-out rows as well as the
oracle's does.  The facade's out= / empty-input / error behaviour follows
umap_rs/tests/test_binding.py:153-180, 521-566, 1073-1136."""
import numpy as np
import pytest

from umapers_amd import Umap

from . import capi

pytestmark = pytest.mark.gpu


def _mixture(n, d, k, seed):
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, 4.0, size=(k, d))
    lab = rng.integers(0, k, n)
    return (centres[lab] + rng.normal(size=(n, d))).astype(np.float32)


@pytest.mark.parametrize("dim,n,d,k", [(2, 500, 16, 15), (3, 800, 40, 10), (2, 3000, 130, 30)])
def test_inverse_init_equals_oracle(gpu, oracle, dim, n, d, k):
    rng = np.random.default_rng(dim * 1000 + n)
    y = (rng.uniform(-10, 10, size=(n, dim))).astype(np.float32)
    x = rng.normal(size=(n, d)).astype(np.float32)
    q = np.concatenate([rng.uniform(-9, 9, size=(200, dim)).astype(np.float32), y[:20]])
    a, b = oracle.find_ab_params(1.0, 0.1)
    got = capi.inverse_init(x, y, q, k, a, b)
    p = oracle.OracleParams(n_components=dim, n_neighbors=k)
    want = oracle.inverse_transform(x, y, np.ones(n), np.zeros(n), q, p, optimize=False)
    assert np.array_equal(got[200:], x[:20])  # fixed rows
    scale = np.abs(want).max()
    err = np.abs(got - want) / scale
    # the weights differ only in the last bit of powf (device vs host libm),
    # which the regression carries through at f32 rounding
    assert err.max() < 1e-5 and err.mean() < 1e-6, (err.max(), err.mean())


def test_embedding_points_map_back_to_their_rows(gpu):
    x = _mixture(2000, 24, 6, 5)
    m = Umap(n_neighbors=12, n_epochs=120, random_seed=3, ann_mode="exact")
    emb = m.fit_transform(x)
    rec = m.inverse_transform(emb[:64])
    assert rec.shape == (64, 24)
    assert np.array_equal(rec, x[:64])


def test_inverse_reconstruction_parity_with_oracle(gpu, oracle):
    # held-out rows -> plain transform -> inverse: the GPU inverse and the
    # oracle's, given the same model (GPU fit; sigma / rho from the oracle's
    # smooth_knn_dist, equal to the fit's to 1e-5), must reconstruct equally well
    x_all = _mixture(4400, 32, 8, 21)
    x, hold = x_all[:4000], x_all[4000:]
    m = Umap(n_neighbors=15, n_epochs=150, random_seed=9, ann_mode="exact")
    emb = m.fit_transform(x)
    q = m.transform(hold)
    gpu_rec = m.inverse_transform(q)
    idx, dist = oracle.knn(x, 15)
    sig, rho = oracle.smooth_knn_dist(dist, 15.0)
    p = oracle.OracleParams(n_neighbors=15, n_epochs=150, random_seed=9)
    orc_rec = oracle.inverse_transform(x, emb, sig, rho, q, p, optimize=True)
    orc_init = oracle.inverse_transform(x, emb, sig, rho, q, p, optimize=False)
    assert np.all(np.isfinite(gpu_rec))
    rmse = lambda r: float(np.sqrt(np.mean((r - hold) ** 2)))  # noqa: E731
    rg, ro = rmse(gpu_rec), rmse(orc_rec)
    assert abs(rg - ro) <= 0.03 * ro, (rg, ro, rmse(orc_init))
    # the SGD ran: the result is not the initialisation
    assert np.median(np.abs(gpu_rec - orc_init).max(axis=1)) > 1e-3
    # and lands where the oracle's does, far closer than the spread of the data
    gap = np.linalg.norm(gpu_rec - orc_rec, axis=1)
    assert np.median(gap) < 0.25 * np.median(np.linalg.norm(hold - hold.mean(0), axis=1))


def test_inverse_api_out_buffer_empty_and_errors(gpu):
    x = _mixture(120, 10, 3, 21)
    m = Umap(n_neighbors=10, n_components=2, n_epochs=60, init="random", random_seed=17,
             ann_mode="exact")
    m.fit(x)
    t = m.transform(x[:18])
    out = np.empty((18, 10), np.float32)
    rec = m.inverse_transform(t, out=out)
    assert rec is out and np.all(np.isfinite(rec))
    empty = m.inverse_transform(np.empty((0, 2), np.float32))
    assert empty.shape == (0, 10)
    out0 = np.empty((0, 10), np.float32)
    assert m.inverse_transform(np.empty((0, 2), np.float32), out=out0) is out0
    with pytest.raises(ValueError, match="embedding dimension mismatch: expected 2, got 3"):
        m.inverse_transform(np.zeros((4, 3), np.float32))
    bad = np.zeros((4, 2), np.float32)
    bad[1, 0] = np.nan
    with pytest.raises(ValueError, match="non-finite value at row 1"):
        m.inverse_transform(bad)
    with pytest.raises(TypeError):
        m.inverse_transform(t, out=np.empty((18, 10), np.float64))


@pytest.mark.parametrize("d", [784, 1500])
def test_wide_rows_fit_and_inverse(gpu, oracle, d):
    # rows wider than 256 features: the GEMM-scored kNN in the fit, then the
    # inverse SGD with 16 (d <= 1024) or 40 features per lane
    x_all = _mixture(2300, d, 6, d)
    x, hold = x_all[:2000], x_all[2000:]
    m = Umap(n_neighbors=15, n_epochs=100, random_seed=4, ann_mode="exact")
    emb = m.fit_transform(x)
    assert np.array_equal(m.inverse_transform(emb[:16]), x[:16])
    q = m.transform(hold)
    gpu_rec = m.inverse_transform(q)
    idx, dist = oracle.knn(x, 15)
    sig, rho = oracle.smooth_knn_dist(dist, 15.0)
    p = oracle.OracleParams(n_neighbors=15, n_epochs=100, random_seed=4)
    orc_rec = oracle.inverse_transform(x, emb, sig, rho, q, p, optimize=True)
    assert np.all(np.isfinite(gpu_rec))
    rmse = lambda r: float(np.sqrt(np.mean((r - hold) ** 2)))  # noqa: E731
    assert abs(rmse(gpu_rec) - rmse(orc_rec)) <= 0.03 * rmse(orc_rec)
