"""The oracle's inverse transform (lib.rs:1800-1915) pinned without the
reference: on data that is an exact affine image of its embedding, the local
linear initialisation (init_inverse_points, lib.rs:4838-4935) must return the
affine image of the query (the 1e-3 ridge term biases it by about 1e-3 / the
local spread squared); queries on a training embedding point return that
training row bit for bit; and the SGD keeps every row finite and moves only
the non-fixed rows."""
import numpy as np
import pytest

from oracle import oracle as O


def _affine_case(dim, n=400, d=12, seed=3):
    rng = np.random.default_rng(seed)
    scale = np.array([50.0, 20.0, 35.0][:dim])
    z = (rng.uniform(-1.0, 1.0, size=(n, dim)) * scale).astype(np.float32)
    a = rng.normal(size=(d, dim)).astype(np.float32)
    c = rng.normal(size=d).astype(np.float32)
    x = (z.astype(np.float64) @ a.T + c).astype(np.float32)
    return z, x, a, c, scale


@pytest.mark.parametrize("dim", [2, 3])
def test_init_reconstructs_affine_data(dim):
    z, x, a, c, scale = _affine_case(dim)
    rng = np.random.default_rng(11)
    q = (rng.uniform(-0.8, 0.8, size=(64, dim)) * scale).astype(np.float32)
    n = x.shape[0]
    p = O.OracleParams(n_components=dim, n_neighbors=15)
    got = O.inverse_transform(x, z, np.ones(n), np.zeros(n), q, p, optimize=False)
    want = q.astype(np.float64) @ a.T + c
    err = np.abs(got - want) / np.abs(want).max()
    # measured ~1e-5 (f32 rounding and the ridge); the weighted neighbour
    # mean alone, i.e. a lost or mis-solved regression term, is off by ~0.1
    assert err.max() < 1e-4, err.max()


def test_fixed_rows_return_the_training_row():
    z, x, _, _, _ = _affine_case(2, seed=5)
    n = x.shape[0]
    p = O.OracleParams(n_components=2, n_neighbors=10, random_seed=4)
    q = z[::7].copy()
    for optimize in (False, True):
        got = O.inverse_transform(x, z, np.full(n, 0.5), np.zeros(n), q, p, optimize=optimize)
        assert np.array_equal(got, x[::7])


def test_sgd_moves_only_free_rows_and_stays_finite():
    rng = np.random.default_rng(8)
    centres = rng.normal(0, 4, size=(6, 10))
    lab = rng.integers(0, 6, 600)
    x = (centres[lab] + rng.normal(size=(600, 10))).astype(np.float32)
    fit = O.fit(x, O.OracleParams(n_epochs=60, random_seed=2, n_neighbors=10))
    q = np.concatenate([fit.embedding[:5], fit.embedding[100:140] + 0.05]).astype(np.float32)
    p = O.OracleParams(n_neighbors=10, random_seed=2)
    init = O.inverse_transform(x, fit.embedding, fit.sigmas, fit.rhos, q, p, optimize=False)
    full = O.inverse_transform(x, fit.embedding, fit.sigmas, fit.rhos, q, p, optimize=True)
    assert np.all(np.isfinite(full))
    assert np.array_equal(full[:5], x[:5])
    moved = np.abs(full[5:] - init[5:]).max(axis=1)
    assert np.all(moved > 0)
