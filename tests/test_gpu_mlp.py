"""The learned transform: HIP gradients against the FP64 oracle, training,
and what the trained encoder gains over the plain transform."""
import time

import numpy as np
import pytest

from oracle import mlp_oracle as M
from umapers_amd import Umap, datasets
from umapers_amd.diagnostics import trustworthiness

from . import capi

pytestmark = pytest.mark.gpu


def _problem(d, h, dim, n_edges, n_neg, seed):
    rng = np.random.default_rng(seed)
    L = M.layout(d, h, dim)
    p = np.zeros(L["total"], np.float32)
    for name, fan in (("w1", d), ("w2", h), ("w3", h)):
        o, size = L[name]
        p[o:o + size] = rng.uniform(-1, 1, size) * np.sqrt(6.0 / fan)
    w1 = p[:L["w1"][1]].reshape(L["dp"], h)
    w1[d:] = 0.0
    for name in ("b1", "b2", "b3"):
        o, size = L[name]
        p[o:o + size] = rng.normal(0, 0.05, size)
    rows = n_edges * (2 + n_neg)
    x = rng.normal(0, 2, size=(rows, d)).astype(np.float32)
    y = rng.normal(0, 3, size=(rows, dim)).astype(np.float32)
    mean = x.mean(0).astype(np.float32)
    inv = (1.0 / x.std(0)).astype(np.float32)
    return p, x, y, mean, inv


@pytest.mark.parametrize("d,h,dim", [(192, 256, 2), (37, 128, 3)])
def test_batch_gradients_match_fp64_oracle(gpu, d, h, dim):
    n_edges, n_neg, a, b, lam = 64, 5, 1.577, 0.895, 1.0
    p, x, y, mean, inv = _problem(d, h, dim, n_edges, n_neg, 11)
    assert capi.mlp_param_count(d, h, dim) == p.size
    loss, g, z = capi.mlp_batch_grad(p, mean, inv, x, y, d, h, dim, n_edges, n_neg, a, b, lam)
    lo, go, zo = M.loss_and_grads(p, x, y, mean, inv, d, h, dim, n_edges, n_neg, a, b, lam,
                                  emulate_bf16=True)
    assert np.allclose(z, zo, rtol=2e-3, atol=2e-3 * np.abs(zo).max())
    assert abs(loss - lo) <= 1e-3 * abs(lo)
    L = M.layout(d, h, dim)
    for name in ("w1", "b1", "w2", "b2", "w3", "b3"):
        o, size = L[name]
        ref = go[o:o + size]
        err = np.linalg.norm(g[o:o + size] - ref) / max(np.linalg.norm(ref), 1e-30)
        assert err <= 1e-2, (name, err)
    # padded input rows of W1 get no gradient
    assert np.all(g[:L["w1"][1]].reshape(L["dp"], h)[d:] == 0)


@pytest.mark.parametrize("d,h,dim,rows", [(192, 256, 2, 70001), (192, 256, 2, 1), (100, 256, 1, 33),
                                           (64, 128, 3, 999), (36, 64, 8, 70), (1024, 192, 5, 257)])
def test_fused_encoder_matches_chain_and_oracle(gpu, d, h, dim, rows):
    """The fused inference kernel against the unfused chain (the training
    forward) and the FP64 oracle of the BF16 network, on ragged row counts
    (one row, partial 32-row tiles, more tiles than one pass of the grid) and
    input widths that end inside a lane's 8-input step."""
    rng = np.random.default_rng(d + h + dim + rows)
    L = M.layout(d, h, dim)
    p = np.zeros(L["total"], np.float32)
    for name, fan in (("w1", d), ("w2", h), ("w3", h)):
        o, size = L[name]
        p[o:o + size] = rng.uniform(-1, 1, size) * np.sqrt(6.0 / fan)
    p[:L["w1"][1]].reshape(L["dp"], h)[d:] = 0.0
    for name in ("b1", "b2", "b3"):
        o, size = L[name]
        p[o:o + size] = rng.normal(0, 0.05, size)
    x = rng.normal(1.5, 2, size=(rows, d)).astype(np.float32)
    mean = (x.mean(0) + 0.1).astype(np.float32)
    inv = (1.0 / (x.std(0) + 0.5)).astype(np.float32)
    z_f = capi.mlp_forward(p, mean, inv, x, d, h, dim, fused=True)
    z_c = capi.mlp_forward(p, mean, inv, x, d, h, dim, fused=False)
    zo, _ = M.forward(p, x, mean, inv, d, h, dim, emulate_bf16=True)
    tol = 2e-3 * np.abs(zo).max()
    assert np.isfinite(z_f).all()
    # FP32 accumulation in a different order: a hidden value may round to the
    # neighbouring BF16 value, so elementwise closeness, not equality
    assert np.abs(z_f - zo).max() <= tol, np.abs(z_f - zo).max() / np.abs(zo).max()
    assert np.abs(z_f - z_c).max() <= tol
    assert np.median(np.abs(z_f - zo)) <= 1e-4 * np.abs(zo).max()


def test_fused_encoder_limits(gpu):
    """Widths the fused kernel does not take are refused by the test entry
    point (the transform itself runs the unfused chain for them)."""
    d, h, dim = 37, 64, 2
    L = M.layout(d, h, dim)
    p = np.zeros(L["total"], np.float32)
    x = np.ones((4, d), np.float32)
    with pytest.raises(ValueError, match="fused encoder"):
        capi.mlp_forward(p, np.zeros(d, np.float32), np.ones(d, np.float32), x, d, h, dim, fused=True)
    z = capi.mlp_forward(p, np.zeros(d, np.float32), np.ones(d, np.float32), x, d, h, dim, fused=False)
    assert z.shape == (4, dim)


def test_learned_transform_trains_and_gains_over_plain(gpu):
    x_fit, x_q = datasets.c5(n_fit=20000, n_query=100000)
    judge = x_q[:2000]
    model = Umap(n_epochs=200, transform_mode="learned", mlp_steps=600, random_seed=3,
                 ann_mode="exact")
    y_fit = model.fit_transform(x_fit)
    first, last, steps = model.learned_transform_loss()
    assert steps == 600 and np.isfinite(last) and last < 0.5 * first, (first, last)

    y_plain = model.transform_plain(judge)          # the judge: kNN + SGD
    y_mlp = model.transform_learned(judge)
    assert np.isfinite(y_mlp).all()
    extent = np.ptp(y_fit, axis=0).max()
    per_row = np.linalg.norm(y_mlp - y_plain, axis=1)
    t_plain, t_mlp = trustworthiness(judge, y_plain), trustworthiness(judge, y_mlp)
    assert np.median(per_row) <= 0.05 * extent, (np.median(per_row), extent)
    assert t_mlp >= t_plain - 0.02, (t_mlp, t_plain)

    # the same encoder after a single step: training is what buys the match
    raw = Umap(n_epochs=200, transform_mode="learned", mlp_steps=1, random_seed=3,
               ann_mode="exact")
    raw.fit(x_fit)
    y_raw = raw.transform_learned(judge)
    assert np.median(np.linalg.norm(y_raw - y_plain, axis=1)) > 4 * np.median(per_row)

    # one pass through the encoder against the per-query kNN + SGD
    model.transform_learned(x_q)
    t0 = time.perf_counter()
    model.transform_learned(x_q)
    mlp_rps = x_q.shape[0] / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    model.transform_plain(judge)
    plain_rps = judge.shape[0] / (time.perf_counter() - t0)
    assert mlp_rps > 10 * plain_rps, (mlp_rps, plain_rps)
