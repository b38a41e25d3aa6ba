"""The FP64 encoder oracle against central finite differences (CPU only)."""
import numpy as np

from oracle import mlp_oracle as M


def _problem(d, h, dim, n_edges, n_neg, seed):
    rng = np.random.default_rng(seed)
    L = M.layout(d, h, dim)
    p = np.zeros(L["total"])
    for name, fan in (("w1", d), ("w2", h), ("w3", h)):
        o, size = L[name]
        p[o:o + size] = rng.uniform(-1, 1, size) * np.sqrt(6.0 / fan)
    if L["dp"] > d:  # padded input rows of W1 are structurally zero
        w1 = p[:L["w1"][1]].reshape(L["dp"], h)
        w1[d:] = 0.0
    for name in ("b1", "b2", "b3"):
        o, size = L[name]
        p[o:o + size] = rng.normal(0, 0.1, size)
    rows = n_edges * (2 + n_neg)
    x = rng.normal(size=(rows, d))
    y = rng.normal(0, 3, size=(rows, dim))
    mean, inv_std = x.mean(0), 1.0 / x.std(0)
    return p, x, y, mean, inv_std


def test_oracle_gradient_matches_finite_differences():
    d, h, dim, B, nn = 13, 64, 2, 6, 3
    a, b, lam = 1.577, 0.895, 0.7
    p, x, y, mean, inv = _problem(d, h, dim, B, nn, 3)
    _, g, _ = M.loss_and_grads(p, x, y, mean, inv, d, h, dim, B, nn, a, b, lam)
    rng = np.random.default_rng(5)
    L = M.layout(d, h, dim)
    picks = []
    for name in ("w1", "b1", "w2", "b2", "w3", "b3"):
        o, size = L[name]
        picks += list(o + rng.choice(size, min(size, 6), replace=False))
    for i in picks:
        if abs(g[i]) < 1e-9:  # padded or dead-unit entries
            continue
        e = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        lp = M.loss_and_grads(pp, x, y, mean, inv, d, h, dim, B, nn, a, b, lam)[0]
        lm = M.loss_and_grads(pm, x, y, mean, inv, d, h, dim, B, nn, a, b, lam)[0]
        fd = (lp - lm) / (2 * e)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_oracle_loss_closed_form_single_edge():
    # one edge, no negatives, identity-like network: the loss is the closed
    # form log(1 + a s^b) + lambda/2 (|z_i - y_i|^2 + |z_j - y_j|^2)
    a, b, lam = 1.2, 0.8, 0.5
    z = np.array([[0.0, 0.0], [3.0, 4.0]])
    y = np.array([[1.0, 0.0], [3.0, 3.0]])
    loss, dz = M.loss_and_dz(z, y, 1, 0, a, b, lam)
    s = 25.0
    assert np.isclose(loss, np.log1p(a * s ** b) + lam / 2 * (1.0 + 1.0))
    g = 2 * a * b * s ** (b - 1) / (1 + a * s ** b)
    assert np.allclose(dz[0], g * (z[0] - z[1]) + lam * (z[0] - y[0]))
