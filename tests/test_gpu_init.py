"""Spectral initialisation on the GPU against the oracle's (lib.rs:3855-4764).

Eigenvectors are fixed only up to sign (and rotation inside a degenerate
eigenspace), so coordinates are compared as subspaces: the smallest cosine of
the principal angles between the two column spaces must exceed 0.99.
"""
import numpy as np
import pytest

from . import capi

pytestmark = pytest.mark.gpu


def _graph(oracle, x, k=15):
    idx, dist = oracle.knn(x, k)
    sig, rho = oracle.smooth_knn_dist(dist, float(k))
    return oracle.fuzzy_graph(idx, dist, sig, rho, 1.0, 200)


def _stretched_blob(rng, n, d, centre=0.0):
    # a Gaussian stretched along its first two axes: the graph's two smoothest
    # non-trivial eigenvectors follow those axes with a clear gap to the next,
    # so the plane they span is well defined (an isotropic blob has d nearly
    # equal eigenvalues there)
    scale = np.ones(d)
    scale[:2] = (4.0, 2.5)
    return (centre + rng.normal(size=(n, d)) * scale).astype(np.float32)


def _subspace_cos(a, b):
    qa, _ = np.linalg.qr(np.asarray(a, np.float64))
    qb, _ = np.linalg.qr(np.asarray(b, np.float64))
    return np.linalg.svd(qa.T @ qb, compute_uv=False).min()


@pytest.mark.parametrize("hubs", [False, True])
def test_spectral_init_connected_graph(gpu, oracle, hubs):
    x = _stretched_blob(np.random.default_rng(21), 4000, 32)
    h, t, w = _graph(oracle, x)
    if hubs:
        # rows of 130 and 700 entries: the product splits them into 64-entry
        # segments whose partial sums are added in a second pass
        rng = np.random.default_rng(4)
        have = set(zip(h.tolist(), t.tolist()))
        extra = []
        for v, m in ((5, 130), (17, 700)):
            nb = [u for u in rng.choice(np.arange(20, 4000), size=m, replace=False)
                  if (v, u) not in have and (u, v) not in have]
            extra += [(v, u) for u in nb] + [(u, v) for u in nb]
        eh, et = np.array(extra).T
        h = np.concatenate([h, eh])
        t = np.concatenate([t, et])
        w = np.concatenate([w, np.full(eh.shape[0], 0.05, np.float32)])
        order = np.lexsort((t, h))
        h, t, w = h[order], t[order], w[order]
        assert np.bincount(h).max() >= 690
    nc, _ = oracle.connected_components(x.shape[0], h, t)
    assert nc == 1
    yo, ok_o = oracle.initialize_embedding(x, 2, h, t, w, "spectral", 42)
    yg, ok_g = capi.init_embedding(x, 2, h, t, w, init=1, seed=42)
    assert ok_o and ok_g
    assert np.isfinite(yg).all()
    # noisy_scale_coords: max |coord| = 10 before N(0, 1e-4) noise
    assert abs(np.abs(yg).max() - 10.0) < 1e-3
    assert _subspace_cos(yg, yo) > 0.99


def test_spectral_init_components(gpu, oracle):
    # one kNN graph per blob, so each blob is exactly one component: the three
    # largest take the GPU subspace iteration, the next two the dense host
    # solve and the 3-row blob the random layout.  Centroids about 1.7 apart
    # keep the meta layout's affinities exp(-d^2) away from zero.
    rng = np.random.default_rng(5)
    sizes = (900, 600, 300, 100, 60, 3)
    centres = rng.normal(0.0, 0.3, size=(len(sizes), 16))
    xs, hs, ts, ws = [], [], [], []
    off = 0
    for c, s in zip(centres, sizes):
        xb = _stretched_blob(rng, s, 16, c)
        h, t, w = _graph(oracle, xb, k=min(15, s - 1))
        xs.append(xb)
        hs.append(h + off)
        ts.append(t + off)
        ws.append(w)
        off += s
    x = np.concatenate(xs)
    h, t, w = np.concatenate(hs), np.concatenate(ts), np.concatenate(ws)
    nc, comp = oracle.connected_components(x.shape[0], h, t)
    assert nc == len(sizes)
    yo, ok_o = oracle.initialize_embedding(x, 2, h, t, w, "spectral", 42)
    yg, ok_g = capi.init_embedding(x, 2, h, t, w, init=1, seed=42)
    assert ok_o and ok_g and np.isfinite(yg).all()
    # The anchors are one eigen-solve of the centroid affinities; each of its
    # columns is fixed only up to sign (the reference takes whatever its
    # eigen-solver returns), so the oracle's anchors may come out reflected
    # per axis.  Find that reflection from the component means.
    means_g = np.array([yg[comp == c].mean(0) for c in range(nc)])
    means_o = np.array([yo[comp == c].mean(0) for c in range(nc)])
    flips = [np.array(f) for f in ((1, 1), (1, -1), (-1, 1), (-1, -1))]
    flip = min(flips, key=lambda f: np.abs(means_g - means_o * f).max())
    for c in range(nc):
        m = comp == c
        if m.sum() < 4:
            # random layout from the same seeded stream around the anchor
            assert np.allclose(yg[m] - means_g[c], yo[m] - means_o[c], atol=1e-5)
            continue
        # local coordinates are eigenvectors scaled to the same range, so
        # component means (the anchor plus a small offset) and extents agree
        # and the centred coordinates span the same plane
        span = np.ptp(yo[m], axis=0).max()
        assert np.abs(means_g[c] - means_o[c] * flip).max() < 0.05 * span
        assert abs(np.ptp(yg[m], axis=0).max() / span - 1.0) < 0.02
        assert _subspace_cos(yg[m] - yg[m].mean(0), yo[m] - yo[m].mean(0)) > 0.99


def test_spectral_init_far_apart_components(gpu, oracle):
    # twelve blobs far apart (centroid affinities exp(-d^2) = 0), as in the
    # 32-cluster configuration: the meta layout puts two components on the
    # axes and packs the other ten within ~1e-4 of the origin, so their local
    # layouts are scaled into boxes of half-width <= 1e-4.  The GPU stops
    # iterating those after 16 products (any layout in such a box is within
    # 2e-4 of any other, the size of the N(0, 1e-4) noise both add) and
    # iterates the two large ones alone to the end, which must still match the
    # oracle's 256 iterations as subspaces.
    rng = np.random.default_rng(9)
    k_blobs, size = 12, 300
    centres = rng.normal(0.0, 20.0, size=(k_blobs, 16))
    xs, hs, ts, ws = [], [], [], []
    for i, c in enumerate(centres):
        xb = _stretched_blob(rng, size, 16, c)
        h, t, w = _graph(oracle, xb)
        xs.append(xb)
        hs.append(h + i * size)
        ts.append(t + i * size)
        ws.append(w)
    x = np.concatenate(xs)
    h, t, w = np.concatenate(hs), np.concatenate(ts), np.concatenate(ws)
    nc, comp = oracle.connected_components(x.shape[0], h, t)
    assert nc == k_blobs
    yo, ok_o = oracle.initialize_embedding(x, 2, h, t, w, "spectral", 42)
    yg, ok_g = capi.init_embedding(x, 2, h, t, w, init=1, seed=42)
    assert ok_o and ok_g and np.isfinite(yg).all()
    means_g = np.array([yg[comp == c].mean(0) for c in range(nc)])
    means_o = np.array([yo[comp == c].mean(0) for c in range(nc)])
    flips = [np.array(f) for f in ((1, 1), (1, -1), (-1, 1), (-1, -1))]
    flip = min(flips, key=lambda f: np.abs(means_g - means_o * f).max())
    spans = np.array([np.ptp(yo[comp == c], axis=0).max() for c in range(nc)])
    large = spans > 1.0
    assert large.sum() == 2 and (spans[~large] < 2e-3).all(), spans
    for c in range(nc):
        m = comp == c
        if large[c]:
            assert abs(np.ptp(yg[m], axis=0).max() / spans[c] - 1.0) < 0.02
            assert _subspace_cos(yg[m] - yg[m].mean(0), yo[m] - yo[m].mean(0)) > 0.99
        else:
            # anchors equal, local layouts within the box plus both noises
            assert np.abs(yg[m] - yo[m] * flip).max() <= 2e-3


@pytest.mark.parametrize("dim,scales", [(1, (4.0,)), (3, (4.0, 3.2, 2.6))])
def test_spectral_init_other_dims(gpu, oracle, dim, scales):
    # (dim + 1) * 6 iterated columns (lib.rs:4231-4234) for 1-D and 3-D output:
    # the leading axes of a blob stretched along dim axes give the smoothest
    # non-trivial eigenvectors, with a gap to the next
    rng = np.random.default_rng(31 + dim)
    sc = np.ones(24)
    sc[:dim] = scales
    x = (rng.normal(size=(3000, 24)) * sc).astype(np.float32)
    h, t, w = _graph(oracle, x)
    yo, ok_o = oracle.initialize_embedding(x, dim, h, t, w, "spectral", 42)
    yg, ok_g = capi.init_embedding(x, dim, h, t, w, init=1, seed=42)
    assert ok_o and ok_g and yg.shape == (3000, dim) and np.isfinite(yg).all()
    assert abs(np.abs(yg).max() - 10.0) < 1e-3
    assert _subspace_cos(yg, yo) > 0.99


def test_three_component_fit_uses_spectral_init(gpu, oracle):
    from umapers_amd import Umap, datasets
    from umapers_amd.diagnostics import trustworthiness

    x = datasets.gaussian_mixture(3000, 32, 5, 3.0, 8, 1.0)
    m = Umap(n_components=3, n_epochs=200, random_seed=7, ann_mode="exact")
    prof = m.profile_fit_transform(x)
    assert prof["gpu"]["init_used_spectral"] == 1
    y = prof["embedding"]
    assert y.shape == (3000, 3) and np.isfinite(y).all()
    yo = oracle.fit(x, oracle.OracleParams(n_components=3, n_epochs=200, random_seed=7)).embedding
    assert abs(trustworthiness(x, y) - trustworthiness(x, yo)) <= 0.01
