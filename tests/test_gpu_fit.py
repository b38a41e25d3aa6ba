"""End-to-end fit / transform through the Umap facade on the GPU."""
import numpy as np
import pytest

from umapers_amd import Umap, datasets
from umapers_amd.diagnostics import knn_preservation, trustworthiness

from . import capi

pytestmark = pytest.mark.gpu


def test_c1_fit_parity_with_oracle(gpu, oracle):
    # i.i.d. Gaussian rows have no structure, so single-seed trustworthiness
    # moves by ~0.006 (the init planes agree to cosine 1.0; the SGD walk
    # amplifies last-bit differences); the parity bar (0.01) is applied to
    # 10-seed means, whose difference has a spread of about 0.002
    x = datasets.c1()
    tg, to, kg, ko = [], [], [], []
    for seed in range(10):
        y = Umap(n_epochs=200, ann_mode="exact", random_seed=seed).fit_transform(x)
        yo = oracle.fit(x, oracle.OracleParams(n_epochs=200, random_seed=seed)).embedding
        assert y.shape == (1000, 2) and y.dtype == np.float32 and np.isfinite(y).all()
        tg.append(trustworthiness(x, y))
        to.append(trustworthiness(x, yo))
        kg.append(knn_preservation(x, y))
        ko.append(knn_preservation(x, yo))
    assert abs(np.mean(tg) - np.mean(to)) <= 0.01, (tg, to)
    assert abs(np.mean(kg) - np.mean(ko)) <= 0.01, (kg, ko)


def test_clustered_fit_parity_and_profile(gpu, oracle):
    x, lab = datasets.gaussian_mixture(6000, 96, 8, 4.0, 3, 1.0, return_labels=True)
    model = Umap(n_epochs=200, ann_mode="exact")
    prof = model.profile_fit_transform(x)
    y = prof["embedding"]
    for key in ("knn_sec", "smooth_knn_sec", "init_sec", "optimize_sec", "total_sec"):
        assert key in prof["timings"]
    assert prof["n_epochs"] == 200 and prof["n_edges"] > 0
    assert prof["used_approximate_knn"] is False
    fo = oracle.fit(x, oracle.OracleParams(n_epochs=200))
    assert abs(trustworthiness(x, y) - trustworthiness(x, fo.embedding)) <= 0.01
    assert abs(knn_preservation(x, y) - knn_preservation(x, fo.embedding)) <= 0.01


def test_bf16_rerank_fit(gpu, oracle):
    x = datasets.gaussian_mixture(5000, 128, 8, 3.0, 11, (0.5, 1.5))
    y = Umap(n_epochs=200, knn_precision="bf16_rerank").fit_transform(x)
    yo = oracle.fit(x, oracle.OracleParams(n_epochs=200)).embedding
    assert abs(trustworthiness(x, y) - trustworthiness(x, yo)) <= 0.01
    assert abs(knn_preservation(x, y) - knn_preservation(x, yo)) <= 0.01


def test_cosine_fit(gpu, oracle):
    x = datasets.gaussian_mixture(3000, 64, 6, 3.0, 12, 1.0)
    y = Umap(n_epochs=200, metric="cosine").fit_transform(x)
    yo = oracle.fit(x, oracle.OracleParams(n_epochs=200, metric="cosine")).embedding
    xn = x / np.linalg.norm(x, axis=1, keepdims=True)
    assert abs(trustworthiness(xn, y) - trustworthiness(xn, yo)) <= 0.01
    assert abs(knn_preservation(xn, y) - knn_preservation(xn, yo)) <= 0.01


def test_precomputed_knn_matches_direct(gpu, oracle):
    x = datasets.gaussian_mixture(2500, 48, 5, 3.0, 13, 1.0)
    idx, dist = oracle.knn(x, 20)
    m1 = Umap(n_neighbors=15, n_epochs=150)
    y1 = m1.fit_transform_with_knn(x, idx, dist)
    y2 = Umap(n_neighbors=15, n_epochs=150).fit_transform(x)
    assert abs(trustworthiness(x, y1) - trustworthiness(x, y2)) <= 0.01
    assert m1.embedding_.shape == (2500, 2) and m1.n_features_in_ == 48


def test_precomputed_knn_errors(gpu):
    x = datasets.gaussian_mixture(200, 8, 2, 3.0, 1, 1.0)
    idx = np.tile(np.arange(15), (200, 1))
    dist = np.tile(np.linspace(0, 1, 15, dtype=np.float32), (200, 1))
    m = Umap(n_neighbors=15)
    with pytest.raises(ValueError, match="must match model metric"):
        m.fit_transform_with_knn(x, idx, dist, knn_metric="cosine")
    with pytest.raises(ValueError, match="knn columns must be >= n_neighbors"):
        m.fit_transform_with_knn(x, idx[:, :10], dist[:, :10])
    bad = idx.copy()
    bad[3, 4] = 999
    with pytest.raises(ValueError, match="index out of range at row 3, col 4"):
        m.fit_transform_with_knn(x, bad, dist)
    dup = idx.copy()
    dup[2, 5] = dup[2, 4]
    with pytest.raises(ValueError, match="row 2 contains duplicate index"):
        m.fit_transform_with_knn(x, dup, dist)
    dd = dist.copy()
    dd[7, 3] = 0.0
    with pytest.raises(ValueError, match="non-decreasing"):
        m.fit_transform_with_knn(x, idx, dd)
    dn = dist.copy()
    dn[1, 1] = np.nan
    with pytest.raises(ValueError, match="knn_dists must contain only finite values"):
        m.fit_transform_with_knn(x, idx, dn)
    neg = idx.copy()
    neg[0, 0] = -1
    with pytest.raises(ValueError, match="knn indices must be non-negative integers"):
        m.fit_transform_with_knn(x, neg, dist)


def test_plain_transform_parity(gpu, oracle):
    x, lab = datasets.gaussian_mixture(4000, 64, 6, 4.0, 14, 1.0, return_labels=True)
    q = datasets.gaussian_mixture(500, 64, 6, 4.0, 14, 1.0)  # same centres, same seed
    q = q + np.random.default_rng(1).normal(0, 0.3, size=q.shape).astype(np.float32)
    model = Umap(n_epochs=200, random_seed=5)
    y = model.fit_transform(x)
    yq = model.transform(q)
    assert yq.shape == (500, 2) and np.isfinite(yq).all()
    yo = oracle.transform(x, y, q, oracle.OracleParams(n_epochs=200, random_seed=5))
    # judged against the oracle's transform of the same fitted embedding
    both_x = np.concatenate([x, q])
    tg = trustworthiness(both_x, np.concatenate([y, yq]))
    to = trustworthiness(both_x, np.concatenate([y, yo]))
    qg, qo = trustworthiness(q, yq), trustworthiness(q, yo)
    print(f"transform: trust (fit + queries) gpu {tg:.4f} oracle {to:.4f}; queries alone gpu "
          f"{qg:.4f} oracle {qo:.4f}", flush=True)
    assert abs(tg - to) <= 0.01, (tg, to)
    assert abs(qg - qo) <= 0.02, (qg, qo)
    # a query lands near the training rows it resembles
    d = np.linalg.norm(yq - yo, axis=1)
    assert np.median(d) < 1.0


def test_transform_errors(gpu):
    m = Umap()
    with pytest.raises(RuntimeError, match="not fitted"):
        m.transform(np.zeros((3, 4), np.float32))
    x = datasets.gaussian_mixture(300, 10, 3, 3.0, 2, 1.0)
    m.fit(x)
    with pytest.raises(ValueError, match="feature mismatch: expected 10, got 9"):
        m.transform(np.zeros((3, 9), np.float32))
    # UmapError::FeatureMismatch { expected: 10, got: 9 } through the ABI
    import ctypes

    from umapers_amd import _lib as L
    row, exp, got = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    code = L.lib().umapx_last_error_detail(ctypes.byref(row), ctypes.byref(exp), ctypes.byref(got))
    assert (code, row.value, exp.value, got.value) == (L.ERR_FEATURE_MISMATCH, -1, 10, 9)
    with pytest.raises(ValueError, match="embedding dimension mismatch: expected 2, got 3"):
        m.inverse_transform(np.zeros((2, 3), np.float32))
    code = L.lib().umapx_last_error_detail(ctypes.byref(row), ctypes.byref(exp), ctypes.byref(got))
    assert (code, row.value, exp.value, got.value) == (L.ERR_EMBEDDING_DIMENSION_MISMATCH, -1, 2, 3)
    q = np.zeros((4, 10), np.float32)
    q[2, 3] = np.inf
    with pytest.raises(ValueError, match="non-finite value at row 2"):
        m.transform(q)
    assert m.transform(np.zeros((0, 10), np.float32)).shape == (0, 2)


def test_output_dens_radii(gpu, oracle):
    x = datasets.gaussian_mixture(2000, 32, 4, 3.0, 15, 1.0)
    emb, ro, re = Umap(n_epochs=100, output_dens=True).fit_transform(x)
    idx, dist = oracle.knn(x, 15)
    assert np.allclose(ro, oracle.local_radii(idx, dist), rtol=1e-6, atol=1e-6)
    assert np.allclose(re, oracle.embedding_radii(emb, idx), rtol=1e-5, atol=1e-5)


def test_densmap_fit(gpu, oracle):
    # three seeds each way: mean trustworthiness within 0.01 both ways and the
    # original / embedded radius correlation within 0.05 both ways
    x = datasets.gaussian_mixture(3000, 96, 5, 4.0, 16, 1.0)
    tg, to, cg, co = [], [], [], []
    for s in (42, 7, 11):
        m = Umap(n_epochs=200, densmap=True, dens_lambda=2.0, dens_frac=0.3, output_dens=True,
                 random_seed=s)
        emb, ro, re = m.fit_transform(x)
        fo = oracle.fit(x, oracle.OracleParams(n_epochs=200, densmap=True, output_dens=True,
                                               random_seed=s))
        tg.append(trustworthiness(x, emb))
        to.append(trustworthiness(x, fo.embedding))
        cg.append(np.corrcoef(ro, re)[0, 1])
        co.append(np.corrcoef(fo.rad_orig, fo.rad_emb)[0, 1])
    print(f"densMAP fit: trust gpu {np.mean(tg):.4f} oracle {np.mean(to):.4f}; radius corr gpu "
          f"{np.mean(cg):.3f} oracle {np.mean(co):.3f}", flush=True)
    assert abs(np.mean(tg) - np.mean(to)) <= 0.01, (tg, to)
    assert abs(np.mean(cg) - np.mean(co)) <= 0.05, (cg, co)


def test_out_buffer_and_params(gpu):
    x = datasets.gaussian_mixture(500, 16, 3, 3.0, 4, 1.0)
    out = np.empty((500, 2), np.float32)
    m = Umap(init="random", n_epochs=50)
    r = m.fit_transform(x, out=out)
    assert r is out and np.isfinite(out).all()
    with pytest.raises(ValueError, match="output buffer shape mismatch"):
        m.fit_transform(x, out=np.empty((10, 2), np.float32))
    assert m.get_params()["n_epochs"] == 50


def test_transform_launch_configuration_determinism(gpu):
    # the analog of lib.rs:6193-6236 (the same transform whatever the rayon
    # thread count).  A query's kNN into the training rows is row-local, so it
    # comes out bit for bit the same alone, in a short batch or in a long one
    # (different grids, block splits and merges); the whole transform of one
    # batch is the same on a repeat and after unrelated GPU work.  (Across
    # batches the transform itself differs, in the reference too: pruning and
    # the sampling schedule use the batch's largest edge weight.)
    x = datasets.gaussian_mixture(3000, 40, 5, 3.0, 21, 1.0)
    q = datasets.gaussian_mixture(700, 40, 5, 3.0, 21, 1.0) + np.float32(0.05)
    i_all, d_all = capi.knn_to_reference(q, x, 15)
    for rows in (slice(0, 1), slice(0, 37), slice(300, 700)):
        i_part, d_part = capi.knn_to_reference(q[rows], x, 15)
        assert np.array_equal(i_part, i_all[rows]) and np.array_equal(d_part, d_all[rows])
    m = Umap(n_epochs=100, random_seed=4, ann_mode="exact")
    m.fit(x)
    full = m.transform(q)
    assert np.array_equal(full, m.transform(q))
    Umap(n_epochs=30, random_seed=9, ann_mode="exact").fit_transform(x[:900])
    assert np.array_equal(full, m.transform(q))


def test_fit_with_50_neighbours(gpu, oracle):
    # n_neighbors = 50 (more than the candidate lists' 32): exact kNN on the
    # vector ALUs, a graph of up to 64 entries per row, and the transforms;
    # quality within 0.01 of the oracle's over three seeds, both ways
    from umapers_amd.diagnostics import knn_preservation

    x = datasets.gaussian_mixture(3000, 32, 6, 3.0, 21, 1.0)
    tg, to, kg, ko = [], [], [], []
    for s in (1, 2, 3):
        m = Umap(n_neighbors=50, n_epochs=200, random_seed=s, ann_mode="exact")
        y = m.fit_transform(x)
        yo = oracle.fit(x, oracle.OracleParams(n_neighbors=50, n_epochs=200,
                                               random_seed=s)).embedding
        tg.append(trustworthiness(x, y))
        to.append(trustworthiness(x, yo))
        kg.append(knn_preservation(x, y))
        ko.append(knn_preservation(x, yo))
    print(f"k=50 fit: trust gpu {np.mean(tg):.4f} oracle {np.mean(to):.4f}; knn-pres gpu "
          f"{np.mean(kg):.4f} oracle {np.mean(ko):.4f}", flush=True)
    assert abs(np.mean(tg) - np.mean(to)) <= 0.01, (tg, to)
    assert abs(np.mean(kg) - np.mean(ko)) <= 0.01, (kg, ko)
    q = x[:200] + np.float32(0.05)
    assert np.isfinite(m.transform(q)).all()
    assert np.isfinite(m.inverse_transform(y[:20])).all()
