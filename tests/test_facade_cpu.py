"""Facade logic that needs no GPU: argument normalisation, errors, out=
buffers and fitted attributes, with a fake core standing in for the GPU (the
reference's binding tests use the same pattern, test_binding.py:214-248)."""
import numpy as np
import pytest

import umapers_amd._api as api
from umapers_amd import Umap


class FakeCore:
    """Routes the facade to the CPU oracle (test-only)."""

    def __init__(self, params):
        from oracle import oracle as O

        self.O = O
        self.p = params
        self.n_components = params.n_components
        self.emb = None
        self.calls = []

    def _op(self):
        return self.O.OracleParams(n_neighbors=self.p.n_neighbors, n_components=self.p.n_components,
                                   n_epochs=self.p.n_epochs or None, init="random")

    def fit(self, arr):
        self.calls.append("fit")
        self.emb = self.O.fit(arr, self._op()).embedding

    def fit_transform_into(self, arr, out):
        self.calls.append("fit_transform")
        out[...] = self.O.fit(arr, self._op()).embedding
        self.emb = out.copy()

    def fit_transform(self, arr):
        out = np.empty((arr.shape[0], self.n_components), np.float32)
        self.fit_transform_into(arr, out)
        return out

    def fit_transform_csr_into(self, c, out):
        # the oracle's sparse kNN, then the fit from that kNN
        self.calls.append("fit_transform_csr")
        ref = self.O.Csr(c.indptr, c.indices, c.data, c.shape[1])
        idx, dist = self.O.knn_csr(ref, self.p.n_neighbors)
        dense = np.zeros(c.shape, np.float32)
        for r in range(c.shape[0]):
            lo, hi = c.indptr[r], c.indptr[r + 1]
            dense[r, c.indices[lo:hi]] = c.data[lo:hi]
        emb = self.O.fit(dense, self._op(), knn_idx=idx, knn_dist=dist).embedding
        if out is not None:
            out[...] = emb
        self.emb = emb.copy()

    @property
    def embedding(self):
        return self.emb

    def radii(self):
        return None


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(api, "_Core", FakeCore)


def test_defaults_and_get_set_params(fake):
    m = Umap()
    p = m.get_params()
    assert p["n_neighbors"] == 15 and p["n_components"] == 2 and p["init"] == "spectral"
    assert p["ann_mode"] == "auto" and p["knn_precision"] == "fp32"
    m.set_params(n_neighbors=7, ann_mode="exact")
    assert m.n_neighbors == 7 and m.use_approximate_knn is False
    with pytest.raises(ValueError, match="Invalid parameter 'nope'"):
        m.set_params(nope=1)


@pytest.mark.parametrize("kw,msg", [
    ({"metric": "bogus"}, "unsupported metric 'bogus'"),
    ({"metric": "minkowski"}, r"requires metric_kwds=\{'p': \.\.\.\}"),
    ({"metric": "minkowski", "metric_kwds": {"p": -1.0}}, "must be finite and > 0 for minkowski"),
    ({"metric": "minkowski", "metric_kwds": {"p": 3, "w": 1}}, "unsupported metric_kwds for minkowski: w"),
    ({"metric": "manhattan", "metric_kwds": {"p": 1}}, "unsupported metric_kwds for manhattan: p"),
    ({"ann_mode": "fast"}, "unsupported ann_mode 'fast'"),
    ({"dens_lambda": -1.0}, "dens_lambda must be finite and >= 0"),
    ({"init": "pca"}, "unsupported init 'pca'"),
    ({"knn_precision": "fp16"}, "unsupported knn_precision"),
    ({"transform_mode": "magic"}, "unsupported transform_mode"),
    ({"target_metric": "categorical"}, "not available on the GPU path"),
])
def test_constructor_errors(fake, kw, msg):
    with pytest.raises(ValueError, match=msg):
        Umap(**kw)


def test_metric_aliases(fake):
    assert Umap(metric="L2").metric == "euclidean"
    assert Umap(metric="cosine").metric == "cosine"
    # _api.py:242-263 aliases; all eight metrics reach the C parameters
    for alias, name, code in [("l1", "manhattan", 2), ("linf", "chebyshev", 3),
                              ("CORRELATION", "correlation", 5), ("canberra", "canberra", 6),
                              ("bray_curtis", "braycurtis", 7)]:
        m = Umap(metric=alias)
        assert m.metric == name and m._c_params().metric == code
    m = Umap(metric="minkowski", metric_kwds={"p": 3})
    assert m._c_params().metric == 4 and m._c_params().metric_p == 3.0


def test_input_normalisation_and_out_buffer(fake):
    x = np.random.default_rng(0).normal(size=(60, 5))  # float64, converted
    m = Umap(n_neighbors=5, n_epochs=20)
    y = m.fit_transform(x.tolist())
    assert y.shape == (60, 2) and y.dtype == np.float32
    assert m.n_features_in_ == 5 and m.n_samples_fit_ == 60
    with pytest.raises(ValueError, match="must be a 2D array"):
        m.fit_transform(np.zeros(5))
    with pytest.raises(TypeError, match="out dtype must be float32"):
        m.fit_transform(x, out=np.empty((60, 2)))
    with pytest.raises(ValueError, match="C-contiguous"):
        m.fit_transform(x, out=np.empty((2, 60), np.float32).T)
    out = np.empty((60, 2), np.float32)
    assert m.fit_transform(x, out=out) is out


def test_sparse_input_rules(fake):
    # the reference's sparse-path rules (test_binding.py:1314-1530): CSR / CSC /
    # COO fit like their dense rows; densMAP and canberra / braycurtis are
    # dense-only; sparse queries need a sparse fit; no inverse for sparse fits
    sp = pytest.importorskip("scipy.sparse")
    x = np.random.default_rng(3).normal(size=(60, 6)).astype(np.float32)
    x[x < 0.2] = 0.0
    for conv in (sp.csr_matrix, sp.csc_matrix, sp.coo_matrix):
        m = Umap(n_neighbors=5, n_epochs=10)
        assert m.fit_transform(conv(x)).shape == (60, 2)
        assert m._core.calls[-1] == "fit_transform_csr" and m._fitted_sparse
    with pytest.raises(ValueError, match="dense input only"):
        Umap(densmap=True).fit(sp.csr_matrix(x))
    with pytest.raises(ValueError, match="dense input only"):
        Umap(output_dens=True).fit(sp.csr_matrix(x))
    with pytest.raises(ValueError, match="Canberra metric is currently supported for dense input only"):
        Umap(metric="canberra").fit(sp.csr_matrix(x))
    # no feature cap: a 3000-column CSR input reaches the sparse entry point
    wide = sp.random(30, 3000, density=0.01, random_state=4, format="csr", dtype=np.float32)
    wm = Umap(n_neighbors=5, n_epochs=10)
    wm.fit(wide)
    assert wm._core.calls[-1] == "fit_transform_csr" and wm._fitted_sparse
    dense = Umap(n_neighbors=5, n_epochs=10)
    dense.fit(x)
    with pytest.raises(ValueError, match="supported only for sparse-fitted models"):
        dense.transform(sp.csr_matrix(x[:3]))
    m = Umap(n_neighbors=5, n_epochs=10)
    m.fit(sp.csr_matrix(x))
    with pytest.raises(ValueError, match="not supported for sparse-trained models"):
        m.inverse_transform(np.zeros((2, 2), np.float32))


def test_missing_library_fails_loudly(monkeypatch):
    from umapers_amd import _lib as L

    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", "/nonexistent/libumapx.so")
    with pytest.raises(L.UmapxNotBuilt):
        Umap()
