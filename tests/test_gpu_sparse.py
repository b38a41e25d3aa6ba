"""Sparse input (CSR / CSC / COO) on the GPU path, after
umap_rs/tests/test_binding.py:1314-1530: the rows are densified and fitted
like their dense copy, so a sparse query and its dense copy transform to the
same embedding, and the dense-only rules raise the reference's errors."""
import numpy as np
import pytest

from umapers_amd import Umap

pytestmark = pytest.mark.gpu
sp = pytest.importorskip("scipy.sparse")


def _data(n, d, seed, cut):
    x = np.random.default_rng(seed).normal(size=(n, d)).astype(np.float32)
    x[x < cut] = 0.0
    return x


@pytest.mark.parametrize("metric,kw", [("chebyshev", None), ("minkowski", {"p": 3.0}),
                                       ("correlation", None)])
def test_sparse_fit_other_metrics(gpu, metric, kw):
    x = _data(64, 12, 56, 0.15)
    m = Umap(n_neighbors=7, n_epochs=30, metric=metric, metric_kwds=kw, init="random",
             random_seed=91, ann_mode="exact")
    emb = m.fit_transform(sp.csr_matrix(x))
    t = m.transform(x[:8])
    assert emb.shape == (64, 2) and t.shape == (8, 2)
    assert np.all(np.isfinite(emb)) and np.all(np.isfinite(t))


def test_sparse_query_equals_dense_query(gpu):
    x = _data(84, 16, 58, 0.2)
    m = Umap(n_neighbors=7, n_epochs=40, metric="cosine", init="random", random_seed=37,
             ann_mode="exact")
    m.fit_transform(sp.csr_matrix(x))
    dense = m.transform(x[:11])
    for conv in (sp.csr_matrix, sp.coo_matrix, sp.csc_matrix):
        assert np.array_equal(m.transform(conv(x[:11])), dense)
    out = np.empty((11, 2), np.float32)
    assert m.transform(sp.csr_matrix(x[:11]), out=out) is out
    assert np.array_equal(out, dense)


def test_sparse_fit_matches_dense_fit(gpu):
    # the same rows fitted dense and sparse run the same kernels on the same
    # bytes; only the Hogwild SGD's float atomics differ between runs
    from umapers_amd.diagnostics import trustworthiness

    # (one run's trustworthiness on these rows varies by about 0.01, so three
    # seeds are averaged)
    x = _data(2000, 20, 7, 0.3)
    ta, tb = [], []
    for seed in (5, 6, 7):
        kw = dict(n_neighbors=10, n_epochs=100, random_seed=seed, ann_mode="exact")
        ta.append(trustworthiness(x, Umap(**kw).fit_transform(x)))
        tb.append(trustworthiness(x, Umap(**kw).fit_transform(sp.csr_matrix(x))))
    assert abs(np.mean(ta) - np.mean(tb)) < 0.01, (ta, tb)


def test_sparse_trained_inverse_errors(gpu):
    x = _data(64, 9, 71, 0.2)
    m = Umap(n_neighbors=6, n_epochs=40, metric="manhattan", init="random", random_seed=29,
             ann_mode="exact")
    m.fit(sp.csr_matrix(x))
    with pytest.raises(ValueError, match="output buffer shape mismatch"):
        m.inverse_transform(np.zeros((5, 2), np.float32), out=np.empty((5, 10), np.float32))
    with pytest.raises(ValueError, match="not supported for sparse-trained models"):
        m.inverse_transform(np.zeros((5, 2), np.float32))
