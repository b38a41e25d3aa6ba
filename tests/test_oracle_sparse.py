"""The oracle's sparse (CSR) kNN (oracle/sparse_oracle.cpp, restating
rust_umap/src/sparse.rs) pinned independently of its own code:

  - a step-by-step restatement in numpy float32 scalars on tiny inputs (the
    merge over nonzeros in column order, the expanded-form euclidean
    |a|^2 + |b|^2 - 2 a.b, the cosine / correlation summaries, the dense-query
    forms), compared bit for bit, so a dropped term, a reordered sum or a
    direct-difference euclidean fails;
  - a large-norm near-duplicate pair whose expanded-form distance differs from
    the direct difference (the cancellation the reference's arithmetic has);
  - the reference's own check that the sparse kNN matches the dense kNN on the
    same values (lib.rs sparse_csr_knn_matches_dense_euclidean, :6374-6400),
    to f32 rounding;
  - SparseCsrMatrix::new's error messages (sparse.rs:20-90).
"""
import numpy as np
import pytest

from oracle import oracle as O

sp = pytest.importorskip("scipy.sparse")
F = np.float32


def _csr(x):
    return O.Csr.of(sp.csr_matrix(x))


def _rows(c, i):
    lo, hi = c.indptr[i], c.indptr[i + 1]
    return c.indices[lo:hi], c.data[lo:hi]


def _summ(vals):
    sq, s = F(0), F(0)
    for v in vals:
        sq = F(sq + F(v * v))
    for v in vals:
        s = F(s + v)
    return sq, s, F(np.sqrt(sq))


def _merge(ai, av, bi, bv):
    """(column, a, b) over the union in column order, missing side None."""
    i = j = 0
    out = []
    while i < len(ai) and j < len(bi):
        if ai[i] < bi[j]:
            out.append((ai[i], av[i], None))
            i += 1
        elif ai[i] > bi[j]:
            out.append((bi[j], None, bv[j]))
            j += 1
        else:
            out.append((ai[i], av[i], bv[j]))
            i += 1
            j += 1
    out += [(ai[t], av[t], None) for t in range(i, len(ai))]
    out += [(bi[t], None, bv[t]) for t in range(j, len(bi))]
    return out


def _corr(nf, sx, qx, sy, qy, dot):
    n = F(nf)
    cd = F(dot - F(F(sx * sy) / n))
    xv = max(F(qx - F(F(sx * sx) / n)), F(0))
    yv = max(F(qy - F(F(sy * sy) / n)), F(0))
    xn, yn = F(np.sqrt(xv)), F(np.sqrt(yv))
    if xn == 0 and yn == 0:
        return F(0)
    if xn == 0 or yn == 0:
        return F(1)
    return F(F(1) - F(np.clip(F(cd / F(xn * yn)), -1, 1)))


def _pair(a, ia, b, ib, metric):
    ai, av = _rows(a, ia)
    bi, bv = _rows(b, ib)
    sqa, sa, la = _summ(av)
    sqb, sb, lb = _summ(bv)
    m = _merge(ai, av, bi, bv)
    dot = F(0)
    for _, x, y in m:
        if x is not None and y is not None:
            dot = F(dot + F(x * y))
    if metric == "euclidean":
        return F(np.sqrt(max(F(F(sqa + sqb) - F(F(2) * dot)), F(0))))
    if metric == "cosine":
        if la == 0 and lb == 0:
            return F(0)
        if la == 0 or lb == 0:
            return F(1)
        return F(F(1) - F(np.clip(F(dot / F(la * lb)), -1, 1)))
    if metric == "manhattan":
        s = F(0)
        for _, x, y in m:
            s = F(s + abs(F(x if y is None else (-y if x is None else F(x - y)))))
        return s
    if metric == "chebyshev":
        s = F(0)
        for _, x, y in m:
            s = max(s, abs(F(x if y is None else (-y if x is None else F(x - y)))))
        return s
    return _corr(a.n_cols, sa, sqa, sb, sqb, dot)


def _dense_pair(q, b, ib, metric):
    bi, bv = _rows(b, ib)
    sqb, sb, lb = _summ(bv)
    sq, l1, s = F(0), F(0), F(0)
    for v in q:
        sq = F(sq + F(v * v))
    for v in q:
        l1 = F(l1 + abs(v))
    for v in q:
        s = F(s + v)
    dot = F(0)
    for c, v in zip(bi, bv):
        dot = F(dot + F(q[c] * v))
    if metric == "euclidean":
        return F(np.sqrt(max(F(F(sq + sqb) - F(F(2) * dot)), F(0))))
    if metric == "cosine":
        lq = F(np.sqrt(sq))
        if lq == 0 and lb == 0:
            return F(0)
        if lq == 0 or lb == 0:
            return F(1)
        return F(F(1) - F(np.clip(F(dot / F(lq * lb)), -1, 1)))
    if metric == "manhattan":
        acc = l1
        for c, v in zip(bi, bv):
            acc = F(acc + F(abs(F(q[c] - v)) - abs(q[c])))
        return acc
    if metric == "chebyshev":
        acc, pos = F(0), 0
        for f in range(len(q)):
            y = F(0)
            if pos < len(bi) and bi[pos] == f:
                y = bv[pos]
                pos += 1
            acc = max(acc, abs(F(q[f] - y)))
        return acc
    return _corr(len(q), s, sq, sb, sqb, dot)


def _topk(d, k):
    order = sorted(range(len(d)), key=lambda j: (d[j], j))[:k]
    return np.array(order), np.array([d[j] for j in order], np.float32)


def _data(n, c, seed, density=0.3, scale=1.0):
    rng = np.random.default_rng(seed)
    x = (rng.normal(size=(n, c)) * scale).astype(np.float32)
    x[rng.random((n, c)) > density] = 0.0
    x[2] = 0.0  # an empty row
    return x


METRICS = ["euclidean", "cosine", "manhattan", "chebyshev", "correlation"]


@pytest.mark.parametrize("metric", METRICS)
def test_fit_knn_bit_exact_vs_scalar_restatement(metric):
    x = _data(14, 23, 1)
    c = _csr(x)
    idx, dist = O.knn_csr(c, 6, metric)
    for i in range(c.n_rows):
        d = [F(0) if i == j else _pair(c, i, c, j, metric) for j in range(c.n_rows)]
        ei, ed = _topk(d, 6)
        assert np.array_equal(idx[i], ei) and np.array_equal(dist[i], ed), (metric, i)


@pytest.mark.parametrize("metric", METRICS)
def test_query_knn_bit_exact_vs_scalar_restatement(metric):
    r = _csr(_data(13, 19, 2))
    qx = _data(5, 19, 3)
    q = _csr(qx)
    idx, dist = O.knn_csr_query(q, r, 5, metric)
    didx, ddist = O.knn_dense_to_csr(qx, r, 5, metric)
    for i in range(q.n_rows):
        ei, ed = _topk([_pair(q, i, r, j, metric) for j in range(r.n_rows)], 5)
        assert np.array_equal(idx[i], ei) and np.array_equal(dist[i], ed), (metric, i)
        ei, ed = _topk([_dense_pair(qx[i], r, j, metric) for j in range(r.n_rows)], 5)
        assert np.array_equal(didx[i], ei) and np.array_equal(ddist[i], ed), (metric, i)


def test_large_norm_near_duplicate_keeps_the_expanded_form():
    # row 1 is row 0 with one value nudged by 1 ulp-scale: the direct
    # difference is tiny, while |a|^2 + |b|^2 - 2 a.b cancels to a different
    # value at this norm; the reference's sparse path computes the latter
    rng = np.random.default_rng(9)
    base = (rng.normal(size=40) * 3000).astype(np.float32)
    base[rng.random(40) > 0.5] = 0
    x = np.stack([base, base.copy(), (rng.normal(size=40)).astype(np.float32)])
    nz = np.flatnonzero(base)
    x[1, nz[3]] = np.nextafter(x[1, nz[3]], np.float32(np.inf)) * np.float32(1.0000005)
    c = _csr(x)
    idx, dist = O.knn_csr(c, 3, "euclidean")
    expanded = _pair(c, 0, c, 1, "euclidean")
    direct = F(np.sqrt(np.sum((x[0].astype(np.float64) - x[1]) ** 2)))
    assert dist[0][list(idx[0]).index(1)] == expanded
    assert expanded != direct
    # the row against itself is 0.0 in the fit's kNN, as exact_nearest_neighbors sets it
    assert idx[0][0] == 0 and dist[0][0] == 0


@pytest.mark.parametrize("metric", METRICS + ["minkowski"])
def test_sparse_knn_matches_dense_knn(metric):
    # lib.rs sparse_csr_knn_matches_dense_euclidean: the same values give the
    # same neighbours; distances agree to f32 rounding of the two arithmetics
    x = _data(300, 120, 5, density=0.1)
    c = _csr(x)
    si, sd = O.knn_csr(c, 10, metric, p=3.0)
    if metric in ("euclidean", "cosine"):
        di, dd = O.knn(x, 10, metric)
    else:
        di, dd = O.knn_metric(x, x, 10, metric, p=3.0)
    assert np.mean(si == di) > 0.99
    assert np.allclose(sd, dd, rtol=1e-4, atol=1e-5)


def test_csr_checks_repeat_the_reference_messages():
    ip = np.array([0, 2, 3], np.int64)
    ix = np.array([0, 2, 1], np.int64)
    dt = np.array([1.0, 2.0, 3.0], np.float32)
    assert O.csr_check(O.Csr(ip, ix, dt, 3)) is None
    assert O.csr_check(O.Csr(ip, ix, dt, 0)) == "csr n_cols must be >= 1"
    assert O.csr_check(O.Csr(np.array([1, 2, 3]), ix, dt, 3)) == "csr indptr must start from 0"
    assert O.csr_check(O.Csr(np.array([0, 2, 2]), ix, dt, 3)) == \
        "csr indptr last value (2) must equal nnz (3)"
    assert O.csr_check(O.Csr(np.array([0, 3, 2, 3]), np.array([0, 1, 2]), dt, 3)) == \
        "csr indptr must be non-decreasing, got indptr[1]=3 > indptr[2]=2"
    assert O.csr_check(O.Csr(ip, np.array([2, 0, 1]), dt, 3)) == \
        "csr row 0 indices must be strictly increasing"
    assert O.csr_check(O.Csr(ip, np.array([0, 5, 1]), dt, 3)) == \
        "csr row 0 has column index 5 out of bounds for n_cols=3"
    assert O.csr_check(O.Csr(ip, ix, np.array([1.0, np.nan, 3.0], np.float32), 3)) == \
        "csr data must be finite"
