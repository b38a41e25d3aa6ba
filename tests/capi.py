"""Thin numpy wrappers over the C ABI's stage entry points (for tests)."""
import ctypes

import numpy as np

from umapers_amd import _lib as L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def knn(x, k, metric=0, precision=0, rerank=2):
    x = _f32(x)
    n, d = x.shape
    idx = np.empty((n, k), np.int64)
    dist = np.empty((n, k), np.float32)
    L.check(L.lib().umapx_knn(L.ptr(x), n, d, k, metric, precision, rerank, L.ptr(idx, L._i64p),
                              L.ptr(dist)))
    return idx, dist


def knn_to_reference(q, r, k, metric=0, precision=0, rerank=2):
    q, r = _f32(q), _f32(r)
    m, d = q.shape
    idx = np.empty((m, k), np.int64)
    dist = np.empty((m, k), np.float32)
    L.check(L.lib().umapx_knn_to_reference(L.ptr(q), m, L.ptr(r), r.shape[0], d, k, metric,
                                           precision, rerank, L.ptr(idx, L._i64p), L.ptr(dist)))
    return idx, dist


def include_self(idx, dist):
    idx, dist = _i64(idx), _f32(dist)
    out = ctypes.c_int32()
    L.check(L.lib().umapx_knn_rows_include_self(L.ptr(idx, L._i64p), L.ptr(dist), idx.shape[0],
                                                idx.shape[1], ctypes.byref(out)))
    return bool(out.value)


def smooth(dist, k_value, local_connectivity=1.0, include_self=True, fp64=True):
    dist = _f32(dist)
    n, k = dist.shape
    sig = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    L.check(L.lib().umapx_smooth_knn_dist(L.ptr(dist), n, k, k_value, local_connectivity, 1.0,
                                          int(include_self), int(fp64), L.ptr(sig), L.ptr(rho)))
    return sig, rho


def fuzzy_graph(idx, dist, sig, rho, mix=1.0, prune_epochs=0):
    idx, dist, sig, rho = _i64(idx), _f32(dist), _f32(sig), _f32(rho)
    n, k = idx.shape
    cap = 2 * n * k
    h = np.empty(cap, np.int64)
    t = np.empty(cap, np.int64)
    w = np.empty(cap, np.float32)
    ne = ctypes.c_int64()
    L.check(L.lib().umapx_fuzzy_graph(L.ptr(idx, L._i64p), L.ptr(dist), L.ptr(sig), L.ptr(rho), n,
                                      k, mix, prune_epochs, cap, L.ptr(h, L._i64p),
                                      L.ptr(t, L._i64p), L.ptr(w), ctypes.byref(ne)))
    m = ne.value
    return h[:m].copy(), t[:m].copy(), w[:m].copy()


def components(n, h, t, w):
    h, t, w = _i64(h), _i64(t), _f32(w)
    lab = np.empty(n, np.int64)
    nc = ctypes.c_int64()
    L.check(L.lib().umapx_connected_components(n, L.ptr(h, L._i64p), L.ptr(t, L._i64p), L.ptr(w),
                                               h.shape[0], L.ptr(lab, L._i64p), ctypes.byref(nc)))
    return nc.value, lab


def init_embedding(x, dim, h, t, w, init=1, seed=42):
    x, h, t, w = _f32(x), _i64(h), _i64(t), _f32(w)
    n, d = x.shape
    y = np.empty((n, dim), np.float32)
    used = ctypes.c_int32()
    L.check(L.lib().umapx_initialize_embedding(L.ptr(x), n, d, dim, L.ptr(h, L._i64p),
                                               L.ptr(t, L._i64p), L.ptr(w), h.shape[0], init,
                                               seed, L.ptr(y), ctypes.byref(used)))
    return y, bool(used.value)


def optimize(y, h, t, w, n_epochs, a, b, lr=1.0, nsr=5, gamma=1.0, seed=7, radii=None,
             dens_start=0, dens_lambda=2.0, var_shift=0.1, counts=False):
    y = np.array(y, dtype=np.float32, order="C", copy=True)
    h, t, w = _i64(h), _i64(t), _f32(w)
    ne = h.shape[0]
    pc = np.zeros(max(ne, 1), np.int32) if counts else None
    nc = np.zeros(max(ne, 1), np.int32) if counts else None
    rad = None if radii is None else _f32(radii)
    L.check(L.lib().umapx_optimize_layout(
        L.ptr(y), y.shape[0], y.shape[1], L.ptr(h, L._i64p), L.ptr(t, L._i64p), L.ptr(w), ne,
        n_epochs, a, b, lr, nsr, gamma, seed, None if rad is None else L.ptr(rad), dens_start,
        dens_lambda, var_shift, None if pc is None else L.ptr(pc, L._i32p),
        None if nc is None else L.ptr(nc, L._i32p)))
    if counts:
        return y, pc[:ne], nc[:ne]
    return y


def local_radii(idx, dist, vs=0.1):
    idx, dist = _i64(idx), _f32(dist)
    out = np.empty(idx.shape[0], np.float32)
    L.check(L.lib().umapx_local_radii(L.ptr(idx, L._i64p), L.ptr(dist), idx.shape[0], idx.shape[1],
                                      vs, L.ptr(out)))
    return out


def embedding_radii(y, idx, vs=0.1):
    y, idx = _f32(y), _i64(idx)
    out = np.empty(y.shape[0], np.float32)
    L.check(L.lib().umapx_embedding_radii(L.ptr(y), y.shape[0], y.shape[1], L.ptr(idx, L._i64p),
                                          idx.shape[1], vs, L.ptr(out)))
    return out


def mlp_param_count(d, h, dim):
    out = ctypes.c_int64()
    L.check(L.lib().umapx_mlp_param_count(d, h, dim, ctypes.byref(out)))
    return out.value


def mlp_batch_grad(params, mean, inv_std, x_rows, y_rows, d, h, dim, n_edges, n_neg, a, b, lam):
    params, mean, inv_std = _f32(params), _f32(mean), _f32(inv_std)
    x_rows, y_rows = _f32(x_rows), _f32(y_rows)
    grads = np.empty_like(params)
    z = np.empty((x_rows.shape[0], dim), np.float32)
    loss = ctypes.c_float()
    L.check(L.lib().umapx_mlp_batch_grad(d, h, dim, L.ptr(params), L.ptr(mean), L.ptr(inv_std),
                                         L.ptr(x_rows), L.ptr(y_rows), n_edges, n_neg, a, b, lam,
                                         L.ptr(grads), ctypes.pointer(loss), L.ptr(z)))
    return loss.value, grads, z


def mlp_forward(params, mean, inv_std, x, d, h, dim, fused):
    params, mean, inv_std, x = _f32(params), _f32(mean), _f32(inv_std), _f32(x)
    z = np.empty((x.shape[0], dim), np.float32)
    L.check(L.lib().umapx_mlp_forward(d, h, dim, L.ptr(params), L.ptr(mean), L.ptr(inv_std),
                                      L.ptr(x), x.shape[0], 1 if fused else 0, L.ptr(z)))
    return z


def inverse_init(x_train, y_train, q, k, a, b):
    x_train, y_train, q = _f32(x_train), _f32(y_train), _f32(q)
    n, d = x_train.shape
    m, dim = q.shape
    out = np.empty((m, d), np.float32)
    L.check(L.lib().umapx_inverse_init(L.ptr(x_train), n, d, L.ptr(y_train), dim, k, a, b,
                                       L.ptr(q), m, L.ptr(out)))
    return out


METRIC_CODES = {"euclidean": 0, "cosine": 1, "manhattan": 2, "chebyshev": 3, "minkowski": 4,
                "correlation": 5, "canberra": 6, "braycurtis": 7}


def knn_metric(q, r, k, metric, p=2.0):
    q, r = _f32(q), _f32(r)
    m, d = q.shape
    idx = np.empty((m, k), np.int64)
    dist = np.empty((m, k), np.float32)
    L.check(L.lib().umapx_knn_metric(L.ptr(q), m, L.ptr(r), r.shape[0], d, k,
                                     METRIC_CODES[metric], p, L.ptr(idx, L._i64p), L.ptr(dist)))
    return idx, dist


def knn_approx(x, k, metric="euclidean", candidate_pool=50, n_iters=14,
               seed=42 ^ 0xC0FEFEED1234ABCD, p=2.0):
    x = _f32(x)
    n, d = x.shape
    idx = np.empty((n, k), np.int64)
    dist = np.empty((n, k), np.float32)
    L.check(L.lib().umapx_knn_approx(L.ptr(x), n, d, k, METRIC_CODES[metric], p, candidate_pool,
                                     n_iters, seed & 0xFFFFFFFFFFFFFFFF, L.ptr(idx, L._i64p),
                                     L.ptr(dist)))
    return idx, dist


def _csr_parts(m):
    m = m.tocsr()
    m.sort_indices()
    return (np.ascontiguousarray(m.indptr, np.int64), np.ascontiguousarray(m.indices, np.int64),
            np.ascontiguousarray(m.data, np.float32), m.shape)


def knn_csr(x, k, metric=0, p=2.0):
    """umapx_knn_csr: the fit's sparse kNN (row i at 0.0 from itself)."""
    ip, ix, dt, (n, c) = _csr_parts(x)
    idx = np.empty((n, k), np.int64)
    dist = np.empty((n, k), np.float32)
    L.check(L.lib().umapx_knn_csr(n, c, L.ptr(ip, L._i64p), L.ptr(ix, L._i64p), L.ptr(dt),
                                  dt.shape[0], k, metric, p, L.ptr(idx, L._i64p), L.ptr(dist)))
    return idx, dist


def knn_csr_query(q, r, k, metric=0, p=2.0):
    qi, qx, qd, (m, c) = _csr_parts(q)
    ri, rx, rd, (n, _) = _csr_parts(r)
    idx = np.empty((m, k), np.int64)
    dist = np.empty((m, k), np.float32)
    L.check(L.lib().umapx_knn_csr_query(m, L.ptr(qi, L._i64p), L.ptr(qx, L._i64p), L.ptr(qd),
                                        qd.shape[0], n, L.ptr(ri, L._i64p), L.ptr(rx, L._i64p),
                                        L.ptr(rd), rd.shape[0], c, k, metric, p,
                                        L.ptr(idx, L._i64p), L.ptr(dist)))
    return idx, dist


def knn_dense_to_csr(q, r, k, metric=0, p=2.0):
    q = _f32(q)
    ri, rx, rd, (n, c) = _csr_parts(r)
    m = q.shape[0]
    idx = np.empty((m, k), np.int64)
    dist = np.empty((m, k), np.float32)
    L.check(L.lib().umapx_knn_dense_to_csr(L.ptr(q), m, n, L.ptr(ri, L._i64p), L.ptr(rx, L._i64p),
                                           L.ptr(rd), rd.shape[0], c, k, metric, p,
                                           L.ptr(idx, L._i64p), L.ptr(dist)))
    return idx, dist
