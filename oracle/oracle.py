"""TEST INFRASTRUCTURE ONLY — ctypes front end of the CPU oracle.

The oracle is a C++ restatement of umapers' dense fit / transform path
(rust_umap/src/lib.rs, cited function by function in umap_oracle.cpp) plus an
FP64 brute-force kNN in numpy.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module; the product path never
does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_umap.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("n_neighbors", ctypes.c_int32),
        ("n_components", ctypes.c_int32),
        ("n_epochs", ctypes.c_int32),
        ("metric", ctypes.c_int32),
        ("learning_rate", ctypes.c_float),
        ("min_dist", ctypes.c_float),
        ("spread", ctypes.c_float),
        ("local_connectivity", ctypes.c_float),
        ("set_op_mix_ratio", ctypes.c_float),
        ("repulsion_strength", ctypes.c_float),
        ("negative_sample_rate", ctypes.c_int32),
        ("init", ctypes.c_int32),
        ("random_seed", ctypes.c_uint64),
        ("densmap", ctypes.c_int32),
        ("dens_lambda", ctypes.c_float),
        ("dens_frac", ctypes.c_float),
        ("dens_var_shift", ctypes.c_float),
        ("output_dens", ctypes.c_int32),
        ("n_threads", ctypes.c_int32),
        ("metric_p", ctypes.c_float),
    ]


class OrcTimings(ctypes.Structure):
    _fields_ = [
        (name, ctypes.c_double)
        for name in (
            "validation_sec", "knn_sec", "curve_params_sec", "knn_validate_trim_sec",
            "smooth_knn_sec", "density_original_sec", "membership_sec", "symmetrize_sec",
            "target_intersection_sec", "prune_sec", "init_sec", "optimize_sec",
            "density_embedding_sec", "output_copy_sec", "store_state_sec", "total_sec",
        )
    ]


def build() -> str:
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_find_ab_params.argtypes = [ctypes.c_float, ctypes.c_float, _f32p, _f32p]
        _u64p = ctypes.POINTER(ctypes.c_uint64)
        L.orc_xoshiro256pp.argtypes = [_u64p, ctypes.c_int64, _u64p]
        L.orc_splitmix64.argtypes = [ctypes.c_uint64, ctypes.c_int64, _u64p]
        L.orc_smallrng_u64.argtypes = [ctypes.c_uint64, ctypes.c_int64, _u64p]
        L.orc_smallrng_index.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, _u64p]
        L.orc_knn.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                              ctypes.c_int32, ctypes.c_int32, _i64p, _f32p]
        L.orc_knn_to_reference.argtypes = [_f32p, ctypes.c_int64, _f32p, ctypes.c_int64,
                                           ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32, _i64p, _f32p]
        L.orc_sq_euclidean.argtypes = [_f32p, _f32p, ctypes.c_int64]
        L.orc_sq_euclidean.restype = ctypes.c_float
        L.orc_cosine_distance.argtypes = [_f32p, _f32p, ctypes.c_int64]
        L.orc_cosine_distance.restype = ctypes.c_float
        L.orc_knn_rows_include_self.argtypes = [_i64p, _f32p, ctypes.c_int64, ctypes.c_int32]
        L.orc_knn_rows_include_self.restype = ctypes.c_int32
        L.orc_smooth_knn_dist.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_int32,
                                          _f32p, _f32p]
        L.orc_smooth_knn_dist_fp64.argtypes = L.orc_smooth_knn_dist.argtypes
        L.orc_normalized_laplacian.argtypes = [ctypes.c_int64, _i64p, _i64p, _f32p,
                                               ctypes.c_int64, ctypes.c_int32,
                                               ctypes.POINTER(ctypes.c_double)]
        L.orc_normalized_laplacian.restype = ctypes.c_int32
        L.orc_fuzzy_graph.argtypes = [_i64p, _f32p, _f32p, _f32p, ctypes.c_int64,
                                      ctypes.c_int32, ctypes.c_float, ctypes.c_int32,
                                      ctypes.c_int64, _i64p, _i64p, _f32p]
        L.orc_fuzzy_graph.restype = ctypes.c_int64
        L.orc_connected_components.argtypes = [ctypes.c_int64, _i64p, _i64p, ctypes.c_int64,
                                               _i64p]
        L.orc_connected_components.restype = ctypes.c_int64
        L.orc_initialize_embedding.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int64,
                                               ctypes.c_int32, _i64p, _i64p, _f32p,
                                               ctypes.c_int64, ctypes.c_int32,
                                               ctypes.c_uint64, _f32p]
        L.orc_initialize_embedding.restype = ctypes.c_int32
        L.orc_make_epochs_per_sample.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int32, _f32p]
        L.orc_schedule_counts.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.c_int32, _i32p, _i32p, _i64p, _i64p]
        L.orc_optimize_layout.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int32, _i64p, _i64p,
                                          _f32p, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                          ctypes.c_int32, ctypes.c_float, ctypes.c_uint64,
                                          _f32p, ctypes.c_int32, ctypes.c_float,
                                          ctypes.c_float]
        L.orc_density_start_epoch.argtypes = [ctypes.c_int32, ctypes.c_float]
        L.orc_density_start_epoch.restype = ctypes.c_int32
        L.orc_local_radii.argtypes = [_i64p, _f32p, ctypes.c_int64, ctypes.c_int32,
                                      ctypes.c_float, _f32p]
        L.orc_embedding_radii.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int32, _i64p,
                                          ctypes.c_int32, ctypes.c_float, _f32p]
        L.orc_fit.argtypes = [ctypes.POINTER(OrcParams), _f32p, ctypes.c_int64, ctypes.c_int64,
                              _i64p, _f32p, ctypes.c_int64, _f32p, _f32p, _f32p, _f32p,
                              _f32p, ctypes.POINTER(OrcTimings)]
        L.orc_fit.restype = ctypes.c_int32
        L.orc_transform.argtypes = [ctypes.POINTER(OrcParams), _f32p, ctypes.c_int64,
                                    ctypes.c_int64, _f32p, ctypes.c_float, ctypes.c_float,
                                    _f32p, ctypes.c_int64, _f32p]
        L.orc_transform.restype = ctypes.c_int32
        L.orc_knn_metric.argtypes = [_f32p, ctypes.c_int64, _f32p, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                     ctypes.c_int32, _i64p, _f32p]
        L.orc_knn_metric.restype = None
        L.orc_knn_approx.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_float, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, _i64p,
                                     _f32p]
        L.orc_knn_approx.restype = None
        L.orc_inverse_transform.argtypes = [ctypes.POINTER(OrcParams), _f32p, ctypes.c_int64,
                                            ctypes.c_int64, _f32p, _f32p, _f32p, ctypes.c_float,
                                            ctypes.c_float, _f32p, ctypes.c_int64, ctypes.c_int32,
                                            _f32p]
        L.orc_inverse_transform.restype = ctypes.c_int32
        L.orc_transform_with_knn.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int64, _f32p,
                                             ctypes.c_float, ctypes.c_float, ctypes.c_int64,
                                             _i64p, _f32p, ctypes.c_int32, _f32p]
        L.orc_transform_with_knn.restype = ctypes.c_int32
        L.orc_knn_csr.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, _i64p, _f32p,
                                  ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_int32,
                                  _i64p, _f32p]
        L.orc_knn_csr.restype = None
        L.orc_knn_csr_query.argtypes = [ctypes.c_int64, _i64p, _i64p, _f32p, ctypes.c_int64,
                                        _i64p, _i64p, _f32p, ctypes.c_int64, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_float, ctypes.c_int32, _i64p,
                                        _f32p]
        L.orc_knn_csr_query.restype = None
        L.orc_knn_dense_to_csr.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int64, _i64p, _i64p,
                                           _f32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_float, ctypes.c_int32, _i64p, _f32p]
        L.orc_knn_dense_to_csr.restype = None
        L.orc_csr_check.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p,
                                    ctypes.c_int64, _f32p, ctypes.c_int64, ctypes.c_char_p,
                                    ctypes.c_int64]
        L.orc_csr_check.restype = ctypes.c_int32
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


METRICS = {"euclidean": 0, "cosine": 1, "manhattan": 2, "chebyshev": 3, "minkowski": 4,
           "correlation": 5, "canberra": 6, "braycurtis": 7}


def _u64_out(n):
    out = np.empty(n, np.uint64)
    return out, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def xoshiro256pp(state, n: int):
    st = np.ascontiguousarray(state, dtype=np.uint64)
    out, p = _u64_out(n)
    lib().orc_xoshiro256pp(st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n, p)
    return out


def splitmix64(seed: int, n: int):
    out, p = _u64_out(n)
    lib().orc_splitmix64(seed, n, p)
    return out


def smallrng_u64(seed: int, n: int):
    out, p = _u64_out(n)
    lib().orc_smallrng_u64(seed, n, p)
    return out


def smallrng_index(seed: int, range_: int, n: int):
    out, p = _u64_out(n)
    lib().orc_smallrng_index(seed, range_, n, p)
    return out


def find_ab_params(spread: float, min_dist: float) -> tuple[float, float]:
    a = ctypes.c_float()
    b = ctypes.c_float()
    lib().orc_find_ab_params(spread, min_dist, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def knn(x, k: int, metric: str = "euclidean", n_threads: int = 0):
    x = _f32(x)
    n, d = x.shape
    idx = np.empty((n, k), np.int64)
    dist = np.empty((n, k), np.float32)
    lib().orc_knn(_p(x, _f32p), n, d, k, METRICS[metric], n_threads, _p(idx, _i64p),
                  _p(dist, _f32p))
    return idx, dist


def knn_metric(q, r, k: int, metric: str, p: float = 2.0, n_threads: int = 0):
    """Exact kNN of q among r under any of the reference's metrics
    (lib.rs:2891-2928; self included when q is r)."""
    q, r = _f32(q), _f32(r)
    m, d = q.shape
    idx = np.empty((m, k), np.int64)
    dist = np.empty((m, k), np.float32)
    lib().orc_knn_metric(_p(q, _f32p), m, _p(r, _f32p), r.shape[0], d, k, METRICS[metric], p,
                         n_threads, _p(idx, _i64p), _p(dist, _f32p))
    return idx, dist


def knn_approx(x, k: int, metric: str = "euclidean", candidate_pool: int = 50,
               n_iters: int = 14, seed: int = 42 ^ 0xC0FEFEED1234ABCD, n_threads: int = 0,
               p: float = 2.0):
    """approximate_nearest_neighbors (lib.rs:2932-3168): the reference's
    neighbours-of-neighbours refinement from random pools; the seed is the
    model's random_seed ^ 0xC0FE_FEED_1234_ABCD (lib.rs:2326)."""
    x = _f32(x)
    n, d = x.shape
    idx = np.empty((n, k), np.int64)
    dist = np.empty((n, k), np.float32)
    lib().orc_knn_approx(_p(x, _f32p), n, d, k, METRICS[metric], p, candidate_pool, n_iters,
                         seed & 0xFFFFFFFFFFFFFFFF, n_threads, _p(idx, _i64p), _p(dist, _f32p))
    return idx, dist


def knn_to_reference(q, r, k: int, metric: str = "euclidean", n_threads: int = 0):
    q = _f32(q)
    r = _f32(r)
    m, d = q.shape
    idx = np.empty((m, k), np.int64)
    dist = np.empty((m, k), np.float32)
    lib().orc_knn_to_reference(_p(q, _f32p), m, _p(r, _f32p), r.shape[0], d, k,
                               METRICS[metric], n_threads, _p(idx, _i64p), _p(dist, _f32p))
    return idx, dist


def sq_euclidean(x, y) -> float:
    x = _f32(x)
    y = _f32(y)
    return lib().orc_sq_euclidean(_p(x, _f32p), _p(y, _f32p), x.shape[0])


def cosine_distance(x, y) -> float:
    x = _f32(x)
    y = _f32(y)
    return lib().orc_cosine_distance(_p(x, _f32p), _p(y, _f32p), x.shape[0])


def rows_include_self(idx, dist) -> bool:
    idx = _i64(idx)
    dist = _f32(dist)
    return bool(lib().orc_knn_rows_include_self(_p(idx, _i64p), _p(dist, _f32p),
                                                idx.shape[0], idx.shape[1]))


def smooth_knn_dist(dist, k_value: float, local_connectivity: float = 1.0,
                    include_self: bool = True, bandwidth: float = 1.0, fp64_sums: bool = False):
    """sigma / rho of every row; fp64_sums: the bisection's psum in f64."""
    dist = _f32(dist)
    n, k = dist.shape
    sig = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    fn = lib().orc_smooth_knn_dist_fp64 if fp64_sums else lib().orc_smooth_knn_dist
    fn(_p(dist, _f32p), n, k, k_value, local_connectivity, bandwidth, int(include_self),
       _p(sig, _f32p), _p(rho, _f32p))
    return sig, rho


def fuzzy_graph(idx, dist, sig, rho, mix: float = 1.0, n_epochs: int = 0):
    """Membership + union (+ prune if n_epochs > 0).  Returns (heads, tails, w)."""
    idx = _i64(idx)
    dist = _f32(dist)
    sig = _f32(sig)
    rho = _f32(rho)
    n, k = idx.shape
    cap = 2 * n * k
    h = np.empty(cap, np.int64)
    t = np.empty(cap, np.int64)
    w = np.empty(cap, np.float32)
    ne = lib().orc_fuzzy_graph(_p(idx, _i64p), _p(dist, _f32p), _p(sig, _f32p),
                               _p(rho, _f32p), n, k, mix, n_epochs, cap, _p(h, _i64p),
                               _p(t, _i64p), _p(w, _f32p))
    assert ne >= 0
    return h[:ne].copy(), t[:ne].copy(), w[:ne].copy()


def normalized_laplacian(n: int, heads, tails, w, from_affinity: bool = False):
    """I - D^-1/2 W D^-1/2 of the max-symmetrised graph, from the edge list
    or from the dense affinity matrix; None when no weight is positive."""
    heads, tails, w = _i64(heads), _i64(tails), _f32(w)
    out = np.zeros((n, n), np.float64)
    ok = lib().orc_normalized_laplacian(n, _p(heads, _i64p), _p(tails, _i64p), _p(w, _f32p),
                                        heads.shape[0], int(from_affinity),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out if ok else None


def connected_components(n: int, heads, tails):
    heads = _i64(heads)
    tails = _i64(tails)
    labels = np.empty(n, np.int64)
    nc = lib().orc_connected_components(n, _p(heads, _i64p), _p(tails, _i64p), heads.shape[0],
                                        _p(labels, _i64p))
    return int(nc), labels


def initialize_embedding(x, dim: int, heads, tails, w, init: str = "spectral", seed: int = 42):
    x = _f32(x)
    heads = _i64(heads)
    tails = _i64(tails)
    w = _f32(w)
    n, d = x.shape
    y = np.empty((n, dim), np.float32)
    ok = lib().orc_initialize_embedding(_p(x, _f32p), n, d, dim, _p(heads, _i64p),
                                        _p(tails, _i64p), _p(w, _f32p), heads.shape[0],
                                        1 if init == "spectral" else 0, seed, _p(y, _f32p))
    return y, bool(ok)


def epochs_per_sample(w, n_epochs: int):
    w = _f32(w)
    out = np.empty_like(w)
    lib().orc_make_epochs_per_sample(_p(w, _f32p), w.shape[0], n_epochs, _p(out, _f32p))
    return out


def schedule_counts(w, n_epochs: int, negative_sample_rate: int = 5):
    w = _f32(w)
    ne = w.shape[0]
    pos = np.empty(ne, np.int32)
    neg = np.empty(ne, np.int32)
    act = np.empty(n_epochs, np.int64)
    negs = np.empty(n_epochs, np.int64)
    lib().orc_schedule_counts(_p(w, _f32p), ne, n_epochs, negative_sample_rate,
                              _p(pos, _i32p), _p(neg, _i32p), _p(act, _i64p), _p(negs, _i64p))
    return pos, neg, act, negs


def optimize_layout(y, heads, tails, w, n_epochs: int, a: float, b: float, *,
                    learning_rate=1.0, negative_sample_rate=5, repulsion_strength=1.0,
                    seed=42 ^ 0x9E3779B97F4A7C15, radii=None, dens_start=0,
                    dens_lambda=2.0, dens_var_shift=0.1):
    y = np.array(y, dtype=np.float32, order="C", copy=True)
    heads = _i64(heads)
    tails = _i64(tails)
    w = _f32(w)
    rad = None if radii is None else _f32(radii)
    lib().orc_optimize_layout(_p(y, _f32p), y.shape[0], y.shape[1], _p(heads, _i64p),
                              _p(tails, _i64p), _p(w, _f32p), heads.shape[0], n_epochs, a, b,
                              learning_rate, negative_sample_rate, repulsion_strength,
                              seed & 0xFFFFFFFFFFFFFFFF,
                              None if rad is None else _p(rad, _f32p), dens_start,
                              dens_lambda, dens_var_shift)
    return y


def density_start_epoch(n_epochs: int, dens_frac: float) -> int:
    return int(lib().orc_density_start_epoch(n_epochs, dens_frac))


def local_radii(idx, dist, var_shift: float = 0.1):
    idx = _i64(idx)
    dist = _f32(dist)
    out = np.empty(idx.shape[0], np.float32)
    lib().orc_local_radii(_p(idx, _i64p), _p(dist, _f32p), idx.shape[0], idx.shape[1],
                          var_shift, _p(out, _f32p))
    return out


def embedding_radii(y, idx, var_shift: float = 0.1):
    y = _f32(y)
    idx = _i64(idx)
    out = np.empty(y.shape[0], np.float32)
    lib().orc_embedding_radii(_p(y, _f32p), y.shape[0], y.shape[1], _p(idx, _i64p),
                              idx.shape[1], var_shift, _p(out, _f32p))
    return out


@dataclass
class OracleParams:
    """The reference's UmapParams defaults (rust_umap/src/lib.rs:164-195)."""

    n_neighbors: int = 15
    n_components: int = 2
    n_epochs: int | None = None
    metric: str = "euclidean"
    learning_rate: float = 1.0
    min_dist: float = 0.1
    spread: float = 1.0
    local_connectivity: float = 1.0
    set_op_mix_ratio: float = 1.0
    repulsion_strength: float = 1.0
    negative_sample_rate: int = 5
    init: str = "spectral"
    random_seed: int = 42
    densmap: bool = False
    dens_lambda: float = 2.0
    dens_frac: float = 0.3
    dens_var_shift: float = 0.1
    output_dens: bool = False
    n_threads: int = 0
    metric_p: float = 2.0

    def c(self) -> OrcParams:
        return OrcParams(
            self.n_neighbors, self.n_components, 0 if self.n_epochs is None else self.n_epochs,
            METRICS[self.metric], self.learning_rate, self.min_dist, self.spread,
            self.local_connectivity, self.set_op_mix_ratio, self.repulsion_strength,
            self.negative_sample_rate, 1 if self.init == "spectral" else 0,
            self.random_seed & 0xFFFFFFFFFFFFFFFF, int(self.densmap), self.dens_lambda,
            self.dens_frac, self.dens_var_shift, int(self.output_dens), self.n_threads,
            self.metric_p)


@dataclass
class OracleFit:
    embedding: np.ndarray
    sigmas: np.ndarray
    rhos: np.ndarray
    rad_orig: np.ndarray | None
    rad_emb: np.ndarray | None
    timings: dict = field(default_factory=dict)


def fit(x, params: OracleParams | None = None, knn_idx=None, knn_dist=None) -> OracleFit:
    params = params or OracleParams()
    x = _f32(x)
    n, d = x.shape
    cp = params.c()
    y = np.empty((n, params.n_components), np.float32)
    sig = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    dens = params.densmap or params.output_dens
    ro = np.empty(n, np.float32) if dens else None
    re = np.empty(n, np.float32) if dens else None
    tm = OrcTimings()
    if knn_idx is not None:
        ki = _i64(knn_idx)
        kd = _f32(knn_dist)
        kp, dp, kc = _p(ki, _i64p), _p(kd, _f32p), ki.shape[1]
    else:
        kp, dp, kc = None, None, 0
    rc = lib().orc_fit(ctypes.byref(cp), _p(x, _f32p), n, d, kp, dp, kc, _p(y, _f32p),
                       _p(sig, _f32p), _p(rho, _f32p),
                       None if ro is None else _p(ro, _f32p),
                       None if re is None else _p(re, _f32p), ctypes.byref(tm))
    assert rc == 0
    timings = {name: getattr(tm, name) for name, _ in OrcTimings._fields_}
    return OracleFit(y, sig, rho, ro, re, timings)


def transform(x_train, y_train, q, params: OracleParams | None = None):
    params = params or OracleParams()
    x_train = _f32(x_train)
    y_train = _f32(y_train)
    q = _f32(q)
    a, b = find_ab_params(params.spread, params.min_dist)
    out = np.empty((q.shape[0], params.n_components), np.float32)
    cp = params.c()
    rc = lib().orc_transform(ctypes.byref(cp), _p(x_train, _f32p), x_train.shape[0],
                             x_train.shape[1], _p(y_train, _f32p), a, b, _p(q, _f32p),
                             q.shape[0], _p(out, _f32p))
    assert rc == 0
    return out


class Csr:
    """CSR rows as the reference's SparseCsrMatrix holds them: int64 indptr
    [rows + 1] and column indices, f32 data (from a scipy matrix or arrays)."""

    def __init__(self, indptr, indices, data, n_cols: int):
        self.indptr = np.ascontiguousarray(indptr, dtype=np.int64)
        self.indices = np.ascontiguousarray(indices, dtype=np.int64)
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        self.n_rows = self.indptr.shape[0] - 1
        self.n_cols = int(n_cols)

    @classmethod
    def of(cls, m) -> "Csr":
        m = m.tocsr()
        return cls(m.indptr, m.indices, m.data, m.shape[1])

    def ptrs(self):
        return _p(self.indptr, _i64p), _p(self.indices, _i64p), _p(self.data, _f32p)


def csr_check(c: Csr) -> str | None:
    """SparseCsrMatrix::new's checks (sparse.rs:20-90): None when valid, else
    the reference's message."""
    buf = ctypes.create_string_buffer(512)
    rc = lib().orc_csr_check(c.n_rows, c.n_cols, _p(c.indptr, _i64p), c.indptr.shape[0],
                             _p(c.indices, _i64p), c.indices.shape[0], _p(c.data, _f32p),
                             c.data.shape[0], buf, 512)
    return None if rc == 0 else buf.value.decode()


def knn_csr(c, k: int, metric: str = "euclidean", p: float = 2.0, n_threads: int = 0):
    """exact_nearest_neighbors (sparse.rs:605-626): the fit's sparse kNN,
    every row itself at distance 0.0."""
    c = c if isinstance(c, Csr) else Csr.of(c)
    idx = np.empty((c.n_rows, k), np.int64)
    dist = np.empty((c.n_rows, k), np.float32)
    lib().orc_knn_csr(c.n_rows, c.n_cols, *c.ptrs(), k, METRICS[metric], p, n_threads,
                      _p(idx, _i64p), _p(dist, _f32p))
    return idx, dist


def knn_csr_query(q, r, k: int, metric: str = "euclidean", p: float = 2.0, n_threads: int = 0):
    """exact_nearest_neighbors_sparse_query (sparse.rs:628-657)."""
    q = q if isinstance(q, Csr) else Csr.of(q)
    r = r if isinstance(r, Csr) else Csr.of(r)
    idx = np.empty((q.n_rows, k), np.int64)
    dist = np.empty((q.n_rows, k), np.float32)
    lib().orc_knn_csr_query(q.n_rows, *q.ptrs(), r.n_rows, *r.ptrs(), r.n_cols, k,
                            METRICS[metric], p, n_threads, _p(idx, _i64p), _p(dist, _f32p))
    return idx, dist


def knn_dense_to_csr(q, r, k: int, metric: str = "euclidean", p: float = 2.0,
                     n_threads: int = 0):
    """exact_nearest_neighbors_dense_query (sparse.rs:753-900): dense rows
    against CSR rows."""
    q = _f32(q)
    r = r if isinstance(r, Csr) else Csr.of(r)
    idx = np.empty((q.shape[0], k), np.int64)
    dist = np.empty((q.shape[0], k), np.float32)
    lib().orc_knn_dense_to_csr(_p(q, _f32p), q.shape[0], r.n_rows, *r.ptrs(), r.n_cols, k,
                               METRICS[metric], p, n_threads, _p(idx, _i64p), _p(dist, _f32p))
    return idx, dist


def transform_with_knn(y_train, idx, dist, params: OracleParams | None = None):
    """transform_from_knn (lib.rs:1738-1780) on a query kNN among the fitted
    rows (how the sparse transforms finish, lib.rs:1655-1737)."""
    params = params or OracleParams()
    y_train = _f32(y_train)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    dist = _f32(dist)
    a, b = find_ab_params(params.spread, params.min_dist)
    out = np.empty((idx.shape[0], params.n_components), np.float32)
    cp = params.c()
    rc = lib().orc_transform_with_knn(ctypes.byref(cp), y_train.shape[0], _p(y_train, _f32p), a,
                                      b, idx.shape[0], _p(idx, _i64p), _p(dist, _f32p),
                                      idx.shape[1], _p(out, _f32p))
    assert rc == 0
    return out


def inverse_transform(x_train, y_train, sigmas, rhos, q, params: OracleParams | None = None,
                      optimize: bool = True):
    """inverse_transform_rows (lib.rs:1800-1915): embedded rows q back to the
    feature space of x_train, given the fitted embedding and the fit's sigma /
    rho.  optimize=False returns the local linear initialisation only."""
    params = params or OracleParams()
    x_train, y_train, q = _f32(x_train), _f32(y_train), _f32(q)
    sigmas, rhos = _f32(sigmas), _f32(rhos)
    a, b = find_ab_params(params.spread, params.min_dist)
    out = np.empty((q.shape[0], x_train.shape[1]), np.float32)
    cp = params.c()
    rc = lib().orc_inverse_transform(ctypes.byref(cp), _p(x_train, _f32p), x_train.shape[0],
                                     x_train.shape[1], _p(y_train, _f32p), _p(sigmas, _f32p),
                                     _p(rhos, _f32p), a, b, _p(q, _f32p), q.shape[0],
                                     1 if optimize else 0, _p(out, _f32p))
    assert rc == 0
    return out
