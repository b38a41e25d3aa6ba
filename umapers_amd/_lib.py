"""ctypes binding of libumapx.so (the C ABI declared in include/umapx.h).

This plays the role of the reference's PyO3 module (umap_rs/src/lib.rs): it
marshals numpy arrays into plain pointers and maps error codes to Python
exceptions the way umap_rs/src/lib.rs:147-157 maps UmapError.  The library is
loaded from the package directory; a missing build raises instead of falling
back to anything else.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libumapx.so")

OK = 0
ERR_EMPTY_DATA = 1
ERR_NEED_TWO_SAMPLES = 2
ERR_INCONSISTENT_DIMENSIONS = 3
ERR_FEATURE_MISMATCH = 4
ERR_EMBEDDING_DIMENSION_MISMATCH = 5
ERR_NOT_FITTED = 6
ERR_INVALID_PARAMETER = 7
ERR_DEVICE = 8

METRIC_EUCLIDEAN = 0
METRIC_COSINE = 1
METRIC_MANHATTAN = 2
METRIC_CHEBYSHEV = 3
METRIC_MINKOWSKI = 4
METRIC_CORRELATION = 5
METRIC_CANBERRA = 6
METRIC_BRAYCURTIS = 7
INIT_RANDOM = 0
INIT_SPECTRAL = 1
KNN_FP32 = 0
KNN_BF16_RERANK = 1
TRANSFORM_PLAIN = 0
TRANSFORM_LEARNED = 1

_f32p = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


class Params(ctypes.Structure):
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
        ("use_approximate_knn", ctypes.c_int32),
        ("approx_knn_candidates", ctypes.c_int32),
        ("approx_knn_iters", ctypes.c_int32),
        ("approx_knn_threshold", ctypes.c_int32),
        ("densmap", ctypes.c_int32),
        ("dens_lambda", ctypes.c_float),
        ("dens_frac", ctypes.c_float),
        ("dens_var_shift", ctypes.c_float),
        ("output_dens", ctypes.c_int32),
        ("knn_precision", ctypes.c_int32),
        ("rerank_factor", ctypes.c_int32),
        ("smooth_fp64", ctypes.c_int32),
        ("transform_mode", ctypes.c_int32),
        ("mlp_hidden", ctypes.c_int32),
        ("mlp_steps", ctypes.c_int32),
        ("mlp_batch", ctypes.c_int32),
        ("mlp_negatives", ctypes.c_int32),
        ("mlp_learning_rate", ctypes.c_float),
        ("mlp_embedding_weight", ctypes.c_float),
        ("metric_p", ctypes.c_float),
    ]


TIMING_KEYS = (
    "validation_sec", "knn_sec", "curve_params_sec", "knn_validate_trim_sec", "smooth_knn_sec",
    "density_original_sec", "membership_sec", "symmetrize_sec", "target_intersection_sec",
    "prune_sec", "init_sec", "optimize_sec", "density_embedding_sec", "output_copy_sec",
    "store_state_sec", "total_sec",
)


class Timings(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in TIMING_KEYS] + [
        ("knn_kernel_ms", ctypes.c_double),
        ("knn_flops", ctypes.c_double),
        ("sgd_kernel_ms", ctypes.c_double),
        ("n_edges", ctypes.c_int64),
        ("n_epochs", ctypes.c_int32),
        ("used_approximate_knn", ctypes.c_int32),
        ("init_used_spectral", ctypes.c_int32),
        ("n_graph_components", ctypes.c_int32),
        ("sgd_positive_updates", ctypes.c_int64),
        ("sgd_negative_samples", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# (name, restype, argtypes) for every symbol include/umapx.h declares.
SIGNATURES = {
    "umapx_default_params": (None, [ctypes.POINTER(Params)]),
    "umapx_last_error": (ctypes.c_char_p, []),
    "umapx_last_error_detail": (ctypes.c_int32, [_i64p, _i64p, _i64p]),
    "umapx_uses_approximate_knn": (ctypes.c_int32, [ctypes.POINTER(Params), ctypes.c_int64,
                                                    ctypes.c_int64]),
    "umapx_version": (ctypes.c_char_p, []),
    "umapx_device_count": (ctypes.c_int32, []),
    "umapx_create": (ctypes.c_int32, [ctypes.POINTER(Params), ctypes.POINTER(_vp)]),
    "umapx_destroy": (None, [_vp]),
    "umapx_fit": (ctypes.c_int32, [_vp, _f32p, ctypes.c_int64, ctypes.c_int64]),
    "umapx_fit_transform": (ctypes.c_int32, [_vp, _f32p, ctypes.c_int64, ctypes.c_int64, _f32p]),
    "umapx_fit_transform_device": (ctypes.c_int32, [_vp, _vp, ctypes.c_int64, ctypes.c_int64, _vp]),
    "umapx_fit_transform_with_knn": (ctypes.c_int32, [
        _vp, _f32p, ctypes.c_int64, ctypes.c_int64, _i64p, ctypes.c_int64, ctypes.c_int64,
        _f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _f32p]),
    "umapx_profile_fit_transform": (ctypes.c_int32, [
        _vp, _f32p, ctypes.c_int64, ctypes.c_int64, _f32p, ctypes.POINTER(Timings)]),
    "umapx_get_timings": (ctypes.c_int32, [_vp, ctypes.POINTER(Timings)]),
    "umapx_fit_transform_csr": (ctypes.c_int32, [_vp, ctypes.c_int64, ctypes.c_int64, _i64p,
                                                 _i64p, _f32p, ctypes.c_int64, _f32p]),
    "umapx_transform_csr": (ctypes.c_int32, [_vp, ctypes.c_int64, ctypes.c_int64, _i64p, _i64p,
                                             _f32p, ctypes.c_int64, _f32p]),
    "umapx_is_sparse_fit": (ctypes.c_int32, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "umapx_knn_csr": (ctypes.c_int32, [ctypes.c_int64, ctypes.c_int64, _i64p, _i64p, _f32p,
                                       ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_float, _i64p, _f32p]),
    "umapx_knn_csr_query": (ctypes.c_int32, [ctypes.c_int64, _i64p, _i64p, _f32p, ctypes.c_int64,
                                             ctypes.c_int64, _i64p, _i64p, _f32p, ctypes.c_int64,
                                             ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_float, _i64p, _f32p]),
    "umapx_knn_dense_to_csr": (ctypes.c_int32, [_f32p, ctypes.c_int64, ctypes.c_int64, _i64p,
                                                _i64p, _f32p, ctypes.c_int64, ctypes.c_int64,
                                                ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                                _i64p, _f32p]),
    "umapx_transform": (ctypes.c_int32, [_vp, _f32p, ctypes.c_int64, ctypes.c_int64, _f32p]),
    "umapx_transform_device": (ctypes.c_int32, [_vp, _vp, ctypes.c_int64, ctypes.c_int64, _vp]),
    "umapx_transform_plain": (ctypes.c_int32, [_vp, _f32p, ctypes.c_int64, ctypes.c_int64, _f32p]),
    "umapx_transform_learned": (ctypes.c_int32, [_vp, _f32p, ctypes.c_int64, ctypes.c_int64, _f32p]),
    "umapx_transform_learned_device": (ctypes.c_int32, [
        _vp, _vp, ctypes.c_int64, ctypes.c_int64, _vp]),
    "umapx_inverse_transform": (ctypes.c_int32, [_vp, _f32p, ctypes.c_int64, ctypes.c_int64,
                                                 _f32p]),
    "umapx_inverse_transform_device": (ctypes.c_int32, [_vp, _vp, ctypes.c_int64, ctypes.c_int64,
                                                        _vp]),
    "umapx_inverse_init": (ctypes.c_int32, [_f32p, ctypes.c_int64, ctypes.c_int64, _f32p,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                            ctypes.c_float, _f32p, ctypes.c_int64, _f32p]),
    "umapx_knn_approx": (ctypes.c_int32, [_f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_float, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_uint64, _i64p, _f32p]),
    "umapx_knn_metric": (ctypes.c_int32, [_f32p, ctypes.c_int64, _f32p, ctypes.c_int64,
                                          ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_float, _i64p, _f32p]),
    "umapx_n_samples": (ctypes.c_int32, [_vp, _i64p]),
    "umapx_n_features": (ctypes.c_int32, [_vp, _i64p]),
    "umapx_get_embedding": (ctypes.c_int32, [_vp, _f32p]),
    "umapx_get_ab": (ctypes.c_int32, [_vp, _f32p, _f32p]),
    "umapx_get_radii": (ctypes.c_int32, [_vp, _f32p, _f32p]),
    "umapx_find_ab_params": (ctypes.c_int32, [ctypes.c_float, ctypes.c_float, _f32p, _f32p]),
    "umapx_mlp_param_count": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.POINTER(ctypes.c_int64)]),
    "umapx_mlp_batch_grad": (ctypes.c_int32, [
        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _f32p, _f32p, _f32p, _f32p, _f32p,
        ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_float, _f32p,
        _f32p, _f32p]),
    "umapx_get_learned_kernel_ms": (ctypes.c_int32, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "umapx_mlp_forward": (ctypes.c_int32, [
        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _f32p, _f32p, _f32p, _f32p, ctypes.c_int64,
        ctypes.c_int32, _f32p]),
    "umapx_get_mlp_loss": (ctypes.c_int32, [_vp, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_int32)]),
    "umapx_knn": (ctypes.c_int32, [_f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _i64p, _f32p]),
    "umapx_knn_to_reference": (ctypes.c_int32, [
        _f32p, ctypes.c_int64, _f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _i64p, _f32p]),
    "umapx_knn_to_reference_device_noalloc": (ctypes.c_int32, [
        _vp, ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp]),
    "umapx_knn_rows_include_self": (ctypes.c_int32, [
        _i64p, _f32p, ctypes.c_int64, ctypes.c_int32, _i32p]),
    "umapx_smooth_knn_dist": (ctypes.c_int32, [
        _f32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
        ctypes.c_int32, ctypes.c_int32, _f32p, _f32p]),
    "umapx_fuzzy_graph": (ctypes.c_int32, [
        _i64p, _f32p, _f32p, _f32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float,
        ctypes.c_int32, ctypes.c_int64, _i64p, _i64p, _f32p, _i64p]),
    "umapx_initialize_embedding": (ctypes.c_int32, [
        _f32p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, _i64p, _i64p, _f32p,
        ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64, _f32p, _i32p]),
    "umapx_connected_components": (ctypes.c_int32, [
        ctypes.c_int64, _i64p, _i64p, _f32p, ctypes.c_int64, _i64p, _i64p]),
    "umapx_optimize_layout": (ctypes.c_int32, [
        _f32p, ctypes.c_int64, ctypes.c_int32, _i64p, _i64p, _f32p, ctypes.c_int64,
        ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int32,
        ctypes.c_float, ctypes.c_uint64, _f32p, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
        _i32p, _i32p]),
    "umapx_local_radii": (ctypes.c_int32, [
        _i64p, _f32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float, _f32p]),
    "umapx_embedding_radii": (ctypes.c_int32, [
        _f32p, ctypes.c_int64, ctypes.c_int32, _i64p, ctypes.c_int32, ctypes.c_float, _f32p]),
}


class UmapxNotBuilt(RuntimeError):
    pass


_lib = None


def lib() -> ctypes.CDLL:
    """Load libumapx.so once; raise when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UmapxNotBuilt(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().umapx_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    """Raise the Python exception that umap_rs maps each error to."""
    if rc == OK:
        return
    msg = last_error()
    if rc in (ERR_NOT_FITTED, ERR_DEVICE):
        raise RuntimeError(msg)
    raise ValueError(msg)


def default_params() -> Params:
    p = Params()
    lib().umapx_default_params(ctypes.byref(p))
    return p


def ptr(a: np.ndarray, t=_f32p):
    return a.ctypes.data_as(t)


def device_count() -> int:
    return int(lib().umapx_device_count())
