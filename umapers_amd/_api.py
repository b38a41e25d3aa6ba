"""Python facade with the umapers API (umap_rs/python/umapers/_api.py).

``Umap`` keeps the reference's constructor keywords, defaults and methods
(``fit``, ``fit_transform``, ``fit_transform_with_knn``, ``transform``,
``profile_fit_transform``, ``get_params`` / ``set_params``) and its fitted
attributes.  Underneath, ``_Core`` plays the part of the PyO3 ``UmapCore``
(umap_rs/src/lib.rs:540-1320): it owns a ``umapx_handle`` and calls the C ABI,
which runs the fit on the GPU.  There is no CPU fallback: a missing library
or GPU raises.

GPU-path keywords that the reference does not have:
  knn_precision   "fp32" (default) or "bf16_rerank" (BF16 MFMA candidates,
                  exact f32 re-rank of rerank_factor * k candidates)
  rerank_factor   candidates per row over k (default 2)
  transform_mode  "plain" (kNN + SGD, the reference's transform) or "learned"
                  (MLP encoder trained on the fitted graph)
  smooth_fp64     FP64 sums in the sigma bisection (default True)
  mlp_*           learned-transform training knobs
"""
from __future__ import annotations

import ctypes
from typing import Any

import numpy as np

from . import _lib as L

_PARAM_NAMES = (
    "n_neighbors", "n_components", "n_epochs", "metric", "metric_kwds", "learning_rate",
    "min_dist", "spread", "local_connectivity", "set_op_mix_ratio", "repulsion_strength",
    "negative_sample_rate", "random_seed", "init", "ann_mode", "use_approximate_knn",
    "approx_knn_candidates", "approx_knn_iters", "approx_knn_threshold", "densmap",
    "dens_lambda", "dens_frac", "dens_var_shift", "output_dens", "target_metric",
    "target_weight", "target_n_neighbors",
    # GPU path
    "knn_precision", "rerank_factor", "transform_mode", "smooth_fp64", "mlp_hidden",
    "mlp_steps", "mlp_batch", "mlp_negatives", "mlp_learning_rate", "mlp_embedding_weight",
)

_GPU_METRICS = {"euclidean": L.METRIC_EUCLIDEAN, "cosine": L.METRIC_COSINE,
                "manhattan": L.METRIC_MANHATTAN, "chebyshev": L.METRIC_CHEBYSHEV,
                "minkowski": L.METRIC_MINKOWSKI, "correlation": L.METRIC_CORRELATION,
                "canberra": L.METRIC_CANBERRA, "braycurtis": L.METRIC_BRAYCURTIS}


def _as_f32_matrix(x: Any, name: str) -> np.ndarray:
    """_api.py:113-117: C-contiguous float32, 2-D."""
    arr = np.asarray(x, dtype=np.float32, order="C")
    if arr.ndim != 2:
        raise ValueError(f"{name} must be a 2D array, got ndim={arr.ndim}")
    return arr


def _as_knn_indices(x: Any, name: str) -> np.ndarray:
    """_api.py:120-124: C-contiguous int64, 2-D."""
    arr = np.asarray(x, dtype=np.int64, order="C")
    if arr.ndim != 2:
        raise ValueError(f"{name} must be a 2D array, got ndim={arr.ndim}")
    return arr


def _is_sparse(x: Any) -> bool:
    return getattr(x, "format", None) in {"csr", "csc", "coo"}


class _CsrParts:
    """Canonical CSR arrays of a sparse input: int64 indptr and column indices,
    float32 data, with the matrix's shape."""

    def __init__(self, indptr, indices, data, n_rows: int, n_cols: int):
        self.indptr, self.indices, self.data = indptr, indices, data
        self.shape = (n_rows, n_cols)


def _as_csr_parts(x: Any, name: str) -> _CsrParts:
    """_api.py:143-190: CSR / CSC / COO normalised to canonical CSR (summed
    duplicates, sorted indices); the rows stay sparse all the way to the GPU."""
    fmt = getattr(x, "format", None)
    matrix = x
    if fmt in {"csc", "coo"}:
        tocsr = getattr(x, "tocsr", None)
        if not callable(tocsr):
            raise ValueError(f"{name} sparse input with format={fmt!r} cannot be converted to CSR")
        matrix = tocsr()
        if getattr(matrix, "format", None) != "csr":
            raise ValueError(f"{name}.tocsr() must return a CSR matrix")
    if getattr(matrix, "has_canonical_format", True) is False:
        matrix = matrix.copy()
        matrix.sum_duplicates()
        matrix.sort_indices()
    elif getattr(matrix, "has_sorted_indices", True) is False:
        matrix = matrix.copy()
        matrix.sort_indices()
    shape = getattr(matrix, "shape", None)
    if shape is None or len(shape) != 2:
        raise ValueError(f"{name} must be a 2D CSR matrix")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    if n_cols <= 0:
        raise ValueError(f"{name} must have at least one column")
    indptr = np.ascontiguousarray(matrix.indptr, dtype=np.int64)
    indices = np.ascontiguousarray(matrix.indices, dtype=np.int64)
    data = np.ascontiguousarray(matrix.data, dtype=np.float32)
    if indptr.ndim != 1 or indices.ndim != 1 or data.ndim != 1:
        raise ValueError(f"{name} CSR arrays must be 1D")
    if indices.shape[0] != data.shape[0]:
        raise ValueError(f"{name} CSR indices/data length mismatch")
    return _CsrParts(indptr, indices, data, n_rows, n_cols)


def _as_out_buffer(out: Any, shape: tuple[int, int]) -> np.ndarray:
    """_api.py:193-206."""
    if not isinstance(out, np.ndarray):
        raise TypeError("out must be a NumPy ndarray")
    if out.dtype != np.float32:
        raise TypeError(f"out dtype must be float32, got {out.dtype}")
    if out.ndim != 2:
        raise ValueError(f"out must be 2D, got ndim={out.ndim}")
    if not out.flags.c_contiguous:
        raise ValueError("out must be C-contiguous")
    if not out.flags.writeable:
        raise ValueError("out must be writeable")
    if tuple(out.shape) != tuple(shape):
        raise ValueError(f"output buffer shape mismatch: expected {shape}, got {tuple(out.shape)}")
    return out


def _normalize_ann_mode(ann_mode: Any, use_approximate_knn: bool,
                        approx_knn_threshold: int) -> tuple[str, bool, int]:
    """_api.py:209-224."""
    mode = ann_mode if isinstance(ann_mode, str) else str(ann_mode)
    mode = mode.lower()
    if mode == "auto":
        return mode, use_approximate_knn, approx_knn_threshold
    if mode == "exact":
        return mode, False, approx_knn_threshold
    if mode == "approximate":
        return mode, True, 0
    raise ValueError(f"unsupported ann_mode '{ann_mode}', expected auto|exact|approximate")


_METRIC_ALIASES = {
    "euclidean": "euclidean", "l2": "euclidean", "manhattan": "manhattan", "l1": "manhattan",
    "cosine": "cosine", "chebyshev": "chebyshev", "linfinity": "chebyshev",
    "linf": "chebyshev", "minkowski": "minkowski", "correlation": "correlation",
    "canberra": "canberra", "braycurtis": "braycurtis", "bray_curtis": "braycurtis",
}


def _normalize_metric_name(metric: Any) -> str:
    """_api.py:242-263."""
    name = str(metric).lower()
    if name not in _METRIC_ALIASES:
        expected = "|".join(sorted(set(_METRIC_ALIASES.values())))
        raise ValueError(f"unsupported metric '{metric}', expected {expected}")
    return _METRIC_ALIASES[name]


def _parse_metric_config(metric: Any, metric_kwds: Any) -> tuple[str, Any, float | None]:
    """_api.py:266-289: minkowski takes exactly metric_kwds={'p': p}, p finite
    and > 0; every other metric takes no keywords."""
    name = _normalize_metric_name(metric)
    kwds = {} if metric_kwds is None else metric_kwds
    if not hasattr(kwds, "keys"):
        raise TypeError("metric_kwds must be a mapping or None")
    keys = set(kwds.keys())
    if name == "minkowski":
        if "p" not in keys:
            raise ValueError("metric='minkowski' requires metric_kwds={'p': ...}")
        unsupported = keys - {"p"}
        if unsupported:
            bad = ", ".join(sorted(str(key) for key in unsupported))
            raise ValueError(f"unsupported metric_kwds for minkowski: {bad}")
        p = float(kwds["p"])
        if not np.isfinite(p) or p <= 0.0:
            raise ValueError("metric_kwds['p'] must be finite and > 0 for minkowski")
        return name, metric_kwds, p
    if keys:
        bad = ", ".join(sorted(str(key) for key in keys))
        raise ValueError(f"unsupported metric_kwds for {name}: {bad}")
    return name, metric_kwds, None


def _validate_dens_params(dens_lambda: float, dens_frac: float,
                          dens_var_shift: float) -> tuple[float, float, float]:
    """_api.py:302-316."""
    dens_lambda, dens_frac, dens_var_shift = float(dens_lambda), float(dens_frac), float(dens_var_shift)
    if not np.isfinite(dens_lambda) or dens_lambda < 0.0:
        raise ValueError("dens_lambda must be finite and >= 0")
    if not np.isfinite(dens_frac) or dens_frac < 0.0 or dens_frac > 1.0:
        raise ValueError("dens_frac must be finite and in [0, 1]")
    if not np.isfinite(dens_var_shift) or dens_var_shift <= 0.0:
        raise ValueError("dens_var_shift must be finite and > 0")
    return dens_lambda, dens_frac, dens_var_shift


class _Core:
    """Owner of one umapx_handle; the UmapCore of this package."""

    def __init__(self, params: L.Params):
        self._lib = L.lib()
        self._params = params
        self._h = ctypes.c_void_p()
        L.check(self._lib.umapx_create(ctypes.byref(params), ctypes.byref(self._h)))
        self.n_components = params.n_components

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.umapx_destroy(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self._h = None

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def fit(self, arr: np.ndarray) -> None:
        L.check(self._lib.umapx_fit(self._h, L.ptr(arr), arr.shape[0], arr.shape[1]))

    def fit_transform_into(self, arr: np.ndarray, out: np.ndarray) -> None:
        L.check(self._lib.umapx_fit_transform(self._h, L.ptr(arr), arr.shape[0], arr.shape[1],
                                              L.ptr(out)))

    def fit_transform(self, arr: np.ndarray) -> np.ndarray:
        out = np.empty((arr.shape[0], self.n_components), np.float32)
        self.fit_transform_into(arr, out)
        return out

    def fit_transform_with_knn_into(self, arr, idx, dist, out, knn_metric, validate) -> None:
        metric = _GPU_METRICS.get(_METRIC_ALIASES.get(str(knn_metric).lower(), ""), -1)
        if metric < 0:
            raise ValueError(f"unsupported knn_metric '{knn_metric}'")
        L.check(self._lib.umapx_fit_transform_with_knn(
            self._h, L.ptr(arr), arr.shape[0], arr.shape[1], L.ptr(idx, L._i64p), idx.shape[0],
            idx.shape[1], L.ptr(dist), dist.shape[0], dist.shape[1], metric, int(bool(validate)),
            L.ptr(out)))

    def profile_fit_transform(self, arr: np.ndarray) -> dict[str, Any]:
        out = np.empty((arr.shape[0], self.n_components), np.float32)
        t = L.Timings()
        L.check(self._lib.umapx_profile_fit_transform(self._h, L.ptr(arr), arr.shape[0],
                                                      arr.shape[1], L.ptr(out), ctypes.byref(t)))
        td = t.as_dict()
        return {
            "embedding": out,
            "timings": {k: td[k] for k in L.TIMING_KEYS},
            "used_approximate_knn": bool(td["used_approximate_knn"]),
            "n_samples": int(arr.shape[0]),
            "n_features": int(arr.shape[1]),
            "n_epochs": int(td["n_epochs"]),
            "n_edges": int(td["n_edges"]),
            "gpu": {k: td[k] for k in ("knn_kernel_ms", "knn_flops", "sgd_kernel_ms",
                                       "init_used_spectral", "n_graph_components")},
        }

    def fit_transform_csr_into(self, c: "_CsrParts", out: np.ndarray | None) -> None:
        L.check(self._lib.umapx_fit_transform_csr(
            self._h, c.shape[0], c.shape[1], L.ptr(c.indptr, L._i64p), L.ptr(c.indices, L._i64p),
            L.ptr(c.data), c.data.shape[0], None if out is None else L.ptr(out)))

    def transform_csr_into(self, c: "_CsrParts", out: np.ndarray) -> None:
        L.check(self._lib.umapx_transform_csr(
            self._h, c.shape[0], c.shape[1], L.ptr(c.indptr, L._i64p), L.ptr(c.indices, L._i64p),
            L.ptr(c.data), c.data.shape[0], L.ptr(out)))

    def timings(self) -> dict[str, Any]:
        t = L.Timings()
        L.check(self._lib.umapx_get_timings(self._h, ctypes.byref(t)))
        return t.as_dict()

    def mlp_loss(self) -> tuple[float, float, int]:
        first, last, steps = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
        L.check(self._lib.umapx_get_mlp_loss(self._h, ctypes.byref(first), ctypes.byref(last),
                                             ctypes.byref(steps)))
        return first.value, last.value, steps.value

    def transform_into(self, arr: np.ndarray, out: np.ndarray, mode: str | None = None) -> None:
        fn = {None: self._lib.umapx_transform, "plain": self._lib.umapx_transform_plain,
              "learned": self._lib.umapx_transform_learned}[mode]
        L.check(fn(self._h, L.ptr(arr), arr.shape[0], arr.shape[1], L.ptr(out)))

    def transform(self, arr: np.ndarray, mode: str | None = None) -> np.ndarray:
        out = np.empty((arr.shape[0], self.n_components), np.float32)
        self.transform_into(arr, out, mode)
        return out

    def inverse_transform_into(self, arr: np.ndarray, out: np.ndarray) -> None:
        L.check(self._lib.umapx_inverse_transform(self._h, L.ptr(arr), arr.shape[0],
                                                  arr.shape[1], L.ptr(out)))

    def inverse_transform(self, arr: np.ndarray) -> np.ndarray:
        d = self.n_features
        if d is None:
            raise RuntimeError("model is not fitted yet")
        out = np.empty((arr.shape[0], d), np.float32)
        self.inverse_transform_into(arr, out)
        return out

    @property
    def embedding(self) -> np.ndarray | None:
        n = ctypes.c_int64()
        if self._lib.umapx_n_samples(self._h, ctypes.byref(n)) != L.OK:
            return None
        out = np.empty((n.value, self.n_components), np.float32)
        L.check(self._lib.umapx_get_embedding(self._h, L.ptr(out)))
        return out

    @property
    def n_features(self) -> int | None:
        d = ctypes.c_int64()
        if self._lib.umapx_n_features(self._h, ctypes.byref(d)) != L.OK:
            return None
        return int(d.value)

    def radii(self) -> tuple[np.ndarray, np.ndarray] | None:
        n = ctypes.c_int64()
        if self._lib.umapx_n_samples(self._h, ctypes.byref(n)) != L.OK:
            return None
        ro = np.empty(n.value, np.float32)
        re = np.empty(n.value, np.float32)
        if self._lib.umapx_get_radii(self._h, L.ptr(ro), L.ptr(re)) != L.OK:
            return None
        return ro, re

    def ab_params(self) -> tuple[float, float] | None:
        a = ctypes.c_float()
        b = ctypes.c_float()
        if self._lib.umapx_get_ab(self._h, ctypes.byref(a), ctypes.byref(b)) != L.OK:
            return None
        return a.value, b.value


class Umap:
    """UMAP on one MI355X GPU with the umapers ``Umap`` interface.

    Example
    -------
    >>> import numpy as np
    >>> from umapers_amd import Umap
    >>> x = np.random.default_rng(42).normal(size=(100, 8)).astype(np.float32)
    >>> emb = Umap(n_neighbors=15, n_components=2).fit_transform(x)
    """

    def __init__(
        self,
        *,
        n_neighbors: int = 15,
        n_components: int = 2,
        n_epochs: int | None = None,
        metric: str = "euclidean",
        metric_kwds: dict[str, Any] | None = None,
        learning_rate: float = 1.0,
        min_dist: float = 0.1,
        spread: float = 1.0,
        local_connectivity: float = 1.0,
        set_op_mix_ratio: float = 1.0,
        repulsion_strength: float = 1.0,
        negative_sample_rate: int = 5,
        random_seed: int = 42,
        init: str = "spectral",
        ann_mode: str = "auto",
        use_approximate_knn: bool = True,
        approx_knn_candidates: int = 50,
        approx_knn_iters: int = 14,
        approx_knn_threshold: int = 4096,
        densmap: bool = False,
        dens_lambda: float = 2.0,
        dens_frac: float = 0.3,
        dens_var_shift: float = 0.1,
        output_dens: bool = False,
        target_metric: str | None = None,
        target_weight: float = 0.5,
        target_n_neighbors: int | None = None,
        knn_precision: str = "fp32",
        rerank_factor: int = 2,
        transform_mode: str = "plain",
        smooth_fp64: bool = True,
        mlp_hidden: int = 256,
        mlp_steps: int = 1500,
        mlp_batch: int = 4096,
        mlp_negatives: int = 5,
        mlp_learning_rate: float = 1e-3,
        mlp_embedding_weight: float = 1.0,
    ) -> None:
        for fitted_attr in ("embedding_", "n_features_in_", "n_samples_fit_", "radii_original_",
                            "radii_embedding_", "rad_orig_", "rad_emb_"):
            if hasattr(self, fitted_attr):
                delattr(self, fitted_attr)
        self.n_neighbors = int(n_neighbors)
        self.n_components = int(n_components)
        self.n_epochs = None if n_epochs is None else int(n_epochs)
        self.metric, self.metric_kwds, self._metric_p = _parse_metric_config(metric, metric_kwds)
        self.learning_rate = float(learning_rate)
        self.min_dist = float(min_dist)
        self.spread = float(spread)
        self.local_connectivity = float(local_connectivity)
        self.set_op_mix_ratio = float(set_op_mix_ratio)
        self.repulsion_strength = float(repulsion_strength)
        self.negative_sample_rate = int(negative_sample_rate)
        self.random_seed = int(random_seed)
        self.init = str(init)
        if self.init.lower() not in ("random", "spectral"):
            raise ValueError(f"unsupported init '{init}', expected random|spectral")
        if target_metric not in (None, "none", ""):
            raise ValueError("supervised (target_metric) fitting is not available on the GPU path")
        self.target_metric = None
        self.target_weight = float(target_weight)
        self.target_n_neighbors = target_n_neighbors
        self.ann_mode, use_approximate_knn, approx_knn_threshold = _normalize_ann_mode(
            ann_mode, use_approximate_knn, approx_knn_threshold)
        self.use_approximate_knn = bool(use_approximate_knn)
        self.approx_knn_candidates = int(approx_knn_candidates)
        self.approx_knn_iters = int(approx_knn_iters)
        self.approx_knn_threshold = int(approx_knn_threshold)
        self.densmap = bool(densmap)
        self.dens_lambda, self.dens_frac, self.dens_var_shift = _validate_dens_params(
            dens_lambda, dens_frac, dens_var_shift)
        self.output_dens = bool(output_dens)
        if knn_precision not in ("fp32", "bf16_rerank"):
            raise ValueError(f"unsupported knn_precision '{knn_precision}', expected fp32|bf16_rerank")
        self.knn_precision = knn_precision
        self.rerank_factor = int(rerank_factor)
        if transform_mode not in ("plain", "learned"):
            raise ValueError(f"unsupported transform_mode '{transform_mode}', expected plain|learned")
        self.transform_mode = transform_mode
        self.smooth_fp64 = bool(smooth_fp64)
        self.mlp_hidden = int(mlp_hidden)
        self.mlp_steps = int(mlp_steps)
        self.mlp_batch = int(mlp_batch)
        self.mlp_negatives = int(mlp_negatives)
        self.mlp_learning_rate = float(mlp_learning_rate)
        self.mlp_embedding_weight = float(mlp_embedding_weight)
        self._core = _Core(self._c_params())

    def _c_params(self) -> L.Params:
        p = L.default_params()
        p.n_neighbors = self.n_neighbors
        p.n_components = self.n_components
        p.n_epochs = 0 if self.n_epochs is None else self.n_epochs
        p.metric = _GPU_METRICS[self.metric]
        p.metric_p = 2.0 if self._metric_p is None else self._metric_p
        p.learning_rate = self.learning_rate
        p.min_dist = self.min_dist
        p.spread = self.spread
        p.local_connectivity = self.local_connectivity
        p.set_op_mix_ratio = self.set_op_mix_ratio
        p.repulsion_strength = self.repulsion_strength
        p.negative_sample_rate = self.negative_sample_rate
        p.init = L.INIT_SPECTRAL if self.init.lower() == "spectral" else L.INIT_RANDOM
        p.random_seed = self.random_seed & 0xFFFFFFFFFFFFFFFF
        p.use_approximate_knn = int(self.use_approximate_knn)
        p.approx_knn_candidates = self.approx_knn_candidates
        p.approx_knn_iters = self.approx_knn_iters
        p.approx_knn_threshold = self.approx_knn_threshold
        p.densmap = int(self.densmap)
        p.dens_lambda = self.dens_lambda
        p.dens_frac = self.dens_frac
        p.dens_var_shift = self.dens_var_shift
        p.output_dens = int(self.output_dens)
        p.knn_precision = L.KNN_FP32 if self.knn_precision == "fp32" else L.KNN_BF16_RERANK
        p.rerank_factor = self.rerank_factor
        p.smooth_fp64 = int(self.smooth_fp64)
        p.transform_mode = L.TRANSFORM_PLAIN if self.transform_mode == "plain" else L.TRANSFORM_LEARNED
        p.mlp_hidden = self.mlp_hidden
        p.mlp_steps = self.mlp_steps
        p.mlp_batch = self.mlp_batch
        p.mlp_negatives = self.mlp_negatives
        p.mlp_learning_rate = self.mlp_learning_rate
        p.mlp_embedding_weight = self.mlp_embedding_weight
        return p

    def get_params(self, deep: bool = True) -> dict[str, Any]:
        return {name: getattr(self, name) for name in _PARAM_NAMES}

    def set_params(self, **params: Any) -> "Umap":
        if not params:
            return self
        valid = set(_PARAM_NAMES)
        for key in params:
            if key not in valid:
                raise ValueError(f"Invalid parameter '{key}' for estimator Umap")
        current = self.get_params(deep=False)
        current.update(params)
        self.__init__(**current)
        return self

    def _sync_fitted(self, embedding: np.ndarray, n_features: int, n_samples: int) -> None:
        self.embedding_ = embedding
        self.n_features_in_ = int(n_features)
        self.n_samples_fit_ = int(n_samples)
        radii = self._core.radii() if self.output_dens else None
        if radii is None:
            for attr in ("radii_original_", "radii_embedding_", "rad_orig_", "rad_emb_"):
                if hasattr(self, attr):
                    delattr(self, attr)
            return
        self.radii_original_, self.radii_embedding_ = radii
        self.rad_orig_, self.rad_emb_ = radii

    def _fit_rows(self, data: Any):
        """Dense rows, or canonical CSR parts for sparse input, which follows
        the reference's sparse-fit rules (_api.py:562-570, lib.rs:2241-2258)."""
        if not _is_sparse(data):
            return _as_f32_matrix(data, "data")
        if self.densmap or self.output_dens:
            raise ValueError("densmap and output_dens currently support dense input only")
        if self.metric in ("canberra", "braycurtis"):
            name = {"canberra": "Canberra", "braycurtis": "BrayCurtis"}[self.metric]
            raise ValueError(f"invalid parameter: {name} metric is currently supported for dense "
                             f"input only")
        return _as_csr_parts(data, "data")

    def fit(self, data: Any, y: Any | None = None) -> "Umap":
        arr = self._fit_rows(data)
        if isinstance(arr, _CsrParts):
            self._core.fit_transform_csr_into(arr, None)
        else:
            self._core.fit(arr)
        self._fitted_sparse = isinstance(arr, _CsrParts)
        self._sync_fitted(self._core.embedding, arr.shape[1], arr.shape[0])
        return self

    def fit_transform(self, data: Any, y: Any | None = None, *,
                      out: np.ndarray | None = None):
        arr = self._fit_rows(data)
        sparse = isinstance(arr, _CsrParts)
        if out is not None and self.output_dens:
            raise ValueError("out cannot be used when output_dens=True")
        shape = (arr.shape[0], self.n_components)
        out_buf = np.empty(shape, np.float32) if out is None else _as_out_buffer(out, shape)
        if sparse:
            self._core.fit_transform_csr_into(arr, out_buf)
        else:
            self._core.fit_transform_into(arr, out_buf)
        self._fitted_sparse = sparse
        self._sync_fitted(out_buf if out is None else out_buf.copy(), arr.shape[1], arr.shape[0])
        if out is None and self.output_dens:
            return out_buf, self.radii_original_, self.radii_embedding_
        return out_buf

    def profile_fit_transform(self, data: Any) -> dict[str, Any]:
        if _is_sparse(data):
            raise ValueError("profile_fit_transform currently supports dense input only")
        if self.densmap or self.output_dens:
            raise ValueError("profile_fit_transform currently supports standard UMAP only")
        arr = _as_f32_matrix(data, "data")
        self._fitted_sparse = False
        result = self._core.profile_fit_transform(arr)
        self._sync_fitted(result["embedding"], arr.shape[1], arr.shape[0])
        return result

    def fit_transform_with_knn(self, data: Any, knn_indices: Any, knn_dists: Any, *,
                               knn_metric: str = "euclidean", validate_precomputed: bool = True,
                               out: np.ndarray | None = None) -> np.ndarray:
        arr = _as_f32_matrix(data, "data")
        idx = _as_knn_indices(knn_indices, "knn_indices")
        dist = _as_f32_matrix(knn_dists, "knn_dists")
        shape = (arr.shape[0], self.n_components)
        out_buf = np.empty(shape, np.float32) if out is None else _as_out_buffer(out, shape)
        self._core.fit_transform_with_knn_into(arr, idx, dist, out_buf, knn_metric,
                                               validate_precomputed)
        self._fitted_sparse = False
        self._sync_fitted(out_buf if out is None else out_buf.copy(), arr.shape[1], arr.shape[0])
        return out_buf

    def transform(self, query: Any, *, out: np.ndarray | None = None) -> np.ndarray:
        if _is_sparse(query):
            if getattr(self, "_fitted_sparse", None) is False:  # fitted on dense rows
                raise ValueError("sparse query transform is supported only for sparse-fitted "
                                 "models")
            c = _as_csr_parts(query, "query")
            shape = (c.shape[0], self.n_components)
            out_buf = np.empty(shape, np.float32) if out is None else _as_out_buffer(out, shape)
            self._core.transform_csr_into(c, out_buf)
            return out_buf
        arr = _as_f32_matrix(query, "query")
        shape = (arr.shape[0], self.n_components)
        if out is None:
            return self._core.transform(arr)
        out_buf = _as_out_buffer(out, shape)
        self._core.transform_into(arr, out_buf)
        return out_buf

    def transform_plain(self, query: Any) -> np.ndarray:
        """The reference's kNN + SGD transform, whatever transform_mode says."""
        return self._core.transform(_as_f32_matrix(query, "query"), "plain")

    def transform_learned(self, query: Any) -> np.ndarray:
        """One pass through the learned encoder (trained on first use)."""
        return self._core.transform(_as_f32_matrix(query, "query"), "learned")

    def learned_transform_loss(self) -> tuple[float, float, int]:
        """Training loss of the learned encoder at its first and last step, and
        the number of steps (raises when it has not been trained)."""
        return self._core.mlp_loss()

    def inverse_transform(self, embedded_query: Any, *, out: np.ndarray | None = None) -> np.ndarray:
        """_api.py:838-873: map embedded samples back to the original feature
        space (kNN among the fitted embedding, a local linear estimate and the
        feature-space SGD, lib.rs:1800-1915).  out: optional writable float32
        buffer of shape (n_samples, n_features); the model must be fit first."""
        arr = _as_f32_matrix(embedded_query, "embedded_query")
        # lib.rs:1806-1810; the out= buffer is checked first, as the reference's
        # facade does before its core call
        sparse_msg = "invalid parameter: inverse_transform is not supported for sparse-trained models yet"
        if out is None:
            if getattr(self, "_fitted_sparse", False):
                raise ValueError(sparse_msg)
            return self._core.inverse_transform(arr)
        n_features = self._core.n_features
        if n_features is None:
            raise RuntimeError("model must be fit before inverse_transform(out=...)")
        out_buf = _as_out_buffer(out, (arr.shape[0], n_features))
        if getattr(self, "_fitted_sparse", False):
            raise ValueError(sparse_msg)
        self._core.inverse_transform_into(arr, out_buf)
        return out_buf

    @property
    def ab_params(self) -> tuple[float, float] | None:
        return self._core.ab_params()


def fit_transform(data: Any, y: Any | None = None, **kwargs: Any):
    """_api.py:1054-1094: embed a dataset in one call."""
    return Umap(**kwargs).fit_transform(data, y=y)
