"""The C-ABI library loads, exports every declared symbol, and reports
validation errors with the reference's messages before touching a GPU."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "umapx.h")).read()
    return sorted(set(re.findall(r"\b(umapx_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(umapx):
    L = umapx.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
    # and the binding declares a signature for each one
    assert set(syms) == set(umapx.SIGNATURES)


def test_default_params_match_reference(umapx):
    p = umapx.default_params()
    # UmapParams::default (rust_umap/src/lib.rs:164-195)
    assert (p.n_neighbors, p.n_components, p.n_epochs) == (15, 2, 0)
    assert p.learning_rate == 1.0 and abs(p.min_dist - 0.1) < 1e-7 and p.spread == 1.0
    assert p.negative_sample_rate == 5 and p.random_seed == 42 and p.init == 1
    assert abs(p.dens_lambda - 2.0) < 1e-7 and abs(p.dens_frac - 0.3) < 1e-7


def test_find_ab_params_bits_on_host(umapx):
    from tests.test_oracle_kat import AB_KAT, f32_bits

    L = umapx.lib()
    for spread, md, ea, eb in AB_KAT:
        a, b = ctypes.c_float(), ctypes.c_float()
        assert L.umapx_find_ab_params(spread, md, ctypes.byref(a), ctypes.byref(b)) == 0
        assert f32_bits(a.value) == ea and f32_bits(b.value) == eb


def _fit_error(umapx, x, **overrides):
    p = umapx.default_params()
    for k, v in overrides.items():
        setattr(p, k, v)
    h = ctypes.c_void_p()
    assert umapx.lib().umapx_create(ctypes.byref(p), ctypes.byref(h)) == 0
    try:
        x = np.ascontiguousarray(x, dtype=np.float32)
        rc = umapx.lib().umapx_fit(h, umapx.ptr(x), x.shape[0], x.shape[1] if x.ndim == 2 else 0)
        return rc, umapx.last_error()
    finally:
        umapx.lib().umapx_destroy(h)


@pytest.mark.parametrize("x,overrides,code,msg", [
    (np.zeros((0, 3)), {}, 1, "input data is empty"),
    (np.zeros((1, 3)), {}, 2, "need at least two samples"),
    (np.zeros((5, 0)), {}, 7, "invalid parameter: input rows must have at least one feature"),
    (np.array([[0, 1], [np.nan, 2], [1, 1]]), {}, 7, "non-finite value at row 1"),
    (np.zeros((20, 3)), {"n_neighbors": 1}, 7, "n_neighbors must be >= 2"),
    (np.zeros((10, 3)), {}, 7, "n_neighbors (15) must be smaller than number of samples (10)"),
    (np.zeros((20, 3)), {"n_neighbors": 5, "learning_rate": -1.0}, 7, "learning_rate must be finite and > 0"),
    (np.zeros((20, 3)), {"n_neighbors": 5, "min_dist": 2.0}, 7, "min_dist must be <= spread"),
    (np.zeros((20, 3)), {"n_neighbors": 5, "set_op_mix_ratio": 1.5}, 7, "set_op_mix_ratio must be finite and in [0, 1]"),
    (np.zeros((20, 3)), {"n_neighbors": 5, "negative_sample_rate": 0}, 7, "negative_sample_rate must be >= 1"),
    (np.zeros((20, 3)), {"n_neighbors": 5, "dens_frac": 2.0}, 7, "dens_frac must be finite and in [0, 1]"),
    (np.zeros((20, 3)), {"n_neighbors": 5, "dens_var_shift": 0.0}, 7, "dens_var_shift must be finite and > 0"),
    # the GPU path's own bound, checked before any device work
    (np.zeros((20, 3)), {"n_neighbors": 5, "n_components": 9}, 7, "n_components must be <= 8"),
    (np.zeros((20, 3)), {"n_neighbors": 5, "approx_knn_threshold": -1}, 7, "approx_knn_threshold must be >= 0"),
])
def test_validation_messages(umapx, x, overrides, code, msg):
    rc, err = _fit_error(umapx, x, **overrides)
    assert rc == code
    assert msg in err


def test_not_fitted_maps_to_runtime_error(umapx):
    p = umapx.default_params()
    h = ctypes.c_void_p()
    assert umapx.lib().umapx_create(ctypes.byref(p), ctypes.byref(h)) == 0
    n = ctypes.c_int64()
    rc = umapx.lib().umapx_n_samples(h, ctypes.byref(n))
    assert rc == umapx.ERR_NOT_FITTED
    with pytest.raises(RuntimeError, match="not fitted"):
        umapx.check(rc)
    umapx.lib().umapx_destroy(h)


def test_injected_device_fault_propagates(tmp_path):
    # a fresh process with the fault hook armed: the first device allocation
    # reports a device error through the C ABI as RuntimeError
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r)\n"
        "from umapers_amd import Umap\n"
        "x = np.random.default_rng(0).normal(size=(50, 4)).astype(np.float32)\n"
        "try:\n"
        "    Umap(n_neighbors=5).fit_transform(x)\n"
        "except RuntimeError as e:\n"
        "    print('RUNTIME', e)\n" % ROOT)
    env = dict(os.environ, UMAPX_FAULT_INJECT="alloc")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=120)
    assert "RUNTIME device error: injected fault in stage alloc" in out.stdout, out.stderr


def test_inverse_transform_requires_fit(umapx):
    # lib.rs:6366-6373 and test_binding.py:1073-1093: NotFitted -> RuntimeError
    from umapers_amd import Umap

    m = Umap()
    with pytest.raises(RuntimeError, match="not fitted"):
        m.inverse_transform(np.zeros((1, 2), np.float32))
    with pytest.raises(RuntimeError, match=r"model must be fit before inverse_transform\(out=\.\.\.\)"):
        m.inverse_transform(np.zeros((1, 2), np.float32), out=np.empty((1, 2), np.float32))
    h = m._core._h
    rc = umapx.lib().umapx_inverse_transform(h, umapx.ptr(np.zeros((1, 2), np.float32)), 1, 2,
                                             umapx.ptr(np.zeros((1, 2), np.float32)))
    assert rc == umapx.ERR_NOT_FITTED


def test_auto_ann_dispatch_matches_reference_predicate(umapx):
    """should_use_approximate_knn (rust_umap/src/lib.rs:2285-2311) on the
    cases of the reference's own test (lib.rs:5816-5840)."""
    L = umapx.lib()
    p = umapx.default_params()
    p.use_approximate_knn = 1
    p.approx_knn_threshold = 4096
    f = lambda n, d: L.umapx_uses_approximate_knn(ctypes.byref(p), n, d)  # noqa: E731
    assert not f(4200, 96)
    assert not f(16_384, 160)
    assert f(16_385, 160)
    assert not f(16_384, 16)
    assert f(16_385, 16)
    p.approx_knn_threshold = 0  # ann_mode="approximate"
    assert f(4200, 96)
    p.approx_knn_threshold = 4096
    p.n_neighbors = 50  # past the GPU rounds' 32 neighbours, auto keeps the exact kNN
    assert not f(100_000, 160)
    p.approx_knn_threshold = 0
    assert f(100_000, 160)  # forced: the rounds report their limit
    p.n_neighbors = 15
    p.approx_knn_threshold = 4096
    p.use_approximate_knn = 0  # ann_mode="exact"
    assert not f(100_000, 160)


def test_last_error_detail_fields(umapx):
    L = umapx.lib()
    row, exp, got = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    x = np.zeros((4, 3), np.float32)
    idx = np.zeros((4, 1), np.int64)
    dist = np.zeros((4, 1), np.float32)
    rc = L.umapx_knn(umapx.ptr(x), 4, 3, 0, 0, 0, 2, umapx.ptr(idx, umapx._i64p), umapx.ptr(dist))
    assert rc == umapx.ERR_INVALID_PARAMETER
    code = L.umapx_last_error_detail(ctypes.byref(row), ctypes.byref(exp), ctypes.byref(got))
    assert code == rc
    assert (row.value, exp.value, got.value) == (-1, -1, -1)
