"""The allocation-free kNN entry point (umapx_knn_to_reference_device_noalloc)
against the allocating one it shares its kernels with
(umapx_knn_to_reference / umapx_knn): identical indices and distances, on two
calls in a row, the second with smaller inputs, so the shared scratch array's
leftovers from the first call must not leak into the second."""
import numpy as np
import pytest

from tests import capi
from tests.devmem import DeviceArray, synchronize
from umapers_amd import _lib as L

pytestmark = pytest.mark.gpu


def _noalloc(q, r, k, metric, precision, same=False):
    r_d = DeviceArray.of(np.ascontiguousarray(r, np.float32))
    q_d = r_d if same else DeviceArray.of(np.ascontiguousarray(q, np.float32))
    m = q_d.shape[0]
    idx_d = DeviceArray.filled((m, k), np.int64, -7)
    dist_d = DeviceArray.filled((m, k), np.float32, -7.0)
    rc = L.lib().umapx_knn_to_reference_device_noalloc(
        q_d.ptr, m, r_d.ptr, r_d.shape[0], r_d.shape[1], k, metric, precision, 2, idx_d.ptr,
        dist_d.ptr, None)
    synchronize()
    L.check(rc)
    return idx_d.numpy(), dist_d.numpy()


def _data(rng, n, d, metric):
    x = rng.normal(size=(n, d)).astype(np.float32)
    x[:, :4] += rng.integers(0, 4, size=(n, 1)).astype(np.float32) * 3.0  # a few clusters
    if metric == 1:
        x[::97] *= 0.0  # zero rows: the cosine special cases
    return x


@pytest.mark.parametrize("metric", [0, 1])
@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_to_reference_matches_allocating_entry(gpu, metric, precision):
    rng = np.random.default_rng(100 + 2 * metric + precision)
    for (m, n, d, k) in [(2600, 4100, 72, 15), (700, 1900, 40, 10)]:
        r = _data(rng, n, d, metric)
        q = _data(rng, m, d, metric)
        print(f"noalloc to-reference metric={metric} prec={precision} m={m} n={n} d={d} k={k}",
              flush=True)
        ia, da = _noalloc(q, r, k, metric, precision)
        ib, db = capi.knn_to_reference(q, r, k, metric, precision)
        assert np.array_equal(ia, ib)
        assert np.array_equal(da, db)


@pytest.mark.parametrize("metric", [0, 1])
def test_noalloc_self_knn_matches_allocating_entry(gpu, metric):
    rng = np.random.default_rng(7 + metric)
    for (n, d, k, precision) in [(5000, 128, 15, 1), (1200, 33, 20, 0)]:
        x = _data(rng, n, d, metric)
        print(f"noalloc self metric={metric} prec={precision} n={n} d={d} k={k}", flush=True)
        ia, da = _noalloc(x, x, k, metric, precision, same=True)
        ib, db = capi.knn(x, k, metric, precision)
        assert np.array_equal(ia, ib)
        assert np.array_equal(da, db)


def test_noalloc_rejects_what_the_scratch_cannot_hold(gpu):
    n, d = 200_000, 256
    x_d = DeviceArray((n, d), np.float32)
    idx_d = DeviceArray((n, 15), np.int64)
    dist_d = DeviceArray((n, 15), np.float32)
    rc = L.lib().umapx_knn_to_reference_device_noalloc(
        x_d.ptr, n, x_d.ptr, n, d, 15, 0, 0, 2, idx_d.ptr, dist_d.ptr, None)
    assert rc == L.ERR_INVALID_PARAMETER
    assert "bytes of scratch" in L.last_error()
    rc = L.lib().umapx_knn_to_reference_device_noalloc(
        x_d.ptr, 10, x_d.ptr, 10, 300, 5, 0, 0, 2, idx_d.ptr, dist_d.ptr, None)
    assert rc == L.ERR_INVALID_PARAMETER and "256 features" in L.last_error()
