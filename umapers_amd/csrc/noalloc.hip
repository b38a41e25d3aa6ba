// Exact kNN into caller-owned device buffers with no device allocation.
//
// umapx_knn_to_reference_device_noalloc runs the same kernels as
// umapx_knn_to_reference (candidate MFMA tiles, then the exact f32 re-rank)
// but takes every temporary from one file-scope device array, `scratch`,
// instead of allocating: it is for callers that must not call hipMalloc
// (code under HIP graph capture, serving loops, harnesses that forbid it).
//
// Every GPU the library runs on holds its own copy of `scratch` until the
// process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress (the host-side name is only a stand-in) and only that
// address is ever handed to kernels.  All calls on one GPU share the array,
// so calls that overlap on one GPU (other streams or threads) race.
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/umapx.h"
#include "common.hpp"
#include "knn.hpp"

namespace umapx {
namespace {

static_assert(UMAPX_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size must be a multiple of 256");
static __device__ __align__(256) unsigned char scratch[UMAPX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
std::array<unsigned char*, kMaxDevices> g_scratch_ptr{};

// Device address of `scratch` on the current device (looked up on the first
// call made with that device current).
unsigned char* scratch_on_current_device() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw_invalid("device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(kDeviceError, "device error: scratch array has no device address");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch_ptr[dev];
}

// Bump allocator over the scratch array (256-byte aligned pieces).
struct Carve {
  unsigned char* base;
  size_t used = 0;
  template <typename T>
  void take(DevBuf<T>& buf, size_t count) {
    if (count == 0) return;
    buf.borrow(reinterpret_cast<T*>(base + used), count);
    used += (count * sizeof(T) + 255) / 256 * 256;
  }
};

size_t padded_bytes(size_t count, size_t elem) { return (count * elem + 255) / 256 * 256; }

__global__ void widen_idx_kernel(const int32_t* __restrict__ in, int64_t total,
                                 int64_t* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < total) out[i] = in[i];
}

}  // namespace
}  // namespace umapx

using namespace umapx;

extern "C" int32_t umapx_knn_to_reference_device_noalloc(
    const float* d_q, int64_t m, const float* d_r, int64_t n, int64_t d, int32_t k, int32_t metric,
    int32_t precision, int32_t rerank_factor, int64_t* d_idx, float* d_dist, void* stream) {
  try {
    if (k < 1) throw_invalid("n_neighbors must be >= 1");
    if (metric != UMAPX_METRIC_EUCLIDEAN && metric != UMAPX_METRIC_COSINE)
      throw_invalid("umapx_knn_to_reference_device_noalloc supports the euclidean and cosine "
                    "metrics only");
    if (precision != UMAPX_KNN_FP32 && precision != UMAPX_KNN_BF16_RERANK)
      throw_invalid("unknown kNN precision " + std::to_string(precision));
    if (m < 0 || n < 1) throw_invalid("need at least one reference row");
    if (d < 1 || knn_padded_dim(d) < 0)
      throw_invalid("umapx_knn_to_reference_device_noalloc supports 1 to 256 features");
    if (k > n) throw_invalid("n_neighbors must not exceed the number of reference rows");
    if (knn_candidate_count(k, rerank_factor) < 0)
      throw_invalid("n_neighbors must be <= 32 and n_neighbors * rerank_factor <= 64");
    if (m == 0) return UMAPX_OK;
    if (!d_q || !d_r || !d_idx || !d_dist) throw_invalid("null argument");
    const bool self = (d_q == d_r) && (m == n);
    const KnnScratchNeed need =
        knn_scratch_need(m, n, d, k, metric, precision, rerank_factor, self);
    const size_t idx32 = static_cast<size_t>(m) * k;
    const size_t bytes = padded_bytes(need.colsum, sizeof(double)) +
                         padded_bytes(need.rprep, 1) + padded_bytes(need.rnorm2, 4) +
                         padded_bytes(need.qprep, 1) + padded_bytes(need.cand, 4) +
                         padded_bytes(need.pairs, sizeof(uint2)) + padded_bytes(need.rnorm, 4) +
                         padded_bytes(need.qnorm, 4) + padded_bytes(idx32, 4);
    if (bytes > static_cast<size_t>(UMAPX_NOALLOC_SCRATCH_BYTES))
      throw_invalid("this kNN needs " + std::to_string(bytes) +
                    " bytes of scratch, more than the " +
                    std::to_string(static_cast<long long>(UMAPX_NOALLOC_SCRATCH_BYTES)) +
                    "-byte scratch array of umapx_knn_to_reference_device_noalloc");
    Carve cv{scratch_on_current_device()};
    KnnWorkspace ws;  // every member a borrowed view: ensure() never allocates
    cv.take(ws.colsum, need.colsum);
    cv.take(ws.rprep, need.rprep);
    cv.take(ws.rnorm2, need.rnorm2);
    cv.take(ws.qprep, need.qprep);
    cv.take(ws.cand, need.cand);
    cv.take(ws.pairs, need.pairs);
    cv.take(ws.rnorm, need.rnorm);
    cv.take(ws.qnorm, need.qnorm);
    DevBuf<int32_t> out_idx;
    cv.take(out_idx, idx32);
    auto zero_cap = [](auto& b) {
      if (!b.borrowed) b.borrow(nullptr, 0);
    };
    zero_cap(ws.colsum);
    zero_cap(ws.rprep);
    zero_cap(ws.rnorm2);
    zero_cap(ws.qprep);
    zero_cap(ws.cand);
    zero_cap(ws.pairs);
    zero_cap(ws.rnorm);
    zero_cap(ws.qnorm);
    zero_cap(ws.msum);
    zero_cap(ws.scores);
    hipStream_t s = static_cast<hipStream_t>(stream);
    knn_device(d_q, m, d_r, n, d, k, metric, precision, rerank_factor, out_idx.get(), d_dist, ws,
               s, nullptr);
    hipLaunchKernelGGL(widen_idx_kernel, dim3(blocks_for(idx32, 256)), dim3(256), 0, s,
                       out_idx.get(), static_cast<int64_t>(idx32), d_idx);
    KERNEL_CHECK();
    return UMAPX_OK;
  } catch (const Error& e) {
    return set_last_error(e);
  } catch (const std::exception& e) {
    return set_last_error(Error(kDeviceError, std::string("device error: ") + e.what()));
  }
}
