// Shared host/device helpers for the umapx HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace umapx {

// Error carried from any stage to the C-ABI boundary, which turns it into a
// return code plus the thread-local message (umapx_last_error).
struct Error : std::runtime_error {
  int code;
  // The sizes of UmapError's structured variants (lib.rs:198-216):
  // InconsistentDimensions{row, expected, got}, FeatureMismatch{expected,
  // got}, EmbeddingDimensionMismatch{expected, got}; -1 where a variant has
  // no such field.
  int64_t row = -1, expected = -1, got = -1;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
  Error(int c, const std::string& m, int64_t r, int64_t e, int64_t g)
      : std::runtime_error(m), code(c), row(r), expected(e), got(g) {}
};

// Codes mirror UmapError (rust_umap/src/lib.rs:198-244) plus a device error.
enum : int {
  kOk = 0,
  kEmptyData = 1,
  kNeedTwoSamples = 2,
  kInconsistentDimensions = 3,
  kFeatureMismatch = 4,
  kEmbeddingDimensionMismatch = 5,
  kNotFitted = 6,
  kInvalidParameter = 7,
  kDeviceError = 8,
};

[[noreturn]] inline void throw_invalid(const std::string& msg) {
  throw Error(kInvalidParameter, "invalid parameter: " + msg);
}

// Records e as the calling thread's last failure (umapx_last_error and
// umapx_last_error_detail) and returns its code.
int32_t set_last_error(const Error& e);

void check_hip(hipError_t e, const char* what, const char* file, int line);
#define HIP_CHECK(expr) ::umapx::check_hip((expr), #expr, __FILE__, __LINE__)
#define KERNEL_CHECK() ::umapx::check_hip(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// Test hook: when UMAPX_FAULT_INJECT names a stage, that stage reports a
// device error before touching the GPU (host-side, so CPU-only tests can
// check error propagation through the C-ABI).
void maybe_inject_fault(const char* stage);

// Stream on which DevBuf allocations of this thread are stream-ordered
// (hipMallocAsync / hipFreeAsync from the device's default memory pool),
// set by AllocStreamScope while a handle works on its own stream; with no
// scope active, DevBuf uses hipMalloc / hipFree.  Stream-ordered buffers are
// freed without a device-wide synchronisation, so a fit's temporaries do not
// stall the queue between stages.
inline hipStream_t& alloc_stream() {
  static thread_local hipStream_t s = nullptr;
  return s;
}

struct AllocStreamScope {
  hipStream_t prev;
  explicit AllocStreamScope(hipStream_t s) : prev(alloc_stream()) { alloc_stream() = s; }
  ~AllocStreamScope() { alloc_stream() = prev; }
  AllocStreamScope(const AllocStreamScope&) = delete;
  AllocStreamScope& operator=(const AllocStreamScope&) = delete;
};

// Keeps memory freed to the current device's default pool for reuse instead
// of returning it to the driver at synchronisation points.
void retain_pool_memory();

// RAII device buffer.  A buffer allocated under an AllocStreamScope is freed
// on the same stream, which must outlive it.  A borrowed buffer (borrow())
// is a fixed-capacity view of memory owned elsewhere: it is never freed, and
// growing it past its capacity is an error instead of an allocation.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipStream_t s = nullptr;  // allocation stream (stream-ordered), or nullptr
  bool borrowed = false;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), s(o.s), borrowed(o.borrowed) {
    o.p = nullptr;
    o.n = 0;
    o.s = nullptr;
    o.borrowed = false;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      s = o.s;
      borrowed = o.borrowed;
      o.p = nullptr;
      o.n = 0;
      o.s = nullptr;
      o.borrowed = false;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void borrow(T* ptr, size_t capacity) {
    release();
    p = ptr;
    n = capacity;
    borrowed = true;
  }
  void alloc(size_t count) {
    if (borrowed) {
      if (count > n)
        throw Error(kInvalidParameter,
                    "invalid parameter: borrowed device buffer of " + std::to_string(n) +
                        " elements cannot hold " + std::to_string(count));
      return;
    }
    release();
    if (count == 0) return;
    maybe_inject_fault("alloc");
    s = alloc_stream();
    if (s)
      HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&p), count * sizeof(T), s));
    else
      HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  // Grow-only reallocation (contents are not kept).
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (borrowed) {
      p = nullptr;
      n = 0;
      borrowed = false;
      return;
    }
    if (p) {
      if (s)
        (void)hipFreeAsync(p, s);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    s = nullptr;
  }
  T* get() const { return p; }
  size_t size() const { return n; }
};

inline unsigned blocks_for(size_t n, unsigned threads) {
  return static_cast<unsigned>((n + threads - 1) / threads);
}

// Device-side scalar helpers with IEEE round-to-nearest and no FMA
// contraction, used where the reference's f32 arithmetic must be reproduced
// bit for bit (hip's __fadd_rn / __fmul_rn are plain operators that the
// compiler may fuse).  f32 division and sqrt are correctly rounded under
// hipcc's default -fhip-fp32-correctly-rounded-divide-sqrt.
__device__ __forceinline__ float fadd(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ float fsub(float a, float b) {
#pragma clang fp contract(off)
  return a - b;
}
__device__ __forceinline__ float fmul(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float fdiv(float a, float b) {
#pragma clang fp contract(off)
  return a / b;
}
__device__ __forceinline__ float fsqrt(float a) { return __builtin_sqrtf(a); }

// f32::total_cmp key (sign-magnitude -> two's complement order).
__device__ __forceinline__ int total_key(float v) {
  int b = __float_as_int(v);
  return b ^ static_cast<int>(static_cast<unsigned>(b >> 31) >> 1);
}

}  // namespace umapx
