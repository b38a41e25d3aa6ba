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
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
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

void check_hip(hipError_t e, const char* what, const char* file, int line);
#define HIP_CHECK(expr) ::umapx::check_hip((expr), #expr, __FILE__, __LINE__)
#define KERNEL_CHECK() ::umapx::check_hip(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// Test hook: when UMAPX_FAULT_INJECT names a stage, that stage reports a
// device error before touching the GPU (host-side, so CPU-only tests can
// check error propagation through the C-ABI).
void maybe_inject_fault(const char* stage);

// RAII device buffer.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    maybe_inject_fault("alloc");
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  // Grow-only reallocation (contents are not kept).
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
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
