// Exact kNN candidates for rows wider than 256 features (euclidean, cosine).
//
// Past 256 features a query row no longer fits the fused candidate kernel's
// registers, so the work splits in two: the score block
// S[i][j] = q_i . r_j of a block of queries against every reference row is a
// plain GEMM (rocBLAS, BF16 or FP32 operands with FP32 accumulation) written
// to HBM, and a selection kernel reads each query's row of S once or twice
// and keeps the kc reference rows with the largest q.r - ||r||^2 / 2 (the
// same score the fused kernel ranks by).  The exact FP32 re-rank then orders
// the kc candidates as for narrow rows.
//
// The GEMM library may accumulate different tiles in different orders, so
// two identical reference rows need not score bit-identically against a
// query; the fused kernel's scores do, which lets the lowest indices win
// among duplicates as the reference's (distance, index) order requires.  The
// selection therefore reads every reference row's score at its first
// identical row (found by hashing the rows, radix-sorting the hashes and
// comparing each row with the head of its run), so identical rows tie
// exactly here too.
//
// Selection, one wave per query row:
//  1. every lane keeps the best kSelT scores of its share of the row in
//     registers; the kc-th best of those 64 lists, theta, is at most the
//     row's true kc-th best score.  The same pass gathers into LDS (in index
//     order, by ballot prefix) every entry at or above the lists' running
//     kc-th best, which never exceeds theta, so LDS ends up holding the
//     row's top kc; a bitonic sort by (score descending, index ascending)
//     gives them;
//  2. when the running thresholds let more than kSelCap entries through, a
//     second pass gathers again against theta itself;
//  3. when more than kSelCap entries tie at or above theta (many duplicate
//     rows), the kc-th best score is found exactly by bisection over the
//     ordered score bits, and the entries above it plus the lowest-index
//     entries equal to it are taken.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <climits>
#include <map>
#include <cstdlib>
#include <mutex>
#include <string>

#include "knn.hpp"

namespace umapx {
namespace {

constexpr int kSelT = 8;        // per-lane list of the first pass
constexpr int kSelCap = 512;    // gathered entries per wave (LDS)
constexpr int kUnroll = 2;      // 16-byte chunks per lane in flight per step
constexpr int kSelWaves = 4;
constexpr size_t kScoreBudget = size_t(4) << 30;  // bytes of one score block

__device__ __forceinline__ uint32_t order_bits(float v) {
  const uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// ascending 64-bit key: higher score first, then lower index
__device__ __forceinline__ uint64_t sel_key(float v, int64_t j) {
  return (static_cast<uint64_t>(~order_bits(v)) << 32) | static_cast<uint32_t>(j);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ void sort_keys(uint64_t* a, int P, int lane) {
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = lane; t < P / 2; t += 64) {
        const int lo = 2 * t - (t & (stride - 1));
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t x = a[lo], y = a[hi];
        if ((x > y) == up) {
          a[lo] = y;
          a[hi] = x;
        }
      }
      wave_lds_sync();
    }
}

__device__ __forceinline__ uint64_t lanes_below(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Order-dependent 64-bit hash of each row's bits, one wave per row.
__global__ void row_hash_kernel(const float* __restrict__ x, int64_t n, int d,
                                uint64_t* __restrict__ keys, int32_t* __restrict__ ids) {
  const int lane = threadIdx.x & 63;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= n) return;
  uint64_t h = 0;
  for (int f = lane; f < d; f += 64)
    h += mix64((static_cast<uint64_t>(f) << 32) | __float_as_uint(x[r * d + f]));
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t lo = __shfl_xor(static_cast<uint32_t>(h), o);
    const uint32_t hi = __shfl_xor(static_cast<uint32_t>(h >> 32), o);
    h += (static_cast<uint64_t>(hi) << 32) | lo;
  }
  if (lane == 0) {
    keys[r] = h;
    ids[r] = static_cast<int32_t>(r);
  }
}

// Position of each sorted entry when it starts a run of equal hashes, else 0
// (a running max then gives every entry the start of its run).
__global__ void run_start_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                 int32_t* __restrict__ start) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  start[p] = (p == 0 || keys[p] != keys[p - 1]) ? static_cast<int32_t>(p) : 0;
}

// rep[j]: the first row identical to row j (j itself when none).  The sort is
// stable, so a run's first entry has its lowest row index.
__global__ void first_copy_kernel(const float* __restrict__ x, int64_t n, int d,
                                  const int32_t* __restrict__ ids,
                                  const int32_t* __restrict__ start, int32_t* __restrict__ rep) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t j = ids[p];
  const int32_t h = ids[start[p]];
  bool same = h != j;
  for (int f = 0; same && f < d; ++f)
    same = __float_as_uint(x[static_cast<int64_t>(j) * d + f]) ==
           __float_as_uint(x[static_cast<int64_t>(h) * d + f]);
  rep[j] = same ? h : j;
}

// Score rows have a stride of n rounded up to 4 (16-byte aligned rows): the
// two main passes read 4 consecutive entries per lane with vector loads,
// kUnroll of them issued together (2 keeps the kernel within 64 VGPRs, 8
// waves per SIMD).
template <int KC>
__global__ __launch_bounds__(kSelWaves * 64) void wide_select_kernel(
    const float* __restrict__ scores, int64_t ld, int64_t rows, int64_t n,
    const float* __restrict__ nr_half, const int32_t* __restrict__ rep,
    int32_t* __restrict__ cand) {
  __shared__ uint64_t s_keys[kSelWaves][kSelCap];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kSelWaves + wv;
  if (i >= rows) return;
  const float* row = scores + i * ld;
  uint64_t* keys = s_keys[wv];
  auto score = [&](int64_t j) { return fadd(row[rep[j]], nr_half[j]); };
  // kUnroll chunks of 4 consecutive entries per lane, base + 256 u + 4 lane
  // (u < kUnroll), all loads issued before any is used; -inf past n
  auto score_chunks = [&](int64_t base, float (&x)[kUnroll][4]) {
    float4 sv[kUnroll], nv[kUnroll];
    int4 rv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t j0 = base + 256 * u + 4 * lane;
      const bool full = j0 + 4 <= n;
      sv[u] = j0 < n ? *reinterpret_cast<const float4*>(row + j0) : make_float4(0.f, 0.f, 0.f, 0.f);
      nv[u] = full ? *reinterpret_cast<const float4*>(nr_half + j0) : make_float4(0.f, 0.f, 0.f, 0.f);
      rv[u] = full ? *reinterpret_cast<const int4*>(rep + j0) : make_int4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t j0 = base + 256 * u + 4 * lane;
      if (j0 + 4 <= n) {
        x[u][0] = fadd(rv[u].x == j0 ? sv[u].x : row[rv[u].x], nv[u].x);
        x[u][1] = fadd(rv[u].y == j0 + 1 ? sv[u].y : row[rv[u].y], nv[u].y);
        x[u][2] = fadd(rv[u].z == j0 + 2 ? sv[u].z : row[rv[u].z], nv[u].z);
        x[u][3] = fadd(rv[u].w == j0 + 3 ? sv[u].w : row[rv[u].w], nv[u].w);
      } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) x[u][t] = j0 + t < n ? score(j0 + t) : -INFINITY;
      }
    }
  };

  // kc-th best (value) of the 64 lane lists, popping the wave's largest head
  // kc times from a copy
  auto kth_of_lists = [&](const float (&lists)[kSelT]) {
    float w[kSelT];
#pragma unroll
    for (int t = 0; t < kSelT; ++t) w[t] = lists[t];
    float th = -INFINITY;
    for (int r = 0; r < KC; ++r) {
      float mx = w[0];
      for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
      const uint64_t b = __ballot(w[0] == mx);
      if (lane == __ffsll(static_cast<unsigned long long>(b)) - 1) {
#pragma unroll
        for (int t = 0; t < kSelT - 1; ++t) w[t] = w[t + 1];
        w[kSelT - 1] = -INFINITY;
      }
      th = mx;
    }
    return th;
  };
  const uint64_t below = lanes_below(lane);
  int total = 0;
  // appends the chunk's entries >= th to keys, in index order (lane by lane,
  // 4 entries each)
  auto gather = [&](int64_t j0, const float (&xs)[4], float th) {
    bool take[4];
    int before = 0, mine = 0, all = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      take[e] = j0 + e < n && xs[e] >= th;  // th is -inf while fewer than kc seen
      const uint64_t b = __ballot(take[e]);
      before += __popcll(b & below);
      all += __popcll(b);
    }
    if (all) {
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (take[e]) {
          const int pos = total + before + mine++;
          if (pos < kSelCap) keys[pos] = sel_key(xs[e], j0 + e);
        }
      total += all;
    }
  };

  // 1. one pass: per-lane best kSelT (descending) and, against a running
  //    threshold (the kc-th best of the lists so far, a lower bound of the
  //    row's kc-th best, refreshed at steps 1, 2, 4, ..., 32 and then every
  //    32), the gathered candidates
  float v[kSelT];
#pragma unroll
  for (int t = 0; t < kSelT; ++t) v[t] = -INFINITY;
  float run_theta = -INFINITY;
  int step = 0, next_refresh = 1;
  for (int64_t base = 0; base < n; base += 256 * kUnroll) {
    float xs[kUnroll][4];
    score_chunks(base, xs);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float x = xs[u][e];
        if (x > v[kSelT - 1]) {
#pragma unroll
          for (int t = kSelT - 1; t > 0; --t)
            v[t] = x > v[t - 1] ? v[t - 1] : (x > v[t] ? x : v[t]);
          v[0] = x > v[0] ? x : v[0];
        }
      }
    if (++step == next_refresh) {
      run_theta = kth_of_lists(v);
      next_refresh = step < 32 ? 2 * step : step + 32;
    }
    if (total <= kSelCap) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) gather(base + 256 * u + 4 * lane, xs[u], run_theta);
    }
  }
  // the row's kc-th best is at least theta, and every entry >= theta was
  // gathered (each met a running threshold <= theta)
  const float theta = kth_of_lists(v);

  // 2. when the running thresholds let too many through, gather again
  //    against the final one
  if (total > kSelCap) {
    total = 0;
    for (int64_t base0 = 0; base0 < n; base0 += 256 * kUnroll) {
      float xu[kUnroll][4];
      score_chunks(base0, xu);
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) gather(base0 + 256 * u + 4 * lane, xu[u], theta);
    }
  }
  int got = total;
  if (total > kSelCap) {
    // 3. exact kc-th best by bisection: the largest U with
    //    #{order_bits(score) >= U} >= KC
    auto count_ge = [&](uint32_t u) {
      int c = 0;
      for (int64_t base = 0; base < n; base += 64) {
        const int64_t j = base + lane;
        const bool hit = j < n && order_bits(score(j)) >= u;
        c += __popcll(__ballot(hit));
      }
      return c;
    };
    uint64_t lo = order_bits(theta), hi = 0xFFFFFFFFull;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo + 1) / 2;
      if (count_ge(static_cast<uint32_t>(mid)) >= KC)
        lo = mid;
      else
        hi = mid - 1;
    }
    const uint32_t u_star = static_cast<uint32_t>(lo);
    got = 0;
    for (int pass = 0; pass < 2 && got < KC; ++pass) {
      for (int64_t base = 0; base < n && got < KC; base += 64) {
        const int64_t j = base + lane;
        float x = -INFINITY;
        bool take = false;
        if (j < n) {
          x = score(j);
          const uint32_t u = order_bits(x);
          take = pass == 0 ? u > u_star : u == u_star;
        }
        const uint64_t b = __ballot(take);
        if (b) {
          const int pos = got + __popcll(b & lanes_below(lane));
          if (take && pos < KC) keys[pos] = sel_key(x, j);
          got += __popcll(b);
        }
      }
      if (got > KC) got = KC;
    }
  }
  wave_lds_sync();
  int P = 64;
  while (P < got) P <<= 1;
  for (int t = got + lane; t < P; t += 64) keys[t] = ~0ull;
  wave_lds_sync();
  sort_keys(keys, P, lane);
  for (int t = lane; t < KC; t += 64)
    cand[i * KC + t] = t < got ? static_cast<int32_t>(keys[t] & 0xFFFFFFFFu) : -1;
}

// The score block's size in bytes: kScoreBudget, or (tests only) the value
// of UMAPX_TEST_WIDE_SCORE_BYTES, so that small inputs run several blocks.
size_t score_budget() {
  const char* e = std::getenv("UMAPX_TEST_WIDE_SCORE_BYTES");
  const long long b = e ? std::atoll(e) : 0;
  return b > 0 ? static_cast<size_t>(b) : kScoreBudget;
}

}  // namespace

void knn_wide_candidates(const void* qp, int64_t m, const void* rp, const float* nr_half,
                         const float* r, int64_t n, int64_t d, int dpad, bool bf16, int kc,
                         int32_t* cand, KnnWorkspace& ws, hipStream_t s) {
  if (kc != 32 && kc != 64) throw_invalid("unsupported candidate count");
  if (n > INT_MAX - 3) throw_invalid("too many reference rows for the kNN");
  // first identical copy of every reference row
  DevBuf<uint64_t> hkeys(static_cast<size_t>(2 * n));
  DevBuf<int32_t> hids(static_cast<size_t>(2 * n)), start(static_cast<size_t>(n)),
      head(static_cast<size_t>(n)), rep(static_cast<size_t>(n));
  hipLaunchKernelGGL(row_hash_kernel, dim3(blocks_for(static_cast<size_t>(n), 4)), dim3(256), 0,
                     s, r, n, static_cast<int>(d), hkeys.get(), hids.get());
  KERNEL_CHECK();
  {
    size_t sort_bytes = 0, scan_bytes = 0;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, hkeys.get(),
                                                 hkeys.get() + n, hids.get(), hids.get() + n,
                                                 static_cast<int>(n), 0, 64, s));
    HIP_CHECK(hipcub::DeviceScan::InclusiveScan(nullptr, scan_bytes, start.get(), head.get(),
                                                hipcub::Max(), static_cast<int>(n), s));
    DevBuf<unsigned char> tmp(std::max(sort_bytes, scan_bytes));
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.get(), sort_bytes, hkeys.get(),
                                                 hkeys.get() + n, hids.get(), hids.get() + n,
                                                 static_cast<int>(n), 0, 64, s));
    hipLaunchKernelGGL(run_start_kernel, dim3(blocks_for(static_cast<size_t>(n), 256)), dim3(256),
                       0, s, hkeys.get() + n, n, start.get());
    KERNEL_CHECK();
    HIP_CHECK(hipcub::DeviceScan::InclusiveScan(tmp.get(), scan_bytes, start.get(), head.get(),
                                                hipcub::Max(), static_cast<int>(n), s));
    hipLaunchKernelGGL(first_copy_kernel, dim3(blocks_for(static_cast<size_t>(n), 256)),
                       dim3(256), 0, s, r, n, static_cast<int>(d), hids.get() + n, head.get(),
                       rep.get());
    KERNEL_CHECK();
  }
  // the workspace's own handle: calls on other workspaces (other umapx
  // handles, other streams) never share it
  if (!ws.blas.h) {
    rocblas_handle b = nullptr;
    if (rocblas_create_handle(&b) != rocblas_status_success)
      throw Error(kDeviceError, "device error: rocblas_create_handle failed");
    ws.blas.h = b;
  }
  rocblas_handle h = static_cast<rocblas_handle>(ws.blas.h);
  if (rocblas_set_stream(h, s) != rocblas_status_success)
    throw Error(kDeviceError, "device error: rocblas_set_stream failed");
  const int64_t ld = (n + 3) / 4 * 4;  // 16-byte aligned score rows
  int64_t mb = static_cast<int64_t>(score_budget() / (sizeof(float) * static_cast<size_t>(ld)));
  mb = std::max<int64_t>(64, mb / 64 * 64);
  mb = std::min<int64_t>(mb, (m + 63) / 64 * 64);
  ws.scores.ensure(static_cast<size_t>(mb) * ld);
  const size_t esz = bf16 ? 2 : 4;
  const rocblas_datatype ty = bf16 ? rocblas_datatype_bf16_r : rocblas_datatype_f32_r;
  const float one = 1.f, zero = 0.f;
  for (int64_t q0 = 0; q0 < m; q0 += mb) {
    const int64_t rows = std::min(mb, m - q0);
    // row-major rows are column-major [dpad x count] matrices, so
    // C = R^T Q_block is column-major [n x rows] (stride ld) with
    // C(j, i) = r_j . q_i: the row-major [rows x ld] score block
    const void* qb = static_cast<const unsigned char*>(qp) + static_cast<size_t>(q0) * dpad * esz;
    const rocblas_status st = rocblas_gemm_ex(
        h, rocblas_operation_transpose, rocblas_operation_none, static_cast<rocblas_int>(n),
        static_cast<rocblas_int>(rows), dpad, &one, rp, ty, dpad, qb, ty, dpad, &zero,
        ws.scores.get(), rocblas_datatype_f32_r, static_cast<rocblas_int>(ld), ws.scores.get(),
        rocblas_datatype_f32_r, static_cast<rocblas_int>(ld), rocblas_datatype_f32_r,
        rocblas_gemm_algo_standard, 0, 0);
    if (st != rocblas_status_success)
      throw Error(kDeviceError, std::string("device error: rocblas_gemm_ex: ") +
                                    rocblas_status_to_string(st));
    const unsigned blocks = static_cast<unsigned>((rows + kSelWaves - 1) / kSelWaves);
    if (kc == 32)
      hipLaunchKernelGGL(wide_select_kernel<32>, dim3(blocks), dim3(kSelWaves * 64), 0, s,
                         ws.scores.get(), ld, rows, n, nr_half, rep.get(), cand + q0 * kc);
    else
      hipLaunchKernelGGL(wide_select_kernel<64>, dim3(blocks), dim3(kSelWaves * 64), 0, s,
                         ws.scores.get(), ld, rows, n, nr_half, rep.get(), cand + q0 * kc);
    KERNEL_CHECK();
  }
}

BlasHandle::~BlasHandle() {
  if (h) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(h));
}

}  // namespace umapx
