// Approximate kNN: the reference's neighbours-of-neighbours refinement
// (approximate_nearest_neighbors_with_distance, lib.rs:2932-3168) under any
// of the reference's metrics (lib.rs:3213-3281).
//
// Each sample starts from a random pool (itself plus pool draws), and each
// round rebuilds every row from the previous table: the row's neighbours and
// their neighbours, itself and a few random explorations, sorted and
// deduplicated, cut to max_candidates by a partial random shuffle, then the k
// nearest by (distance, index).  Rounds stop when no row changes.
//
// One wave per sample.  The draws follow the reference's stream exactly:
// SmallRng (xoshiro256++ seeded by SplitMix64) seeded per (row, round) by
// mix_approx_seed, and rand's gen_range (Lemire's widening multiply with its
// rejection zone), drawn by one lane in the reference's order.  Candidate
// lists are sorted in LDS (bitonic), distances are computed one candidate per
// lane with the reference's f32 operations (eight lane sums for d >= 64), so
// a table equals the oracle's restatement entry for entry (minkowski aside:
// its powf is the device's).
#include <hip/hip_runtime.h>

#include <climits>

#include "knn.hpp"

namespace umapx {
namespace {

constexpr int kApproxWaves = 4;
constexpr int kCandCap = 2048;  // k + k^2 + 1 + explorations, rounded up (k <= 32)
constexpr int kKeyCap = 1024;   // max_candidates rounded up (candidate pool <= 255)

struct DevSmallRng {
  uint64_t s[4];
  __device__ explicit DevSmallRng(uint64_t state) {
    seed(state);
    if ((s[0] | s[1] | s[2] | s[3]) == 0) seed(0);
  }
  __device__ void seed(uint64_t state) {
    const uint64_t phi = 0x9e3779b97f4a7c15ULL;
    for (int i = 0; i < 4; ++i) {
      state += phi;
      uint64_t z = state;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      s[i] = z ^ (z >> 31);
    }
  }
  __device__ static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  __device__ uint64_t next() {
    const uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // gen_range(0..range) for usize
  __device__ uint64_t index(uint64_t range) {
    const uint64_t zone = (range << __clzll(range)) - 1;
    for (;;) {
      const uint64_t v = next();
      const uint64_t lo = v * range;
      if (lo <= zone) return __umul64hi(v, range);
    }
  }
};

// mix_approx_seed (lib.rs:3016-3023)
__device__ __forceinline__ uint64_t approx_seed(uint64_t seed, uint64_t row, uint64_t round) {
  uint64_t v = seed ^ (row * 0x9E3779B97F4A7C15ULL) ^ (round * 0xBF58476D1CE4E5B9ULL);
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
  return v ^ (v >> 31);
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ascending bitonic sort of a[0, P) (P a power of two) by one wave.
template <typename T>
__device__ void wave_sort(T* a, int P, int lane) {
  for (int size = 2; size <= P; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = lane; t < P / 2; t += 64) {
        const int lo = 2 * t - (t & (stride - 1));
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const T x = a[lo], y = a[hi];
        if ((x > y) == up) {
          a[lo] = y;
          a[hi] = x;
        }
      }
      wave_sync_lds();
    }
}

// squared_distance / dot / l2 sums (lib.rs:2445-2492, 2603-2699): sequential
// below 64 features; eight lane sums added in lane order, then the tail.
template <bool DOT>
__device__ float pair_sum(const float* __restrict__ x, const float* __restrict__ y, int d) {
  auto term = [](float a, float b) { return DOT ? fmul(a, b) : fmul(fsub(a, b), fsub(a, b)); };
  if (d < 64) {
    float acc = 0.f;
    for (int f = 0; f < d; ++f) acc = fadd(acc, term(x[f], y[f]));
    return acc;
  }
  float lanes[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int f = 0;
  for (; f + 8 <= d; f += 8) {
#pragma unroll
    for (int l = 0; l < 8; ++l) lanes[l] = fadd(lanes[l], term(x[f + l], y[f + l]));
  }
  float sum = 0.f;
#pragma unroll
  for (int l = 0; l < 8; ++l) sum = fadd(sum, lanes[l]);
  for (; f < d; ++f) sum = fadd(sum, term(x[f], y[f]));
  return sum;
}

// The other metrics one pair at a time in the reference's f32 operations
// (lib.rs:2494-2698): manhattan's eight lane sums for d >= 64, chebyshev's
// running max, minkowski (device powf), correlation's five sums, canberra,
// braycurtis.
__device__ float metric_pair(const float* __restrict__ a, const float* __restrict__ b, int d,
                             int metric, float p) {
  if (metric == 2) {
    if (d < 64) {
      float acc = 0.f;
      for (int f = 0; f < d; ++f) acc = fadd(acc, fabsf(fsub(a[f], b[f])));
      return acc;
    }
    float lanes[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int f = 0;
    for (; f + 8 <= d; f += 8) {
#pragma unroll
      for (int l = 0; l < 8; ++l) lanes[l] = fadd(lanes[l], fabsf(fsub(a[f + l], b[f + l])));
    }
    float sum = 0.f;
#pragma unroll
    for (int l = 0; l < 8; ++l) sum = fadd(sum, lanes[l]);
    for (; f < d; ++f) sum = fadd(sum, fabsf(fsub(a[f], b[f])));
    return sum;
  }
  if (metric == 3) {
    float m = 0.f;
    for (int f = 0; f < d; ++f) m = fmaxf(m, fabsf(fsub(a[f], b[f])));
    return m;
  }
  if (metric == 4) {
    float acc = 0.f;
    for (int f = 0; f < d; ++f) acc = fadd(acc, powf(fabsf(fsub(a[f], b[f])), p));
    return powf(acc, fdiv(1.f, p));
  }
  if (metric == 5) {
    float sx = 0.f, sy = 0.f, qx = 0.f, qy = 0.f, dt = 0.f;
    for (int f = 0; f < d; ++f) {
      sx = fadd(sx, a[f]);
      sy = fadd(sy, b[f]);
      qx = fadd(qx, fmul(a[f], a[f]));
      qy = fadd(qy, fmul(b[f], b[f]));
      dt = fadd(dt, fmul(a[f], b[f]));
    }
    const float nf = static_cast<float>(d);
    const float cdot = fsub(dt, fdiv(fmul(sx, sy), nf));
    const float xv = fmaxf(fsub(qx, fdiv(fmul(sx, sx), nf)), 0.f);
    const float yv = fmaxf(fsub(qy, fdiv(fmul(sy, sy), nf)), 0.f);
    const float xn = sqrtf(xv), yn = sqrtf(yv);
    if (xn == 0.f && yn == 0.f) return 0.f;
    if (xn == 0.f || yn == 0.f) return 1.f;
    float c = fdiv(cdot, fmul(xn, yn));
    c = c < -1.f ? -1.f : (c > 1.f ? 1.f : c);
    return fsub(1.f, c);
  }
  if (metric == 6) {
    float acc = 0.f;
    for (int f = 0; f < d; ++f) {
      const float den = fadd(fabsf(a[f]), fabsf(b[f]));
      if (den > 0.f) acc = fadd(acc, fdiv(fabsf(fsub(a[f], b[f])), den));
    }
    return acc;
  }
  float num = 0.f, den = 0.f;
  for (int f = 0; f < d; ++f) {
    num = fadd(num, fabsf(fsub(a[f], b[f])));
    den = fadd(den, fabsf(fadd(a[f], b[f])));
  }
  return den == 0.f ? 0.f : fdiv(num, den);
}

// euclidean (sqrt of the squared sum), cosine_distance_with_norms
// (lib.rs:2740-2765) or one of the other metrics
__device__ float approx_distance(const float* __restrict__ x, int64_t d, int metric,
                                 const float* __restrict__ norms, int64_t i, int64_t j, float p) {
  const float* a = x + i * d;
  const float* b = x + j * d;
  if (metric == 0) return sqrtf(pair_sum<false>(a, b, static_cast<int>(d)));
  if (metric >= 2) return metric_pair(a, b, static_cast<int>(d), metric, p);
  const float nx = norms[i], ny = norms[j];
  const float dt = pair_sum<true>(a, b, static_cast<int>(d));
  if (nx == 0.f && ny == 0.f) return 0.f;
  if (nx == 0.f || ny == 0.f) return 1.f;
  float c = fdiv(dt, fmul(nx, ny));
  c = c < -1.f ? -1.f : (c > 1.f ? 1.f : c);
  return fsub(1.f, c);
}

__device__ __forceinline__ uint64_t pair_key(float dist, int32_t idx) {
  const int b = __float_as_int(dist);
  const int key = b ^ static_cast<int>(static_cast<unsigned>(b >> 31) >> 1);  // total order
  return (static_cast<uint64_t>(static_cast<uint32_t>(key) ^ 0x80000000u) << 32) |
         static_cast<uint32_t>(idx);
}
__device__ __forceinline__ float key_dist(uint64_t k) {
  const int key = static_cast<int>(static_cast<uint32_t>(k >> 32) ^ 0x80000000u);
  return __int_as_float(key ^ static_cast<int>(static_cast<unsigned>(key >> 31) >> 1));
}

__global__ void row_norms_kernel(const float* __restrict__ x, int64_t n, int64_t d,
                                 float* __restrict__ norms) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) norms[i] = sqrtf(pair_sum<true>(x + i * d, x + i * d, static_cast<int>(d)));
}

// Distances of the wave's C candidates in cand[0, C) to row i, the k best by
// (distance, index) into the table row.  Returns whether the row changed.
__device__ bool finish_row(const float* __restrict__ x, int64_t d, int metric,
                           const float* __restrict__ norms, int64_t i, const int32_t* cand, int C,
                           uint64_t* keys, int k, const int32_t* __restrict__ old_idx,
                           const float* __restrict__ old_dist, int32_t* __restrict__ new_idx,
                           float* __restrict__ new_dist, int lane, float p) {
  int P = 64;
  while (P < C) P <<= 1;
  for (int c = lane; c < P; c += 64)
    keys[c] = c < C ? pair_key(approx_distance(x, d, metric, norms, i, cand[c], p), cand[c])
                    : ~uint64_t{0};
  wave_sync_lds();
  wave_sort(keys, P, lane);
  bool same = true;
  if (lane < k) {
    const uint64_t key = keys[lane];
    const int32_t id = static_cast<int32_t>(key & 0xffffffffu);
    const float dist = key_dist(key);
    new_idx[i * k + lane] = id;
    new_dist[i * k + lane] = dist;
    if (old_idx) same = old_idx[i * k + lane] == id && old_dist[i * k + lane] == dist;
  }
  return __ballot(!same) != 0ULL;
}

// sample_unique_candidates (lib.rs:3032-3043): the row itself, then draws
// until pool + 1 distinct candidates; their k nearest start the table.
__global__ __launch_bounds__(64 * kApproxWaves) void approx_init_kernel(
    const float* __restrict__ x, int64_t n, int64_t d, int metric, const float* __restrict__ norms,
    int k, int pool, uint64_t seed, int32_t* __restrict__ tab_idx, float* __restrict__ tab_dist,
    float p) {
  __shared__ int32_t s_cand[kApproxWaves][kKeyCap];
  __shared__ uint64_t s_keys[kApproxWaves][kKeyCap];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kApproxWaves + wv;
  if (i >= n) return;  // the whole wave; no block barriers follow
  int32_t* cand = s_cand[wv];
  if (lane == 0) {
    DevSmallRng rng(approx_seed(seed, static_cast<uint64_t>(i), 0));
    int cnt = 0;
    cand[cnt++] = static_cast<int32_t>(i);
    while (cnt < pool + 1) {
      const int32_t idx = static_cast<int32_t>(rng.index(static_cast<uint64_t>(n)));
      bool dup = false;
      for (int c = 0; c < cnt && !dup; ++c) dup = cand[c] == idx;
      if (!dup) cand[cnt++] = idx;
    }
  }
  wave_sync_lds();
  finish_row(x, d, metric, norms, i, cand, pool + 1, s_keys[wv], k, nullptr, nullptr, tab_idx,
             tab_dist, lane, p);
}

// One round (lib.rs:3046-3081 collect_approx_candidate_indices, 3120-3140).
__global__ __launch_bounds__(64 * kApproxWaves) void approx_round_kernel(
    const float* __restrict__ x, int64_t n, int64_t d, int metric, const float* __restrict__ norms,
    int k, int max_cand, int exploration, uint64_t seed, int round,
    const int32_t* __restrict__ old_idx, const float* __restrict__ old_dist,
    int32_t* __restrict__ new_idx, float* __restrict__ new_dist, int32_t* __restrict__ changed,
    float p) {
  __shared__ int32_t s_cand[kApproxWaves][kCandCap];
  __shared__ uint64_t s_keys[kApproxWaves][kKeyCap];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kApproxWaves + wv;
  if (i >= n) return;  // the whole wave; no block barriers follow
  int32_t* cand = s_cand[wv];
  const int raw = k + k * k + 1 + exploration;
  int P = 64;
  while (P < raw) P <<= 1;
  // the row's neighbours, their neighbours, itself; the rest empty
  for (int c = lane; c < P; c += 64) {
    int32_t v = INT_MAX;
    if (c < k) {
      v = old_idx[i * k + c];
    } else if (c < k + k * k) {
      const int a = (c - k) / k, b = (c - k) - a * k;
      v = old_idx[static_cast<int64_t>(old_idx[i * k + a]) * k + b];
    } else if (c == k + k * k) {
      v = static_cast<int32_t>(i);
    }
    cand[c] = v;
  }
  DevSmallRng rng(approx_seed(seed, static_cast<uint64_t>(i), static_cast<uint64_t>(round) + 1));
  if (lane == 0)
    for (int e = 0; e < exploration; ++e)
      cand[k + k * k + 1 + e] = static_cast<int32_t>(rng.index(static_cast<uint64_t>(n)));
  wave_sync_lds();
  wave_sort(cand, P, lane);
  // drop repeats: each lane compacts a contiguous chunk, offsets by prefix count
  const int chunk = P / 64;
  int32_t vals[kCandCap / 64];
  int keep = 0;
#pragma unroll
  for (int q = 0; q < kCandCap / 64; ++q) {
    if (q >= chunk) break;
    const int c = lane * chunk + q;
    const int32_t v = cand[c];
    vals[q] = v;
    if (v != INT_MAX && (c == 0 || cand[c - 1] != v)) ++keep;
  }
  int before = keep;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(before, off);
    if (lane >= off) before += o;
  }
  const int total = __shfl(before, 63);
  before -= keep;
  int32_t prev = lane == 0 ? INT_MIN : cand[lane * chunk - 1];
  wave_sync_lds();
#pragma unroll
  for (int q = 0; q < kCandCap / 64; ++q) {
    if (q >= chunk) break;
    const int32_t v = vals[q];
    if (v != INT_MAX && v != prev) cand[before++] = v;
    prev = v;
  }
  wave_sync_lds();
  // cut to max_candidates by a partial shuffle on the same stream
  if (total > max_cand && lane == 0) {
    for (int s = 0; s < max_cand; ++s) {
      const int r = s + static_cast<int>(rng.index(static_cast<uint64_t>(total - s)));
      const int32_t t = cand[s];
      cand[s] = cand[r];
      cand[r] = t;
    }
  }
  wave_sync_lds();
  const int C = total < max_cand ? total : max_cand;
  if (finish_row(x, d, metric, norms, i, cand, C, s_keys[wv], k, old_idx, old_dist, new_idx,
                 new_dist, lane, p) &&
      lane == 0)
    atomicOr(changed, 1);
}

}  // namespace

void knn_approx_device(const float* x, int64_t n, int64_t d, int k, int metric, float metric_p,
                       int candidate_pool, int n_iters, uint64_t seed, int32_t* out_idx,
                       float* out_dist, KnnWorkspace& ws, hipStream_t s) {
  if (metric < 0 || metric > 7) throw_invalid("unknown metric code");
  if (k < 1 || k > 32 || k >= n) throw_invalid("n_neighbors must be <= 32 and below the sample count");
  if (candidate_pool < 1 || candidate_pool > 255)
    throw_invalid("approx_knn_candidates must be in [1, 255] on the GPU path");
  if (n >= (int64_t{1} << 31) - 1) throw_invalid("approximate kNN supports fewer than 2^31 rows");
  const int pool = static_cast<int>(std::min<int64_t>(std::max(candidate_pool, k), n - 1));
  const int max_cand = std::max(candidate_pool, k) * 4 + 1;
  const int exploration = std::min(std::max(k / 2, 4), max_cand);
  const float* norms = nullptr;
  if (metric == 1) {
    ws.msum.ensure(static_cast<size_t>(n));
    hipLaunchKernelGGL(row_norms_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, x, n, d,
                       ws.msum.get());
    KERNEL_CHECK();
    norms = ws.msum.get();
  }
  DevBuf<int32_t> idx2(static_cast<size_t>(n) * k), flag(1);
  DevBuf<float> dist2(static_cast<size_t>(n) * k);
  int32_t* cur_i = out_idx;
  float* cur_d = out_dist;
  int32_t* nxt_i = idx2.get();
  float* nxt_d = dist2.get();
  const dim3 grid(blocks_for(n, kApproxWaves)), block(64 * kApproxWaves);
  hipLaunchKernelGGL(approx_init_kernel, grid, block, 0, s, x, n, d, metric, norms, k, pool, seed,
                     cur_i, cur_d, metric_p);
  KERNEL_CHECK();
  for (int round = 0; round < n_iters; ++round) {
    HIP_CHECK(hipMemsetAsync(flag.get(), 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(approx_round_kernel, grid, block, 0, s, x, n, d, metric, norms, k, max_cand,
                       exploration, seed, round, cur_i, cur_d, nxt_i, nxt_d, flag.get(), metric_p);
    KERNEL_CHECK();
    int32_t changed = 0;
    HIP_CHECK(hipMemcpyAsync(&changed, flag.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::swap(cur_i, nxt_i);
    std::swap(cur_d, nxt_d);
    if (!changed) break;
  }
  if (cur_i != out_idx) {
    HIP_CHECK(hipMemcpyAsync(out_idx, cur_i, sizeof(int32_t) * n * k, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(out_dist, cur_d, sizeof(float) * n * k, hipMemcpyDeviceToDevice, s));
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace umapx
