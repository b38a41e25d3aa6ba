// Exact kNN under the reference's non-inner-product metrics (manhattan,
// chebyshev, minkowski, correlation, canberra, braycurtis):
// exact_nearest_neighbors_by / exact_nearest_neighbors_to_reference_by
// (lib.rs:2891-2928, 3346-3399) with the distances of lib.rs:2494-2698.
//
// These distances are not inner products, so they run on the vector ALUs,
// not on MFMA.  One wave per block; lane l owns query l of the block and keeps
// its best H (distance, index) pairs sorted in registers.  The block's 64
// query rows sit in LDS transposed ([feature / 4][lane][4] when d is a
// multiple of 4, so each lane reads its 4 features as one 16-byte word and
// the wave a contiguous 1 KB), and the reference rows stream through LDS 16
// at a time (row-major, so every lane reads the same 16 bytes: a broadcast).
// Each lane evaluates its query against the 16 rows with 16 independent
// accumulators, feature by feature in the reference's order and without
// contraction, so every distance except minkowski's equals the reference's
// bit for bit: manhattan keeps the eight AVX2 lane sums of lib.rs:2518-2541
// for d >= 64, correlation takes its per-row sums from a separate pass and
// the pair's dot product in feature order.  Minkowski raises each |x - y| to
// p with the hardware log2 / exp2, so its distances agree to a tolerance.
#include <hip/hip_runtime.h>

#include <climits>

#include "knn.hpp"

namespace umapx {
namespace {

constexpr int kMetricRows = 16;  // reference rows per LDS tile

// f32::total_cmp key and its inverse (the map is an involution).
__device__ __forceinline__ int total_key_f(float v) {
  const int b = __float_as_int(v);
  return b ^ static_cast<int>(static_cast<unsigned>(b >> 31) >> 1);
}
__device__ __forceinline__ float key_to_float(int key) {
  return __int_as_float(key ^ static_cast<int>(static_cast<unsigned>(key >> 31) >> 1));
}

// Rust's f32::clamp: NaN passes through.
__device__ __forceinline__ float clamp_pm1(float c) { return c < -1.f ? -1.f : (c > 1.f ? 1.f : c); }

// |x|^p for the minkowski sum from the hardware log2 / exp2 (about 1 ulp
// each); minkowski distances are compared to a tolerance, not bit for bit.
__device__ __forceinline__ float abs_pow(float v, float p) {
  const float a = fabsf(v);
  return a == 0.f ? 0.f : __builtin_amdgcn_exp2f(p * __builtin_amdgcn_logf(a));
}

// One feature of one (query, row) pair, in the reference's f32 operations.
template <int METRIC>
__device__ __forceinline__ void pair_step(float& acc, float& acc2, float xv, float yv, float p) {
  if (METRIC == 0) {
    const float df = fsub(xv, yv);
    acc = fadd(acc, fmul(df, df));
  } else if (METRIC == 1) {
    acc = fadd(acc, fmul(xv, yv));
  } else if (METRIC == 2) {
    acc = fadd(acc, fabsf(fsub(xv, yv)));
  } else if (METRIC == 3) {
    acc = fmaxf(acc, fabsf(fsub(xv, yv)));
  } else if (METRIC == 4) {
    acc = fadd(acc, abs_pow(fsub(xv, yv), p));
  } else if (METRIC == 5) {
    acc = fadd(acc, fmul(xv, yv));
  } else if (METRIC == 6) {
    const float den = fadd(fabsf(xv), fabsf(yv));
    if (den > 0.f) acc = fadd(acc, fdiv(fabsf(fsub(xv, yv)), den));
  } else {
    acc = fadd(acc, fabsf(fsub(xv, yv)));
    acc2 = fadd(acc2, fabsf(fadd(xv, yv)));
  }
}

// One feature's term of the metrics summed in eight lanes for d >= 64:
// squared euclidean (lib.rs:2440-2492), dot product (lib.rs:2700-2737) and
// manhattan (lib.rs:2518-2541).
template <int METRIC>
__device__ __forceinline__ float lane_term(float x, float y) {
  if (METRIC == 0) {
    const float df = fsub(x, y);
    return fmul(df, df);
  }
  if (METRIC == 1) return fmul(x, y);
  return fabsf(fsub(x, y));
}

// L2 norms of rows, the reference's dot product of a row with itself then
// sqrt (eight lane sums for d >= 64; lib.rs:2700-2765).
__global__ void row_norms_kernel(const float* __restrict__ x, int64_t rows, int d,
                                 float* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const float* r = x + i * d;
  float s = 0.f;
  if (d >= 64) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int f = 0;
    for (; f + 8 <= d; f += 8)
      for (int l = 0; l < 8; ++l) acc[l] = fadd(acc[l], fmul(r[f + l], r[f + l]));
    for (int l = 0; l < 8; ++l) s = fadd(s, acc[l]);
    for (; f < d; ++f) s = fadd(s, fmul(r[f], r[f]));
  } else {
    for (int f = 0; f < d; ++f) s = fadd(s, fmul(r[f], r[f]));
  }
  out[i] = fsqrt(s);
}

// Per-row sum and sum of squares in feature order (the row halves of
// correlation_distance, lib.rs:2655-2676).
__global__ void row_summaries_kernel(const float* __restrict__ x, int64_t rows, int d,
                                     float* __restrict__ sum, float* __restrict__ sq) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const float* r = x + i * d;
  float s = 0.f, q = 0.f;
  for (int f = 0; f < d; ++f) {
    s = fadd(s, r[f]);
    q = fadd(q, fmul(r[f], r[f]));
  }
  sum[i] = s;
  sq[i] = q;
}

// METRIC: 0 euclidean (squared, the root taken at output), 1 cosine, 2
// manhattan, 3 chebyshev, 4 minkowski, 5 correlation, 6 canberra, 7
// braycurtis.  H: kept candidates per query (>= k; 64 serves n_neighbors of
// 33-64 under every metric, euclidean and cosine included).  QG: the transposed
// query tile lives in the block's slice of qglob (HBM, read through the
// caches) instead of LDS, for rows too wide for both tiles in LDS.
template <int METRIC, int H, bool QG>
__global__ __launch_bounds__(64) void knn_metric_kernel(
    const float* __restrict__ q, int64_t m, const float* __restrict__ r, int64_t n, int d, int k,
    float p, const float* __restrict__ qsum, const float* __restrict__ qsq,
    const float* __restrict__ rsum, const float* __restrict__ rsq, int32_t* __restrict__ out_idx,
    float* __restrict__ out_dist, float* __restrict__ qglob) {
  extern __shared__ float lds[];
  // [d / 4][64][4], or [d][64] when d % 4 != 0
  float* qt = QG ? qglob + static_cast<int64_t>(blockIdx.x) * 64 * d : lds;
  float* rt = QG ? lds : lds + 64 * d;  // [kMetricRows][d]
  const int lane = threadIdx.x;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * 64;
  const int64_t i = q0 + lane;
  const bool act = i < m;
  const int qrows = static_cast<int>(m - q0 < 64 ? m - q0 : 64);
  const bool v4 = (d & 3) == 0;
  for (int e = lane; e < 64 * d; e += 64) {
    const int row = e / d, f = e - row * d;
    const float v = row < qrows ? q[(q0 + row) * d + f] : 0.f;
    if (v4)
      qt[((f >> 2) * 64 + row) * 4 + (f & 3)] = v;
    else
      qt[f * 64 + row] = v;
  }
  const float4* qt4 = reinterpret_cast<const float4*>(qt);
  const float4* rt4 = reinterpret_cast<const float4*>(rt);
  const int d4 = d >> 2;
  float xs = 0.f, xq = 0.f;
  if (METRIC == 5 && act) {
    xs = qsum[i];
    xq = qsq[i];
  }
  if (METRIC == 1 && act) xq = qsq[i];  // the query row's L2 norm
  int keys[H];
  int32_t ids[H];
#pragma unroll
  for (int s = 0; s < H; ++s) {
    keys[s] = INT_MAX;
    ids[s] = INT_MAX;
  }
  const int nb = d / 8;
  for (int64_t r0 = 0; r0 < n; r0 += kMetricRows) {
    const int rows = static_cast<int>(n - r0 < kMetricRows ? n - r0 : kMetricRows);
    __syncthreads();
    for (int e = lane; e < kMetricRows * d; e += 64) {
      const int row = e / d;
      rt[e] = row < rows ? r[r0 * d + e] : 0.f;
    }
    __syncthreads();
    float dist[kMetricRows];
    if ((METRIC == 0 || METRIC == 1 || METRIC == 2) && d >= 64 && v4) {
      // the eight lane sums as two float4 halves of each 8-feature block,
      // four rows at a time; the lanes added in order, then the tail
      constexpr int kPass = 4;
      const int nb8 = d >> 3;
#pragma unroll
      for (int t0 = 0; t0 < kMetricRows; t0 += kPass) {
        float4 lo[kPass], hi[kPass];
#pragma unroll
        for (int t = 0; t < kPass; ++t) {
          lo[t] = make_float4(0.f, 0.f, 0.f, 0.f);
          hi[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        for (int b = 0; b < nb8; ++b) {
          const float4 xa = qt4[(2 * b) * 64 + lane], xb = qt4[(2 * b + 1) * 64 + lane];
#pragma unroll
          for (int t = 0; t < kPass; ++t) {
            const float4 ya = rt4[(t0 + t) * d4 + 2 * b], yb = rt4[(t0 + t) * d4 + 2 * b + 1];
            lo[t].x = fadd(lo[t].x, lane_term<METRIC>(xa.x, ya.x));
            lo[t].y = fadd(lo[t].y, lane_term<METRIC>(xa.y, ya.y));
            lo[t].z = fadd(lo[t].z, lane_term<METRIC>(xa.z, ya.z));
            lo[t].w = fadd(lo[t].w, lane_term<METRIC>(xa.w, ya.w));
            hi[t].x = fadd(hi[t].x, lane_term<METRIC>(xb.x, yb.x));
            hi[t].y = fadd(hi[t].y, lane_term<METRIC>(xb.y, yb.y));
            hi[t].z = fadd(hi[t].z, lane_term<METRIC>(xb.z, yb.z));
            hi[t].w = fadd(hi[t].w, lane_term<METRIC>(xb.w, yb.w));
          }
        }
#pragma unroll
        for (int t = 0; t < kPass; ++t) {
          float sum = 0.f;
          sum = fadd(sum, lo[t].x);
          sum = fadd(sum, lo[t].y);
          sum = fadd(sum, lo[t].z);
          sum = fadd(sum, lo[t].w);
          sum = fadd(sum, hi[t].x);
          sum = fadd(sum, hi[t].y);
          sum = fadd(sum, hi[t].z);
          sum = fadd(sum, hi[t].w);
          if (nb8 * 2 < d4) {  // a 4-feature tail
            const float4 xa = qt4[(2 * nb8) * 64 + lane];
            const float4 ya = rt4[(t0 + t) * d4 + 2 * nb8];
            sum = fadd(sum, lane_term<METRIC>(xa.x, ya.x));
            sum = fadd(sum, lane_term<METRIC>(xa.y, ya.y));
            sum = fadd(sum, lane_term<METRIC>(xa.z, ya.z));
            sum = fadd(sum, lane_term<METRIC>(xa.w, ya.w));
          }
          dist[t0 + t] = sum;
        }
      }
    } else if ((METRIC == 0 || METRIC == 1 || METRIC == 2) && d >= 64) {
      // eight lane sums over the full 8-feature blocks, added in lane order,
      // then the tail one feature at a time
      float tot[kMetricRows];
#pragma unroll
      for (int t = 0; t < kMetricRows; ++t) tot[t] = 0.f;
      for (int l = 0; l < 8; ++l) {
        float acc[kMetricRows];
#pragma unroll
        for (int t = 0; t < kMetricRows; ++t) acc[t] = 0.f;
        for (int b = 0; b < nb; ++b) {
          const int f = 8 * b + l;
          const float xv = qt[f * 64 + lane];
#pragma unroll
          for (int t = 0; t < kMetricRows; ++t) acc[t] = fadd(acc[t], lane_term<METRIC>(xv, rt[t * d + f]));
        }
#pragma unroll
        for (int t = 0; t < kMetricRows; ++t) tot[t] = fadd(tot[t], acc[t]);
      }
      for (int f = 8 * nb; f < d; ++f) {
        const float xv = qt[f * 64 + lane];
#pragma unroll
        for (int t = 0; t < kMetricRows; ++t) tot[t] = fadd(tot[t], lane_term<METRIC>(xv, rt[t * d + f]));
      }
#pragma unroll
      for (int t = 0; t < kMetricRows; ++t) dist[t] = tot[t];
    } else {
      float acc[kMetricRows], acc2[kMetricRows];
#pragma unroll
      for (int t = 0; t < kMetricRows; ++t) {
        acc[t] = 0.f;
        acc2[t] = 0.f;
      }
      if (v4) {
        for (int f4 = 0; f4 < d4; ++f4) {
          const float4 xv = qt4[f4 * 64 + lane];
#pragma unroll
          for (int t = 0; t < kMetricRows; ++t) {
            const float4 yv = rt4[t * d4 + f4];
            pair_step<METRIC>(acc[t], acc2[t], xv.x, yv.x, p);
            pair_step<METRIC>(acc[t], acc2[t], xv.y, yv.y, p);
            pair_step<METRIC>(acc[t], acc2[t], xv.z, yv.z, p);
            pair_step<METRIC>(acc[t], acc2[t], xv.w, yv.w, p);
          }
        }
      } else {
        for (int f = 0; f < d; ++f) {
          const float xv = qt[f * 64 + lane];
#pragma unroll
          for (int t = 0; t < kMetricRows; ++t) pair_step<METRIC>(acc[t], acc2[t], xv, rt[t * d + f], p);
        }
      }
#pragma unroll
      for (int t = 0; t < kMetricRows; ++t) {
        if (METRIC == 4) {
          dist[t] = powf(acc[t], fdiv(1.f, p));
        } else if (METRIC == 5) {
          // correlation_distance_from_summaries (lib.rs:2678-2698)
          const float nf = static_cast<float>(d);
          const int64_t rj = r0 + (t < rows ? t : 0);
          const float ys = rsum[rj], yq = rsq[rj];
          const float cdot = fsub(acc[t], fdiv(fmul(xs, ys), nf));
          const float xvar = fmaxf(fsub(xq, fdiv(fmul(xs, xs), nf)), 0.f);
          const float yvar = fmaxf(fsub(yq, fdiv(fmul(ys, ys), nf)), 0.f);
          const float xn = sqrtf(xvar), yn = sqrtf(yvar);
          if (xn == 0.f && yn == 0.f)
            dist[t] = 0.f;
          else if (xn == 0.f || yn == 0.f)
            dist[t] = 1.f;
          else
            dist[t] = fsub(1.f, clamp_pm1(fdiv(cdot, fmul(xn, yn))));
        } else if (METRIC == 7) {
          dist[t] = acc2[t] == 0.f ? 0.f : fdiv(acc[t], acc2[t]);
        } else {
          dist[t] = acc[t];
        }
      }
    }
    if (METRIC == 1) {  // cosine_distance_with_norms (lib.rs:2739-2765) on the dot
#pragma unroll
      for (int t = 0; t < kMetricRows; ++t) {
        const float ny = rsq[r0 + (t < rows ? t : 0)];
        if (xq == 0.f && ny == 0.f)
          dist[t] = 0.f;
        else if (xq == 0.f || ny == 0.f)
          dist[t] = 1.f;
        else
          dist[t] = fsub(1.f, clamp_pm1(fdiv(dist[t], fmul(xq, ny))));
      }
    }
    // top-H by (distance total order, index), rows in increasing index order;
    // euclidean ranks the squared distances (lib.rs:2829-2834)
#pragma unroll
    for (int t = 0; t < kMetricRows; ++t) {
      if (t >= rows) break;
      const int key = total_key_f(dist[t]);
      const int32_t id = static_cast<int32_t>(r0 + t);
      if (key < keys[H - 1] || (key == keys[H - 1] && id < ids[H - 1])) {
        keys[H - 1] = key;
        ids[H - 1] = id;
#pragma unroll
        for (int s = H - 1; s > 0; --s) {
          const bool sw = keys[s] < keys[s - 1] || (keys[s] == keys[s - 1] && ids[s] < ids[s - 1]);
          if (sw) {
            const int tk = keys[s];
            keys[s] = keys[s - 1];
            keys[s - 1] = tk;
            const int32_t ti = ids[s];
            ids[s] = ids[s - 1];
            ids[s - 1] = ti;
          }
        }
      }
    }
  }
  if (!act) return;
#pragma unroll
  for (int s = 0; s < H; ++s)
    if (s < k) {
      out_idx[i * k + s] = ids[s];
      out_dist[i * k + s] = METRIC == 0 ? fsqrt(key_to_float(keys[s])) : key_to_float(keys[s]);
    }
}

// Rows up to kMetricLdsFeatures keep both tiles in LDS; wider rows (up to
// kMetricMaxFeatures, the reference tile alone in LDS) keep the query tile
// in HBM.
constexpr int kMetricLdsFeatures = 160 * 1024 / ((64 + kMetricRows) * 4);  // 512
constexpr int kMetricMaxFeatures = 160 * 1024 / (kMetricRows * 4);         // 2560

template <int METRIC, int H>
void launch_metric(const float* q, int64_t m, const float* r, int64_t n, int d, int k, float p,
                   const float* qs, const float* qq, const float* rs, const float* rq,
                   int32_t* out_idx, float* out_dist, float* qglob, hipStream_t s) {
  const bool qg = d > kMetricLdsFeatures;
  const size_t bytes = static_cast<size_t>(qg ? kMetricRows : 64 + kMetricRows) * d * sizeof(float);
  auto kern = qg ? knn_metric_kernel<METRIC, H, true> : knn_metric_kernel<METRIC, H, false>;
  HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
  hipLaunchKernelGGL(kern, dim3(blocks_for(m, 64)), dim3(64), bytes, s, q, m, r, n, d, k, p, qs,
                     qq, rs, rq, out_idx, out_dist, qglob);
  KERNEL_CHECK();
}

template <int H>
void dispatch_metric(int metric, const float* q, int64_t m, const float* r, int64_t n, int d,
                     int k, float p, const float* qs, const float* qq, const float* rs,
                     const float* rq, int32_t* out_idx, float* out_dist, float* qg,
                     hipStream_t s) {
  switch (metric) {
    case 0: launch_metric<0, H>(q, m, r, n, d, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s); break;
    case 1: launch_metric<1, H>(q, m, r, n, d, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s); break;
    case 2: launch_metric<2, H>(q, m, r, n, d, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s); break;
    case 3: launch_metric<3, H>(q, m, r, n, d, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s); break;
    case 4: launch_metric<4, H>(q, m, r, n, d, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s); break;
    case 5: launch_metric<5, H>(q, m, r, n, d, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s); break;
    case 6: launch_metric<6, H>(q, m, r, n, d, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s); break;
    default: launch_metric<7, H>(q, m, r, n, d, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s); break;
  }
}

}  // namespace

void knn_metric_device(const float* q, int64_t m, const float* r, int64_t n, int64_t d, int k,
                       int metric, float p, int32_t* out_idx, float* out_dist, KnnWorkspace& ws,
                       hipStream_t s) {
  if (metric < 0 || metric > 7) throw_invalid("unknown metric code");
  if (d < 1 || d > kMetricMaxFeatures)
    throw_invalid(metric < 2 ? "exact GPU kNN with n_neighbors > 32 supports at most 2560 features"
                             : "exact GPU kNN under this metric supports at most 2560 features");
  if (k < 1 || k > 64 || k > n) throw_invalid("n_neighbors must be <= 64 on the GPU path");
  if (n >= (int64_t{1} << 31)) throw_invalid("exact GPU kNN supports fewer than 2^31 rows");
  if (m == 0) return;
  const float* qs = nullptr;
  const float* qq = nullptr;
  const float* rs = nullptr;
  const float* rq = nullptr;
  if (metric == 5) {
    ws.msum.ensure(static_cast<size_t>(2 * (n + m)));
    float* base = ws.msum.get();
    hipLaunchKernelGGL(row_summaries_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, r, n,
                       static_cast<int>(d), base, base + n);
    hipLaunchKernelGGL(row_summaries_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, s, q, m,
                       static_cast<int>(d), base + 2 * n, base + 2 * n + m);
    KERNEL_CHECK();
    rs = base;
    rq = base + n;
    qs = base + 2 * n;
    qq = base + 2 * n + m;
  } else if (metric == 1) {  // L2 norms of both row sets
    ws.msum.ensure(static_cast<size_t>(n + m));
    float* base = ws.msum.get();
    hipLaunchKernelGGL(row_norms_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, r, n,
                       static_cast<int>(d), base);
    hipLaunchKernelGGL(row_norms_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, s, q, m,
                       static_cast<int>(d), base + n);
    KERNEL_CHECK();
    rq = base;
    qq = base + n;
  }
  const int dd = static_cast<int>(d);
  float* qg = nullptr;  // per-block transposed query tiles of wide rows
  if (dd > kMetricLdsFeatures) {
    ws.scores.ensure(static_cast<size_t>(blocks_for(m, 64)) * 64 * dd);
    qg = ws.scores.get();
  }
  if (k <= 16)
    dispatch_metric<16>(metric, q, m, r, n, dd, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s);
  else if (k <= 32)
    dispatch_metric<32>(metric, q, m, r, n, dd, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s);
  else
    dispatch_metric<64>(metric, q, m, r, n, dd, k, p, qs, qq, rs, rq, out_idx, out_dist, qg, s);
}

}  // namespace umapx
