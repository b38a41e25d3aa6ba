// Inverse transform: embedded points back to the training feature space
// (inverse_transform_rows, lib.rs:1800-1915).
//
// 1. kNN of the embedded queries among the fitted embedding (the exact FP32
//    kNN kernel with n_components features), k = min(n_neighbors, n).
// 2. inverse_weights_kernel: w = 1 / (1 + a d^2b) per neighbour, the
//    per-row normalised copy, the fixed rows (nearest at distance <= 1e-6)
//    and the largest weight of all edges (lib.rs:1856-1890).
// 3. inverse_init_kernel: the local linear initialisation
//    (init_inverse_points, lib.rs:4838-4935).  One wave per query, lanes over
//    the features; the weighted means, the small Gram matrix and its LU are
//    wave-uniform and the right-hand side is solved lane by lane.  The f32
//    operations follow the reference's order without contraction, so the
//    result equals the oracle's up to the last bit of powf.
// 4. inverse_sgd_kernel: optimize_layout_inverse (lib.rs:5547-5645), a group
//    of 16-64 lanes per query moving that query's row in feature space
//    (Gauss-Seidel over its own edges, as the reference's edge order does for
//    one head).  The
//    per-edge schedule is the reference's exact f32 one; negatives come from
//    a counter-based stream per (query, epoch, edge) instead of the
//    reference's single SmallRng stream.
#include "inverse.hpp"

#include <hip/hip_runtime.h>

#include <cmath>

namespace umapx {
namespace {

constexpr int kInvWaves = 4;     // queries (waves) per block
constexpr int kMaxLowDim = 8;    // largest n_components

__device__ __forceinline__ uint64_t inv_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Uniform index in [0, n) for draw s of stream key (multiply-shift on the top
// 32 bits; n < 2^31).
__device__ __forceinline__ uint32_t inv_draw(uint64_t key, uint32_t s, uint32_t n) {
  const uint64_t h = inv_mix64(key + (static_cast<uint64_t>(s) + 1ULL) * 0x9e3779b97f4a7c15ULL);
  return static_cast<uint32_t>(((h >> 32) * static_cast<uint64_t>(n)) >> 32);
}

__device__ __forceinline__ float inv_clip4(float v) { return fminf(fmaxf(v, -4.f), 4.f); }

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ float lane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__global__ __launch_bounds__(256) void inverse_weights_kernel(
    const float* __restrict__ dist, int64_t m, int k, float a, float two_b,
    float* __restrict__ raw, float* __restrict__ norm, int32_t* __restrict__ fixed,
    uint32_t* __restrict__ wmax_bits) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  float mx = 0.f;
  if (i < m) {
    const float* dr = dist + i * k;
    float row_sum = 0.f;
    for (int j = 0; j < k; ++j) {
      const float w = fdiv(1.f, fadd(1.f, fmul(a, powf(dr[j], two_b))));
      const bool ok = w > 0.f && isfinite(w);
      raw[i * k + j] = ok ? w : 0.f;
      if (ok) {
        row_sum = fadd(row_sum, w);
        mx = fmaxf(mx, w);
      }
    }
    for (int j = 0; j < k; ++j) {
      const float r = raw[i * k + j];
      norm[i * k + j] = row_sum > 0.f ? fdiv(r, row_sum) : r;
    }
    fixed[i] = dr[0] <= 1e-6f ? 1 : 0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  // weights are >= 0, so their bit patterns order as unsigned integers
  if ((threadIdx.x & 63) == 0 && mx > 0.f) atomicMax(wmax_bits, __float_as_uint(mx));
}

// Partial-pivot LU of the wave-uniform dim x dim matrix held in L (row-major,
// in place), in the steps of nalgebra 0.33's LU::new: pivot = the first row
// of largest |.| in the column, a zero pivot skips the column, the column
// below the pivot is scaled by 1 / pivot and the trailing rows are updated
// as (-pivot_row[c]) * l + row[c].  pv[i] is the row swapped with row i (or
// -1).  Every index is a compile-time constant so L stays in registers.
template <int MAXD>
__device__ __forceinline__ void small_lu(float (&L)[MAXD][MAXD], int (&pv)[MAXD], int dim) {
#pragma unroll
  for (int i = 0; i < MAXD; ++i) {
    pv[i] = -1;
    if (i >= dim) continue;
    int piv = i;
    float best = fabsf(L[i][i]);
#pragma unroll
    for (int r = i + 1; r < MAXD; ++r)
      if (r < dim && fabsf(L[r][i]) > best) {
        best = fabsf(L[r][i]);
        piv = r;
      }
    float diag = L[i][i];
#pragma unroll
    for (int r = i + 1; r < MAXD; ++r)
      if (r == piv) diag = L[r][i];
    if (diag == 0.f) continue;
    if (piv != i) {
      pv[i] = piv;
#pragma unroll
      for (int r = i + 1; r < MAXD; ++r)
        if (r == piv) {
#pragma unroll
          for (int c = 0; c < MAXD; ++c) {
            const float t = L[i][c];
            L[i][c] = L[r][c];
            L[r][c] = t;
          }
        }
    }
    const float inv = fdiv(1.f, diag);
#pragma unroll
    for (int r = i + 1; r < MAXD; ++r)
      if (r < dim) L[r][i] = fmul(L[r][i], inv);
#pragma unroll
    for (int c = i + 1; c < MAXD; ++c) {
      if (c >= dim) continue;
      const float na = -L[i][c];
#pragma unroll
      for (int r = i + 1; r < MAXD; ++r)
        if (r < dim) L[r][c] = fadd(fmul(na, L[r][i]), L[r][c]);
    }
  }
}

// LU::solve for one right-hand side: the row swaps in order, the unit-lower
// then the upper substitution (column axpy form).  The caller has checked
// that U has no zero diagonal.
template <int MAXD>
__device__ __forceinline__ void small_lu_solve(const float (&L)[MAXD][MAXD], const int (&pv)[MAXD],
                                               int dim, float (&x)[MAXD]) {
#pragma unroll
  for (int i = 0; i < MAXD; ++i) {
    if (i >= dim || pv[i] < 0) continue;
#pragma unroll
    for (int r = i + 1; r < MAXD; ++r)
      if (r == pv[i]) {
        const float t = x[i];
        x[i] = x[r];
        x[r] = t;
      }
  }
#pragma unroll
  for (int i = 0; i + 1 < MAXD; ++i) {
    if (i + 1 >= dim) continue;
    const float nc = -x[i];
#pragma unroll
    for (int r = i + 1; r < MAXD; ++r)
      if (r < dim) x[r] = fadd(fmul(nc, L[r][i]), x[r]);
  }
#pragma unroll
  for (int i = MAXD - 1; i >= 0; --i) {
    if (i >= dim) continue;
    const float coeff = fdiv(x[i], L[i][i]);
    x[i] = coeff;
#pragma unroll
    for (int r = 0; r < i; ++r) x[r] = fadd(fmul(-coeff, L[r][i]), x[r]);
  }
}

// Skip rule of init_inverse_points: weight <= 0 or not finite.
__device__ __forceinline__ bool live(float w) { return w > 0.f && isfinite(w); }

template <int DIM>
__global__ __launch_bounds__(64 * kInvWaves) void inverse_init_kernel(
    const float* __restrict__ q, int64_t m, int dim_rt, const int32_t* __restrict__ idx,
    const float* __restrict__ norm, const int32_t* __restrict__ fixed, int k,
    const float* __restrict__ y_train, const float* __restrict__ x_train, int64_t d,
    float* __restrict__ out) {
  constexpr int MAXD = DIM > 0 ? DIM : kMaxLowDim;
  const int dim = DIM > 0 ? DIM : dim_rt;
  const int lane = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kInvWaves + (threadIdx.x >> 6);
  if (i >= m) return;  // the whole wave
  const int32_t* ir = idx + i * k;
  const float* wr = norm + i * k;
  float* o = out + i * d;
  if (fixed[i]) {
    const float* src = x_train + static_cast<int64_t>(ir[0]) * d;
    for (int64_t f = lane; f < d; f += 64) o[f] = src[f];
    return;
  }
  // wave-uniform: weighted embedding mean, weight sum, Gram matrix
  float zb[MAXD];
#pragma unroll
  for (int c = 0; c < MAXD; ++c) zb[c] = 0.f;
  float wsum = 0.f;
  for (int j = 0; j < k; ++j) {
    const float w = wr[j];
    if (!live(w)) continue;
    wsum = fadd(wsum, w);
    const float* yr = y_train + static_cast<int64_t>(ir[j]) * dim;
#pragma unroll
    for (int c = 0; c < MAXD; ++c)
      if (c < dim) zb[c] = fadd(zb[c], fmul(w, yr[c]));
  }
  if (!(wsum > 0.f)) {
    for (int64_t f = lane; f < d; f += 64) o[f] = __int_as_float(0x7fc00000);
    return;
  }
  float L[MAXD][MAXD];
#pragma unroll
  for (int r = 0; r < MAXD; ++r)
#pragma unroll
    for (int c = 0; c < MAXD; ++c) L[r][c] = 0.f;
  int nz = 0;
  for (int j = 0; j < k; ++j) {
    const float w = wr[j];
    if (!live(w)) continue;
    ++nz;
    const float* yr = y_train + static_cast<int64_t>(ir[j]) * dim;
#pragma unroll
    for (int a = 0; a < MAXD; ++a) {
      if (a >= dim) continue;
      const float za = fsub(yr[a], zb[a]);
#pragma unroll
      for (int b = 0; b < MAXD; ++b)
        if (b < dim) L[a][b] = fadd(L[a][b], fmul(fmul(w, za), fsub(yr[b], zb[b])));
    }
  }
#pragma unroll
  for (int c = 0; c < MAXD; ++c)
    if (c < dim) L[c][c] = fadd(L[c][c], 1e-3f);
  int pv[MAXD];
  small_lu<MAXD>(L, pv, dim);
  bool solvable = nz > dim;
#pragma unroll
  for (int c = 0; c < MAXD; ++c)
    if (c < dim && L[c][c] == 0.f) solvable = false;
  float qd[MAXD];
#pragma unroll
  for (int c = 0; c < MAXD; ++c) qd[c] = c < dim ? fsub(q[i * dim + c], zb[c]) : 0.f;

  bool bad = false;
  for (int64_t f0 = 0; f0 < d; f0 += 64) {
    const int64_t f = f0 + lane;
    const bool act = f < d;
    float xb = 0.f;
    if (act)
      for (int j = 0; j < k; ++j) {
        const float w = wr[j];
        if (live(w)) xb = fadd(xb, fmul(w, x_train[static_cast<int64_t>(ir[j]) * d + f]));
      }
    float pred = xb;
    if (solvable) {
      float rh[MAXD];
#pragma unroll
      for (int c = 0; c < MAXD; ++c) rh[c] = 0.f;
      if (act)
        for (int j = 0; j < k; ++j) {
          const float w = wr[j];
          if (!live(w)) continue;
          const int64_t r = ir[j];
          const float xd = fsub(x_train[r * d + f], xb);
          const float* yr = y_train + r * dim;
#pragma unroll
          for (int a = 0; a < MAXD; ++a)
            if (a < dim) rh[a] = fadd(rh[a], fmul(fmul(w, fsub(yr[a], zb[a])), xd));
        }
      small_lu_solve<MAXD>(L, pv, dim, rh);
      float v = xb;
#pragma unroll
      for (int a = 0; a < MAXD; ++a)
        if (a < dim) v = fadd(v, fmul(qd[a], rh[a]));
      pred = v;
      bad = bad || (act && !isfinite(v));
    }
    if (act) o[f] = pred;
  }
  // a non-finite value anywhere in the row falls back to the weighted mean
  if (__ballot(bad) != 0ULL)
    for (int64_t f = lane; f < d; f += 64) {
      float xb = 0.f;
      for (int j = 0; j < k; ++j) {
        const float w = wr[j];
        if (live(w)) xb = fadd(xb, fmul(w, x_train[static_cast<int64_t>(ir[j]) * d + f]));
      }
      o[f] = xb;
    }
}

// G lanes per query (64 / G queries per wave), PER features per lane
// (d <= G * PER): small groups keep the per-step reductions short and let one
// wave serve several queries.  The rows of up to kNegBatch negatives are
// gathered before the first of them is applied, so their loads overlap; the
// moves are then applied one after another as the reference does.  Each
// query's edge state (weight, tail, eps and both schedule words) sits in LDS.
constexpr int kNegBatch = 4;
constexpr int kInvMaxK = 64;  // the kNN width bound (n_neighbors <= 64)

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

template <int PER, int G>
__global__ __launch_bounds__(64 * kInvWaves) void inverse_sgd_kernel(
    float* __restrict__ out, int64_t m, const float* __restrict__ x_train, int64_t d,
    uint32_t n, const int32_t* __restrict__ idx, const float* __restrict__ raw,
    const int32_t* __restrict__ fixed, int k, const uint32_t* __restrict__ wmax_bits,
    const float* __restrict__ sigmas, const float* __restrict__ rhos, int n_epochs, float alpha0,
    float nsr, float gamma, uint64_t seed) {
  constexpr int kQ = 64 / G;  // queries per wave
  __shared__ float s_w[kInvWaves * kQ][kInvMaxK], s_eps[kInvWaves * kQ][kInvMaxK];
  __shared__ float s_eons[kInvWaves * kQ][kInvMaxK], s_eonns[kInvWaves * kQ][kInvMaxK];
  __shared__ int32_t s_t[kInvWaves * kQ][kInvMaxK];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int grp = lane / G, sl = lane % G;
  const int slot = wv * kQ + grp;
  const int64_t i = (static_cast<int64_t>(blockIdx.x) * kInvWaves + wv) * kQ + grp;
  if (i >= m || fixed[i]) return;  // the whole group; no block barriers follow
  const uint64_t gmask = (G == 64 ? ~uint64_t{0} : ((uint64_t{1} << G) - 1)) << (grp * G);
  float* o = out + i * d;
  float cur[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int64_t f = sl + G * t;
    cur[t] = f < d ? o[f] : 0.f;
  }
  // make_epochs_per_sample (lib.rs:4937-4960) over all inverse edges
  const float wmax = fmaxf(__uint_as_float(*wmax_bits), 1e-12f);
  const float ne = static_cast<float>(n_epochs);
  for (int j = sl; j < k; j += G) {
    const float w = raw[i * k + j];
    const float ns = fmul(ne, fdiv(w, wmax));
    const float e = ns > 0.f ? fdiv(ne, ns) : INFINITY;
    s_w[slot][j] = w;
    s_t[slot][j] = idx[i * k + j];
    s_eps[slot][j] = e;
    s_eons[slot][j] = e;
    s_eonns[slot][j] = fdiv(e, nsr);
  }
  // the edge state is written and read by lanes of one wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint64_t key_head = inv_mix64(seed ^ (static_cast<uint64_t>(i) * 0x9E3779B97F4A7C15ULL));
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const float alpha = fmul(alpha0, fsub(1.f, fdiv(static_cast<float>(epoch), ne)));
    const float ef = static_cast<float>(epoch);
    for (int j = 0; j < k; ++j) {
      const float eo = s_eons[slot][j];
      if (eo > ef) continue;
      bool fin = true;
#pragma unroll
      for (int t = 0; t < PER; ++t) fin = fin && isfinite(cur[t]);
      if ((__ballot(!fin) & gmask) != 0ULL) continue;
      const int32_t tj = s_t[slot][j];
      const float wj = s_w[slot][j];
      const float* other = x_train + static_cast<int64_t>(tj) * d;
      float ov[PER];
      float part = 0.f;
#pragma unroll
      for (int t = 0; t < PER; ++t) {
        const int64_t f = sl + G * t;
        ov[t] = f < d ? other[f] : 0.f;
        const float df = cur[t] - ov[t];
        part += df * df;
      }
      const float gd = 1e-6f + sqrtf(group_sum<G>(part));
      const float gc = -(1.f / (wj * sigmas[tj] + 1e-6f));
#pragma unroll
      for (int t = 0; t < PER; ++t)
        if (sl + G * t < d) cur[t] += inv_clip4(gc * ((cur[t] - ov[t]) / gd)) * alpha;
      const float ej = s_eps[slot][j];
      if (sl == 0) s_eons[slot][j] = fadd(eo, ej);
      const float en = fdiv(ej, nsr);
      if (!isfinite(en) || en <= 0.f) continue;
      const float eonn = s_eonns[slot][j];
      const float fl = floorf(fdiv(fsub(ef, eonn), en));
      const int nneg = fl > 0.f ? static_cast<int>(fl) : 0;
      if (sl == 0) s_eonns[slot][j] = fadd(eonn, fmul(static_cast<float>(nneg), en));
      const uint64_t key = inv_mix64(key_head + static_cast<uint64_t>(epoch) * 0xD1B54A32D192ED03ULL +
                                     static_cast<uint64_t>(j));
      for (int s0 = 0; s0 < nneg; s0 += kNegBatch) {
        float nv[kNegBatch][PER];
        float sk[kNegBatch], rk[kNegBatch];
#pragma unroll
        for (int u = 0; u < kNegBatch; ++u) {
          sk[u] = 1.f;
          rk[u] = 0.f;
#pragma unroll
          for (int t = 0; t < PER; ++t) nv[u][t] = 0.f;
          if (s0 + u < nneg) {
            const uint32_t kk = inv_draw(key, static_cast<uint32_t>(s0 + u), n);
            const float* o2 = x_train + static_cast<int64_t>(kk) * d;
#pragma unroll
            for (int t = 0; t < PER; ++t) {
              const int64_t f = sl + G * t;
              if (f < d) nv[u][t] = o2[f];
            }
            sk[u] = sigmas[kk];
            rk[u] = rhos[kk];
          }
        }
#pragma unroll
        for (int u = 0; u < kNegBatch; ++u) {
          if (s0 + u >= nneg) break;
          float p2 = 0.f;
#pragma unroll
          for (int t = 0; t < PER; ++t) {
            const float df = cur[t] - nv[u][t];
            p2 += df * df;
          }
          const float dn = sqrtf(group_sum<G>(p2));
          const float gd2 = 1e-6f + dn;
          const float wh = expf(-fmaxf(dn - rk[u], 1e-6f) / (sk[u] + 1e-6f));
          const float gc2 = -gamma * ((0.f - wh) / ((1.f - wh) * sk[u] + 1e-6f));
#pragma unroll
          for (int t = 0; t < PER; ++t)
            if (sl + G * t < d) cur[t] += inv_clip4(gc2 * ((cur[t] - nv[u][t]) / gd2)) * alpha;
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int64_t f = sl + G * t;
    if (f < d) o[f] = cur[t];
  }
}

}  // namespace

void inverse_transform_device(const float* q, int64_t m, int dim, const float* y_train,
                              const float* x_train, int64_t n, int64_t d, const float* sigmas,
                              const float* rhos, const InverseArgs& args, float* out,
                              KnnWorkspace& kws, hipStream_t s) {
  maybe_inject_fault("inverse_transform");
  if (m == 0) return;
  if (dim < 1 || dim > kMaxLowDim) throw_invalid("n_components must be <= 8 on the GPU path");
  if (args.n_epochs > 0 && d > kInverseMaxFeatures)
    throw_invalid("inverse_transform on the GPU supports at most 2560 features");
  if (n >= (int64_t{1} << 31)) throw_invalid("inverse_transform supports fewer than 2^31 training rows");
  const int k = args.k;
  if (k < 1 || k > kInvMaxK || k > n) throw_invalid("inverse_transform: invalid neighbour count");
  DevBuf<int32_t> idx(static_cast<size_t>(m) * k);
  DevBuf<float> dist(static_cast<size_t>(m) * k);
  knn_device(q, m, y_train, n, dim, k, 0, 0, 2, idx.get(), dist.get(), kws, s, nullptr);
  DevBuf<float> raw(static_cast<size_t>(m) * k), norm(static_cast<size_t>(m) * k);
  DevBuf<int32_t> fixed(static_cast<size_t>(m));
  DevBuf<uint32_t> wmax(1);
  HIP_CHECK(hipMemsetAsync(wmax.get(), 0, sizeof(uint32_t), s));
  hipLaunchKernelGGL(inverse_weights_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, s, dist.get(),
                     m, k, args.a, 2.f * args.b, raw.get(), norm.get(), fixed.get(), wmax.get());
  KERNEL_CHECK();
  const dim3 grid(blocks_for(m, kInvWaves));
  if (dim == 2)
    hipLaunchKernelGGL(inverse_init_kernel<2>, grid, dim3(64 * kInvWaves), 0, s, q, m, dim,
                       idx.get(), norm.get(), fixed.get(), k, y_train, x_train, d, out);
  else
    hipLaunchKernelGGL(inverse_init_kernel<0>, grid, dim3(64 * kInvWaves), 0, s, q, m, dim,
                       idx.get(), norm.get(), fixed.get(), k, y_train, x_train, d, out);
  KERNEL_CHECK();
  if (args.n_epochs > 0) {
    auto sgd = [&](auto kern, int group) {
      const dim3 sgrid(blocks_for(m, kInvWaves * (64 / group)));
      hipLaunchKernelGGL(kern, sgrid, dim3(64 * kInvWaves), 0, s, out, m, x_train, d,
                         static_cast<uint32_t>(n), idx.get(), raw.get(), fixed.get(), k,
                         wmax.get(), sigmas, rhos, args.n_epochs, args.learning_rate,
                         static_cast<float>(args.negative_sample_rate), args.repulsion_strength,
                         args.seed);
      KERNEL_CHECK();
    };
    if (d <= 64)
      sgd(inverse_sgd_kernel<4, 16>, 16);
    else if (d <= 128)
      sgd(inverse_sgd_kernel<8, 16>, 16);
    else if (d <= 256)
      sgd(inverse_sgd_kernel<8, 32>, 32);
    else if (d <= 512)
      sgd(inverse_sgd_kernel<8, 64>, 64);
    else if (d <= 1024)
      sgd(inverse_sgd_kernel<16, 64>, 64);
    else
      sgd(inverse_sgd_kernel<40, 64>, 64);
  }
  // the temporaries above are released in stream order after the launches
}

}  // namespace umapx
