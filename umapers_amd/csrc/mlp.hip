// Learned transform on gfx950: MLP encoder training and inference.
//
// Network: x (standardised, BF16, features padded to dp) -> H1 = relu(x W1 +
// b1) -> H2 = relu(H1 W2 + b2) -> z = H2 W3 + b3.  The two wide layers run on
// one hand-written BF16 MFMA GEMM (32x32x16, FP32 accumulation, 64x64 block
// tiles staged through LDS) with fused epilogues: bias + ReLU in the forward
// pass, the ReLU mask in the backward pass, and split-K atomic accumulation
// for the weight gradients.  The narrow output layer and the loss gradient
// are VALU kernels.  FP32 master weights are updated by Adam; the GEMMs read
// their BF16 copies.
//
// Loss per batch of B edges (i, j) drawn with probability proportional to
// their fuzzy-graph weight, each with n_neg uniformly drawn negatives k:
//   L = (1/B) sum_e [ log(1 + a s_ij^b) + sum_k log(1 + 1 / (a (s_ik + eps)^b)) ]
//       + lambda / S * sum_rows |z - y|^2
// with s = |z_u - z_v|^2, eps = 1e-3, S = B (2 + n_neg) rows and y the fitted
// embedding of each row.  The first term is UMAP's cross-entropy with
// q = 1 / (1 + a s^b) (attraction -log q, repulsion -log(1 - q)); the second
// keeps the encoder in the coordinates of the plain fit, as the reference's
// parametric model regresses on its teacher (parametric.rs:391-393).
#include "mlp.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace umapx {

MlpShape MlpShape::make(int d, int h, int dim) {
  MlpShape s;
  s.d = d;
  s.dp = (d + 31) / 32 * 32;
  s.h = h;
  s.dim = dim;
  int64_t o = 0;
  s.off_w1 = o;
  o += static_cast<int64_t>(s.dp) * h;
  s.off_b1 = o;
  o += h;
  s.off_w2 = o;
  o += static_cast<int64_t>(h) * h;
  s.off_b2 = o;
  o += h;
  s.off_w3 = o;
  o += static_cast<int64_t>(h) * dim;
  s.off_b3 = o;
  o += dim;
  s.total = o;
  return s;
}

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxDim = 8;
constexpr float kRepulsionEps = 1e-3f;

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}
__device__ __forceinline__ uint16_t lane16(const u32x4& v, int i) {
  return static_cast<uint16_t>(v[i >> 1] >> (16 * (i & 1)));
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
uint64_t mix64_host(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t key) {
  return static_cast<double>(mix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

constexpr int kBM = 64, kBN = 64, kBK = 32, kLdsK = kBK + 8;
enum { kEpiBiasRelu = 0, kEpiMask = 1, kEpiAtomic = 2 };

// C = A B, A [M x K] and B [K x N] in BF16, FP32 accumulation on MFMA
// 32x32x16.  A_T: A is stored [K x M] (row stride lda); B_T: B is stored
// [N x K].  Each of the 4 waves owns a 32x32 quarter of the 64x64 block tile;
// K advances 32 at a time through LDS, where both operands sit with K
// contiguous so each lane's MFMA fragment is one 16-byte read.  blockIdx.z
// selects the K range [z * k_chunk, (z + 1) * k_chunk).  Requires M, N, K
// and k_chunk to be multiples of 8 (every 16-byte vector lies inside them).
template <bool A_T, bool B_T, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(
    const uint16_t* __restrict__ A, int lda, const uint16_t* __restrict__ B, int ldb, int M,
    int N, int K, int k_chunk, const float* __restrict__ bias, const uint16_t* __restrict__ mask,
    uint16_t* __restrict__ Cb, float* __restrict__ Cf, int ldc) {
  __shared__ __align__(16) uint16_t As[kBM][kLdsK];
  __shared__ __align__(16) uint16_t Bs[kBN][kLdsK];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int kb = blockIdx.z * k_chunk;
  const int ke = min(K, kb + k_chunk);
  floatx16 acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.f;
  const u32x4 zero = {0u, 0u, 0u, 0u};
  for (int k0 = kb; k0 < ke; k0 += kBK) {
    if (!A_T) {
      const int r = tid >> 2, kv = (tid & 3) * 8;
      const int gm = m0 + r, gk = k0 + kv;
      u32x4 v = zero;
      if (gm < M && gk < ke) v = *reinterpret_cast<const u32x4*>(A + static_cast<int64_t>(gm) * lda + gk);
      *reinterpret_cast<u32x4*>(&As[r][kv]) = v;
    } else {
      const int kk = tid >> 3, mv = (tid & 7) * 8;
      const int gk = k0 + kk, gm = m0 + mv;
      u32x4 v = zero;
      if (gk < ke && gm < M) v = *reinterpret_cast<const u32x4*>(A + static_cast<int64_t>(gk) * lda + gm);
#pragma unroll
      for (int i = 0; i < 8; ++i) As[mv + i][kk] = lane16(v, i);
    }
    if (!B_T) {
      const int kk = tid >> 3, nv = (tid & 7) * 8;
      const int gk = k0 + kk, gn = n0 + nv;
      u32x4 v = zero;
      if (gk < ke && gn < N) v = *reinterpret_cast<const u32x4*>(B + static_cast<int64_t>(gk) * ldb + gn);
#pragma unroll
      for (int i = 0; i < 8; ++i) Bs[nv + i][kk] = lane16(v, i);
    } else {
      const int r = tid >> 2, kv = (tid & 3) * 8;
      const int gn = n0 + r, gk = k0 + kv;
      u32x4 v = zero;
      if (gn < N && gk < ke) v = *reinterpret_cast<const u32x4*>(B + static_cast<int64_t>(gn) * ldb + gk);
      *reinterpret_cast<u32x4*>(&Bs[r][kv]) = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      const u32x4 af = *reinterpret_cast<const u32x4*>(&As[wm * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
      const u32x4 bf = *reinterpret_cast<const u32x4*>(&Bs[wn * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, af),
                                                    __builtin_bit_cast(bf16x8, bf), acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = m0 + wm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
    const int col = n0 + wn * 32 + (lane & 31);
    if (row >= M || col >= N) continue;
    const int64_t o = static_cast<int64_t>(row) * ldc + col;
    if (EPI == kEpiBiasRelu) {
      Cb[o] = to_bf16(fmaxf(acc[i] + bias[col], 0.f));
    } else if (EPI == kEpiMask) {
      const uint16_t mv = mask[o];
      Cb[o] = ((mv & 0x8000u) == 0 && (mv & 0x7FFFu) != 0) ? to_bf16(acc[i]) : uint16_t{0};
    } else {
      atomicAdd(&Cf[o], acc[i]);
    }
  }
}

template <bool A_T, bool B_T, int EPI>
void gemm(const uint16_t* A, int lda, const uint16_t* B, int ldb, int M, int N, int K, int splits,
          const float* bias, const uint16_t* mask, uint16_t* Cb, float* Cf, int ldc, hipStream_t s) {
  const int k_chunk = ((K + splits - 1) / splits + kBK - 1) / kBK * kBK;
  const dim3 grid((N + kBN - 1) / kBN, (M + kBM - 1) / kBM, (K + k_chunk - 1) / k_chunk);
  hipLaunchKernelGGL((gemm_bf16_kernel<A_T, B_T, EPI>), grid, dim3(256), 0, s, A, lda, B, ldb, M,
                     N, K, k_chunk, bias, mask, Cb, Cf, ldc);
  KERNEL_CHECK();
}

// Standardised, BF16, zero-padded copy of the listed rows (row index < 0, or
// feature >= d, gives 0); the fitted embedding of each row alongside.
__global__ void gather_rows_kernel(const float* __restrict__ x, int d, int dp,
                                   const int32_t* __restrict__ rows, int64_t first_row,
                                   int64_t count, const float* __restrict__ mean,
                                   const float* __restrict__ inv_std, uint16_t* __restrict__ xb,
                                   const float* __restrict__ y, int dim, float* __restrict__ yt) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count * dp) return;
  const int64_t s = t / dp;
  const int f = static_cast<int>(t % dp);
  const int64_t r = rows ? static_cast<int64_t>(rows[s]) : first_row + s;
  float v = 0.f;
  if (r >= 0 && f < d) v = (x[r * d + f] - mean[f]) * inv_std[f];
  xb[t] = to_bf16(v);
  if (yt && f < dim) yt[s * dim + f] = (r >= 0) ? y[r * dim + f] : 0.f;
}

// z = H2 W3 + b3, one wave per row (W3 from its BF16 copy).
__global__ __launch_bounds__(256) void out_layer_kernel(const uint16_t* __restrict__ h2,
                                                        int64_t rows, int h, int dim,
                                                        const uint16_t* __restrict__ w3h,
                                                        const float* __restrict__ b3,
                                                        float* __restrict__ z) {
  const int lane = threadIdx.x & 63;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;
  float acc[kMaxDim];
#pragma unroll
  for (int c = 0; c < kMaxDim; ++c) acc[c] = 0.f;
  for (int j = lane; j < h; j += 64) {
    const float hv = from_bf16(h2[r * h + j]);
#pragma unroll
    for (int c = 0; c < kMaxDim; ++c)
      if (c < dim) acc[c] += hv * from_bf16(w3h[j * dim + c]);
  }
#pragma unroll
  for (int c = 0; c < kMaxDim; ++c)
    for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o);
  if (lane < dim) {
    float v = 0.f;
#pragma unroll
    for (int c = 0; c < kMaxDim; ++c)
      if (c == lane) v = acc[c];
    z[r * dim + lane] = v + b3[lane];
  }
}

// dL/dz for every row of the batch and the batch loss (one thread per edge;
// every row belongs to exactly one edge, so no atomics on dz).
__global__ void loss_grad_kernel(const float* __restrict__ z, const float* __restrict__ yt,
                                 int n_edges, int n_neg, int dim, int s_real, float a, float b,
                                 float lambda, float* __restrict__ dz, double* __restrict__ loss) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (e < n_edges) {
    const float inv_b = 1.f / static_cast<float>(n_edges);
    const float mse_c = 2.f * lambda / static_cast<float>(s_real);
    const int64_t hi = e, ti = static_cast<int64_t>(n_edges) + e;
    float zi[kMaxDim], gi[kMaxDim];
    for (int c = 0; c < dim; ++c) {
      zi[c] = z[hi * dim + c];
      gi[c] = 0.f;
    }
    auto mse = [&](int64_t row, const float* zr, float* g) {
      float sq = 0.f;
      for (int c = 0; c < dim; ++c) {
        const float df = zr[c] - yt[row * dim + c];
        g[c] += mse_c * df;
        sq += df * df;
      }
      return static_cast<double>(lambda) * sq / s_real;
    };
    // attraction
    {
      float zj[kMaxDim], gj[kMaxDim];
      float s = 0.f;
      for (int c = 0; c < dim; ++c) {
        zj[c] = z[ti * dim + c];
        gj[c] = 0.f;
        const float df = zi[c] - zj[c];
        s += df * df;
      }
      if (s > 0.f) {
        const float pb = powf(s, b);
        const float P = a * pb;
        l += log1p(static_cast<double>(P)) * inv_b;
        const float dlds = a * b * (pb / s) / (1.f + P);
        for (int c = 0; c < dim; ++c) {
          const float g = 2.f * dlds * (zi[c] - zj[c]) * inv_b;
          gi[c] += g;
          gj[c] -= g;
        }
      }
      l += mse(ti, zj, gj);
      for (int c = 0; c < dim; ++c) dz[ti * dim + c] = gj[c];
    }
    // repulsion
    for (int t = 0; t < n_neg; ++t) {
      const int64_t ki = 2 * static_cast<int64_t>(n_edges) + static_cast<int64_t>(e) * n_neg + t;
      float zk[kMaxDim], gk[kMaxDim];
      float s = 0.f;
      for (int c = 0; c < dim; ++c) {
        zk[c] = z[ki * dim + c];
        gk[c] = 0.f;
        const float df = zi[c] - zk[c];
        s += df * df;
      }
      const float se = s + kRepulsionEps;
      const float P = a * powf(se, b);
      l += log1p(1.0 / static_cast<double>(P)) * inv_b;
      const float dlds = -b / (se * (1.f + P));
      for (int c = 0; c < dim; ++c) {
        const float g = 2.f * dlds * (zi[c] - zk[c]) * inv_b;
        gi[c] += g;
        gk[c] -= g;
      }
      l += mse(ki, zk, gk);
      for (int c = 0; c < dim; ++c) dz[ki * dim + c] = gk[c];
    }
    l += mse(hi, zi, gi);
    for (int c = 0; c < dim; ++c) dz[hi * dim + c] = gi[c];
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if (loss && (threadIdx.x & 63) == 0 && l != 0.0) atomicAdd(loss, l);
}

// dH2 = (dz W3^T) masked by H2 > 0, in BF16.
__global__ void out_layer_back_kernel(const float* __restrict__ dz, const uint16_t* __restrict__ h2,
                                      int64_t rows, int h, int dim,
                                      const uint16_t* __restrict__ w3h, uint16_t* __restrict__ dh2) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= rows * h) return;
  const int64_t r = t / h;
  const int j = static_cast<int>(t % h);
  float v = 0.f;
  if (from_bf16(h2[t]) > 0.f)
    for (int c = 0; c < dim; ++c) v += dz[r * dim + c] * from_bf16(w3h[j * dim + c]);
  dh2[t] = to_bf16(v);
}

// Column sums over a block of rows, added atomically: out[j] += sum_r m[r][j]
// (BF16 input), and, for the output layer, w[j][c] += sum_r h2[r][j] dz[r][c]
// and bias[c] += sum_r dz[r][c].
// Column reductions of the backward pass in two stages: block b sums rows
// [b * kRedRows, (b + 1) * kRedRows) of every column into part[b][q] (one
// thread per column, no atomics), then red_final_kernel adds the blocks'
// partials per output in block order.  Outputs q: with WGRAD, q = j * dim + c
// is sum_r h[r][j] dz[r][c] (the output layer's weight gradient) and
// q = h * dim + c is sum_r dz[r][c] (its bias gradient); without, q = j is
// sum_r m[r][j] (a hidden layer's bias gradient).
constexpr int kRedRows = 128;

template <bool WGRAD>
__global__ __launch_bounds__(256) void red_partial_kernel(const uint16_t* __restrict__ m,
                                                          const float* __restrict__ dz,
                                                          int64_t rows, int cols, int dim,
                                                          int q_len, float* __restrict__ part) {
  const int j = blockIdx.y * 256 + threadIdx.x;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kRedRows;
  const int64_t r1 = min(rows, r0 + kRedRows);
  float* out = part + static_cast<int64_t>(blockIdx.x) * q_len;
  if (WGRAD) {
    float acc[kMaxDim];
#pragma unroll
    for (int c = 0; c < kMaxDim; ++c) acc[c] = 0.f;
    if (j < cols) {
#pragma unroll 8
      for (int64_t r = r0; r < r1; ++r) {
        const float hv = from_bf16(m[r * cols + j]);
#pragma unroll
        for (int c = 0; c < kMaxDim; ++c)
          if (c < dim) acc[c] += hv * dz[r * dim + c];
      }
#pragma unroll
      for (int c = 0; c < kMaxDim; ++c)
        if (c < dim) out[j * dim + c] = acc[c];
    }
    if (blockIdx.y == 0 && threadIdx.x < dim) {
      float sb = 0.f;
      for (int64_t r = r0; r < r1; ++r) sb += dz[r * dim + threadIdx.x];
      out[cols * dim + threadIdx.x] = sb;
    }
  } else if (j < cols) {
    float acc = 0.f;
#pragma unroll 8
    for (int64_t r = r0; r < r1; ++r) acc += from_bf16(m[r * cols + j]);
    out[j] = acc;
  }
}

// eight independent running sums keep eight loads in flight; they are
// combined in a fixed order
__global__ void red_final_kernel(const float* __restrict__ part, int nb, int q_len,
                                 float* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= q_len) return;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int b = 0;
  for (; b + 8 <= nb; b += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += part[static_cast<int64_t>(b + u) * q_len + q];
  }
  for (; b < nb; ++b) acc[0] += part[static_cast<int64_t>(b) * q_len + q];
  out[q] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

// Adam on the FP32 masters (bias-corrected step), then the BF16 copies.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr_t,
                            float b1, float b2, float eps, uint16_t* __restrict__ ph) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr_t * mi / (sqrtf(vi) + eps);
  p[i] = pi;
  ph[i] = to_bf16(pi);
}

__global__ void to_bf16_kernel(const float* __restrict__ p, int64_t n, uint16_t* __restrict__ ph) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) ph[i] = to_bf16(p[i]);
}

// He-uniform weights U(-sqrt(6 / fan_in), +) from a counter hash; zero
// biases and zero rows of W1 past the real input width.
__global__ void init_params_kernel(float* __restrict__ p, int64_t n, MlpShape sh, uint64_t seed) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = 0.f;
  float fan = 0.f;
  if (i >= sh.off_w1 && i < sh.off_b1) {
    const int64_t row = (i - sh.off_w1) / sh.h;
    if (row < sh.d) fan = static_cast<float>(sh.d);
  } else if (i >= sh.off_w2 && i < sh.off_b2) {
    fan = static_cast<float>(sh.h);
  } else if (i >= sh.off_w3 && i < sh.off_b3) {
    fan = static_cast<float>(sh.h);
  }
  if (fan > 0.f) {
    const float lim = sqrtf(6.f / fan);
    v = static_cast<float>(2.0 * uniform01(seed ^ mix64(static_cast<uint64_t>(i) + 1)) - 1.0) * lim;
  }
  p[i] = v;
}

// Batch rows: edge e draws u ~ U(0, total) and takes the first edge whose
// inclusive weight prefix exceeds u; its head and tail fill slots e and
// B + e, and its negatives (uniform over the n rows) slots 2B + e n_neg + t.
__global__ void sample_batch_kernel(const double* __restrict__ cdf, int64_t ne,
                                    const int32_t* __restrict__ head,
                                    const int32_t* __restrict__ tail, int n_edges, int n_neg,
                                    int64_t n, uint64_t key, int32_t* __restrict__ rows) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_edges) return;
  const double total = cdf[ne - 1];
  const double u = uniform01(key ^ mix64(2 * static_cast<uint64_t>(e) + 1)) * total;
  int64_t lo = 0, hi = ne - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (cdf[mid] > u)
      hi = mid;
    else
      lo = mid + 1;
  }
  rows[e] = head[lo];
  rows[n_edges + e] = tail[lo];
  for (int t = 0; t < n_neg; ++t) {
    const uint64_t h = mix64(key ^ mix64(2 * (static_cast<uint64_t>(e) * n_neg + t) + 2));
    rows[2 * static_cast<int64_t>(n_edges) + static_cast<int64_t>(e) * n_neg + t] =
        static_cast<int32_t>(((h >> 32) * static_cast<uint64_t>(n)) >> 32);
  }
}

__global__ void fill_i32_kernel(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Column mean and 1/std of the training rows (f64 sums); a zero-variance
// column gets 1/std = 1.
__global__ void col_moments_kernel(const float* __restrict__ x, int64_t n, int d,
                                   double* __restrict__ sums) {
  const int f = blockIdx.y * 256 + threadIdx.x;
  if (f >= d) return;
  double s = 0.0, s2 = 0.0;
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const double v = x[r * d + f];
    s += v;
    s2 += v * v;
  }
  atomicAdd(&sums[f], s);
  atomicAdd(&sums[d + f], s2);
}
__global__ void col_finish_kernel(const double* __restrict__ sums, int64_t n, int d,
                                  float* __restrict__ mean, float* __restrict__ inv_std) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d) return;
  const double mu = sums[f] / static_cast<double>(n);
  const double var = std::max(sums[d + f] / static_cast<double>(n) - mu * mu, 0.0);
  mean[f] = static_cast<float>(mu);
  inv_std[f] = var > 1e-24 ? static_cast<float>(1.0 / std::sqrt(var)) : 1.f;
}

// Device buffers of one forward / backward pass over s_pad rows.
struct Pass {
  DevBuf<uint16_t> xb, h1, h2, dh2, dh1;
  DevBuf<float> yt, z, dz, part;
  DevBuf<double> loss;
  void ensure(const MlpShape& sh, int64_t s_pad) {
    xb.ensure(s_pad * sh.dp);
    h1.ensure(s_pad * sh.h);
    h2.ensure(s_pad * sh.h);
    dh2.ensure(s_pad * sh.h);
    dh1.ensure(s_pad * sh.h);
    yt.ensure(s_pad * sh.dim);
    z.ensure(s_pad * sh.dim);
    dz.ensure(s_pad * sh.dim);
    loss.ensure(1);
  }
};

void forward(const MlpShape& sh, const float* params, const uint16_t* wh, Pass& ps, int64_t rows,
             hipStream_t s) {
  const int r = static_cast<int>(rows);
  gemm<false, false, kEpiBiasRelu>(ps.xb.get(), sh.dp, wh + sh.off_w1, sh.h, r, sh.h, sh.dp, 1,
                                   params + sh.off_b1, nullptr, ps.h1.get(), nullptr, sh.h, s);
  gemm<false, false, kEpiBiasRelu>(ps.h1.get(), sh.h, wh + sh.off_w2, sh.h, r, sh.h, sh.h, 1,
                                   params + sh.off_b2, nullptr, ps.h2.get(), nullptr, sh.h, s);
  hipLaunchKernelGGL(out_layer_kernel, dim3(static_cast<unsigned>((rows + 3) / 4)), dim3(256), 0, s,
                     ps.h2.get(), rows, sh.h, sh.dim, wh + sh.off_w3, params + sh.off_b3, ps.z.get());
  KERNEL_CHECK();
}

// Column reduction into out (overwritten): see red_partial_kernel.
template <bool WGRAD>
void reduce_cols(const uint16_t* m, const float* dz, int64_t rows, int cols, int dim, Pass& ps,
                 float* out, hipStream_t s) {
  const int q_len = WGRAD ? (cols + 1) * dim : cols;
  const int nb = static_cast<int>((rows + kRedRows - 1) / kRedRows);
  ps.part.ensure(static_cast<size_t>(nb) * q_len);
  hipLaunchKernelGGL(red_partial_kernel<WGRAD>, dim3(nb, (cols + 255) / 256), dim3(256), 0, s, m, dz,
                     rows, cols, dim, q_len, ps.part.get());
  hipLaunchKernelGGL(red_final_kernel, dim3((q_len + 255) / 256), dim3(256), 0, s, ps.part.get(), nb,
                     q_len, out);
  KERNEL_CHECK();
}

// Gradients of the loss into grads (zeroed here), given dz.
void backward(const MlpShape& sh, const uint16_t* wh, Pass& ps, int64_t rows, float* grads,
              hipStream_t s) {
  HIP_CHECK(hipMemsetAsync(grads, 0, sizeof(float) * sh.total, s));
  const int r = static_cast<int>(rows);
  // dW3 and db3 are adjacent in the flat layout: one reduction writes both
  reduce_cols<true>(ps.h2.get(), ps.dz.get(), rows, sh.h, sh.dim, ps, grads + sh.off_w3, s);
  hipLaunchKernelGGL(out_layer_back_kernel, dim3(blocks_for(rows * sh.h, 256)), dim3(256), 0, s,
                     ps.dz.get(), ps.h2.get(), rows, sh.h, sh.dim, wh + sh.off_w3, ps.dh2.get());
  reduce_cols<false>(ps.dh2.get(), nullptr, rows, sh.h, 0, ps, grads + sh.off_b2, s);
  KERNEL_CHECK();
  const int splits = std::max(1, r / 1024);
  // dW2 = H1^T dH2
  gemm<true, false, kEpiAtomic>(ps.h1.get(), sh.h, ps.dh2.get(), sh.h, sh.h, sh.h, r, splits,
                                nullptr, nullptr, nullptr, grads + sh.off_w2, sh.h, s);
  // dH1 = (dH2 W2^T) masked by H1 > 0
  gemm<false, true, kEpiMask>(ps.dh2.get(), sh.h, wh + sh.off_w2, sh.h, r, sh.h, sh.h, 1, nullptr,
                              ps.h1.get(), ps.dh1.get(), nullptr, sh.h, s);
  reduce_cols<false>(ps.dh1.get(), nullptr, rows, sh.h, 0, ps, grads + sh.off_b1, s);
  KERNEL_CHECK();
  // dW1 = X^T dH1
  gemm<true, false, kEpiAtomic>(ps.xb.get(), sh.dp, ps.dh1.get(), sh.h, sh.dp, sh.h, r, splits,
                                nullptr, nullptr, nullptr, grads + sh.off_w1, sh.h, s);
}

int64_t pad64(int64_t v) { return (v + 63) / 64 * 64; }

void check_shape(int d, int h, int dim) {
  if (dim < 1 || dim > kMaxDim) throw_invalid("the learned transform supports n_components <= 8");
  if (h < 64 || h % 64 != 0) throw_invalid("mlp_hidden must be a positive multiple of 64");
  if (d < 1) throw_invalid("input rows must have at least one feature");
}

}  // namespace

void pack_model(MlpModel& m, hipStream_t s);

void mlp_train(MlpModel& m, const float* x, int64_t n, int64_t d, const float* y, int dim,
               const Graph& g, float a, float b, const umapx_params& p, hipStream_t s) {
  maybe_inject_fault("mlp_train");
  check_shape(static_cast<int>(d), p.mlp_hidden, dim);
  if (p.mlp_steps < 1 || p.mlp_batch < 1 || p.mlp_negatives < 0)
    throw_invalid("mlp_steps and mlp_batch must be >= 1 and mlp_negatives >= 0");
  if (!(p.mlp_learning_rate > 0.f) || !(p.mlp_embedding_weight >= 0.f))
    throw_invalid("mlp_learning_rate must be > 0 and mlp_embedding_weight >= 0");
  if (g.n_edges < 1) throw_invalid("the fitted graph has no edges to train the learned transform on");
  const MlpShape sh = MlpShape::make(static_cast<int>(d), p.mlp_hidden, dim);
  m.shape = sh;
  m.params.ensure(sh.total);
  m.grads.ensure(sh.total);
  m.adam_m.ensure(sh.total);
  m.adam_v.ensure(sh.total);
  m.wh.ensure(sh.total);
  m.mean.ensure(d);
  m.inv_std.ensure(d);
  const uint64_t seed = p.random_seed ^ 0x6D6C7045ACE0F00DULL;
  hipLaunchKernelGGL(init_params_kernel, dim3(blocks_for(sh.total, 256)), dim3(256), 0, s,
                     m.params.get(), sh.total, sh, seed);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(sh.total, 256)), dim3(256), 0, s,
                     m.params.get(), sh.total, m.wh.get());
  HIP_CHECK(hipMemsetAsync(m.adam_m.get(), 0, sizeof(float) * sh.total, s));
  HIP_CHECK(hipMemsetAsync(m.adam_v.get(), 0, sizeof(float) * sh.total, s));
  {
    DevBuf<double> sums(2 * d);
    HIP_CHECK(hipMemsetAsync(sums.get(), 0, sizeof(double) * 2 * d, s));
    hipLaunchKernelGGL(col_moments_kernel,
                       dim3(static_cast<unsigned>(std::min<int64_t>(n, 512)),
                            static_cast<unsigned>((d + 255) / 256)),
                       dim3(256), 0, s, x, n, static_cast<int>(d), sums.get());
    hipLaunchKernelGGL(col_finish_kernel, dim3(blocks_for(d, 256)), dim3(256), 0, s, sums.get(), n,
                       static_cast<int>(d), m.mean.get(), m.inv_std.get());
    KERNEL_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));
  }
  // inclusive prefix sums of the edge weights, in f64 on the host (once)
  const int64_t ne = g.n_edges;
  DevBuf<double> cdf(ne);
  {
    std::vector<float> w(ne);
    HIP_CHECK(hipMemcpyAsync(w.data(), g.weight.get(), 4 * ne, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<double> c(ne);
    double run = 0.0;
    for (int64_t i = 0; i < ne; ++i) {
      run += static_cast<double>(w[i]);
      c[i] = run;
    }
    HIP_CHECK(hipMemcpyAsync(cdf.get(), c.data(), 8 * ne, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
  const int B = p.mlp_batch, nn = p.mlp_negatives;
  const int64_t s_real = static_cast<int64_t>(B) * (2 + nn);
  const int64_t s_pad = pad64(s_real);
  Pass ps;
  ps.ensure(sh, s_pad);
  DevBuf<int32_t> rows(s_pad);
  hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(s_pad, 256)), dim3(256), 0, s, rows.get(),
                     s_pad, -1);
  HIP_CHECK(hipMemsetAsync(ps.dz.get(), 0, sizeof(float) * s_pad * dim, s));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  std::vector<double> loss_log;
  for (int step = 0; step < p.mlp_steps; ++step) {
    const uint64_t key = mix64_host(seed + 0x9E3779B97F4A7C15ULL * (static_cast<uint64_t>(step) + 1));
    hipLaunchKernelGGL(sample_batch_kernel, dim3(blocks_for(B, 256)), dim3(256), 0, s, cdf.get(),
                       ne, g.head.get(), g.tail.get(), B, nn, n, key, rows.get());
    hipLaunchKernelGGL(gather_rows_kernel, dim3(blocks_for(s_pad * sh.dp, 256)), dim3(256), 0, s, x,
                       static_cast<int>(d), sh.dp, rows.get(), int64_t{0}, s_pad, m.mean.get(),
                       m.inv_std.get(), ps.xb.get(), y, dim, ps.yt.get());
    KERNEL_CHECK();
    forward(sh, m.params.get(), m.wh.get(), ps, s_pad, s);
    const bool log_loss = step == 0 || step + 1 == p.mlp_steps;
    if (log_loss) HIP_CHECK(hipMemsetAsync(ps.loss.get(), 0, sizeof(double), s));
    hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks_for(B, 256)), dim3(256), 0, s, ps.z.get(),
                       ps.yt.get(), B, nn, dim, static_cast<int>(s_real), a, b,
                       p.mlp_embedding_weight, ps.dz.get(), log_loss ? ps.loss.get() : nullptr);
    KERNEL_CHECK();
    backward(sh, m.wh.get(), ps, s_pad, m.grads.get(), s);
    const double t1 = static_cast<double>(step) + 1.0;
    const float lr_t = static_cast<float>(p.mlp_learning_rate * std::sqrt(1.0 - std::pow(b2, t1)) /
                                          (1.0 - std::pow(b1, t1)));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(sh.total, 256)), dim3(256), 0, s,
                       m.params.get(), m.grads.get(), m.adam_m.get(), m.adam_v.get(), sh.total,
                       lr_t, b1, b2, eps, m.wh.get());
    KERNEL_CHECK();
    if (log_loss) {
      double l = 0.0;
      HIP_CHECK(hipMemcpyAsync(&l, ps.loss.get(), sizeof(double), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      if (step == 0) m.first_loss = l;
      m.last_loss = l;
    }
  }
  pack_model(m, s);
  HIP_CHECK(hipStreamSynchronize(s));
  m.steps_done = p.mlp_steps;
  m.trained = true;
}

namespace {

// ---------------------------------------------------------------------------
// Fused inference: raw f32 rows in, z out, in one kernel.
//
// Each wave carries 32 rows through all three layers in the transposed form
// H^T = W^T X^T: the weights are the MFMA's A operand (32 output features x
// 16 inputs) and the rows its B operand (16 inputs x 32 rows), so a lane
// ends a layer holding one row's outputs for 16 features of every 32-feature
// tile: features 8q + 4g + r (q, r = 0..3) of the tile, g = lane / 32.  The
// sum inside one MFMA step does not depend on which 8 inputs a lane supplies,
// as long as the A and B operands agree; so the next layer takes that
// register footprint as its B operand unchanged (step 2t + h uses the lane's
// values 8h .. 8h + 7 of tile t) and its weights are packed in the matching
// input order.  H1 and H2 never leave the registers; the only HBM traffic is
// the rows in (4 d bytes) and z out (4 dim bytes).
//
// W2's packed fragments (H^2 BF16 values), the biases, W3 (f32 copies of the
// BF16 values) and the standardisation live in LDS; W1's fragments (dp H
// values) are read through the caches, one 16-byte vector per lane per MFMA.
// Layer 1 takes inputs 16 s + 8 g .. + 7 at step s, so the two lanes of a row
// read the two halves of one 64-byte piece of it (one cache line per row per
// load, where a lane-contiguous order touched two).
// The arithmetic is the unfused chain's: standardise in f32, round to BF16,
// FP32 accumulation, bias + ReLU rounded to BF16 for H1 and H2, an f32 dot
// product with W3 (only the order of the f32 sums differs).
constexpr int kFusedMaxDp = 1024;
constexpr int kFusedRowsPerWave = 32;
constexpr int kFusedWaves = 8;

// Packed weights (g = l / 32, ks1 = dp / 16, ks2 = h / 16, p(s, g, j) =
// 32 (s / 2) + 16 (s % 2) + 8 (j / 4) + 4 g + j % 4, the input order of the
// register footprint):
//   w1p[((o ks1 + s) 64 + l) 8 + j] = W1[16 s + 8 g + j][32 o + l % 32]
//   w2p[((o ks2 + s) 64 + l) 8 + j] = W2[p(s, g, j)][32 o + l % 32]
//   w3p[(s 64 + l) 8 + j]          = W3[p(s, g, j)][l % 32], 0 for l % 32 >= dim
__global__ void mlp_pack_kernel(const uint16_t* __restrict__ wh, MlpShape sh,
                                uint16_t* __restrict__ w1p, uint16_t* __restrict__ w2p,
                                uint16_t* __restrict__ w3p) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n1 = static_cast<int64_t>(sh.dp) * sh.h, n2 = static_cast<int64_t>(sh.h) * sh.h;
  const int64_t n3 = static_cast<int64_t>(sh.h) * 32;
  const int ks1 = sh.dp / 16, ks2 = sh.h / 16;
  if (i < n1) {
    const int j = static_cast<int>(i & 7), l = static_cast<int>((i >> 3) & 63);
    const int64_t os = i >> 9;
    const int o = static_cast<int>(os / ks1), s = static_cast<int>(os % ks1);
    const int in = 16 * s + 8 * (l >> 5) + j;
    w1p[i] = wh[sh.off_w1 + static_cast<int64_t>(in) * sh.h + 32 * o + (l & 31)];
  } else if (i < n1 + n2) {
    const int64_t k = i - n1;
    const int j = static_cast<int>(k & 7), l = static_cast<int>((k >> 3) & 63);
    const int64_t os = k >> 9;
    const int o = static_cast<int>(os / ks2), s = static_cast<int>(os % ks2);
    const int in = 32 * (s >> 1) + 16 * (s & 1) + 8 * (j >> 2) + 4 * (l >> 5) + (j & 3);
    w2p[k] = wh[sh.off_w2 + static_cast<int64_t>(in) * sh.h + 32 * o + (l & 31)];
  } else if (i < n1 + n2 + n3) {
    const int64_t k = i - n1 - n2;
    const int j = static_cast<int>(k & 7), l = static_cast<int>((k >> 3) & 63);
    const int s = static_cast<int>(k >> 9);
    const int in = 32 * (s >> 1) + 16 * (s & 1) + 8 * (j >> 2) + 4 * (l >> 5) + (j & 3);
    const int c = l & 31;
    w3p[k] = c < sh.dim ? wh[sh.off_w3 + static_cast<int64_t>(in) * sh.dim + c] : uint16_t{0};
  }
}

// One workgroup of kFusedWaves waves per CU; wave w of block b takes the
// 32-row tiles b * kFusedWaves + w, then every gridDim.x * kFusedWaves-th.
// Requires d % 4 == 0, x 16-byte aligned, dp <= kFusedMaxDp, dim <= kMaxDim.
template <int H>
__global__ __launch_bounds__(kFusedWaves * 64) void mlp_infer_kernel(
    const float* __restrict__ x, int64_t rows, int d, int dp, int dim,
    const uint16_t* __restrict__ w1p, const uint16_t* __restrict__ w2p,
    const float* __restrict__ b1, const float* __restrict__ b2, const uint16_t* __restrict__ w3p,
    const float* __restrict__ b3, const float* __restrict__ mean,
    const float* __restrict__ inv_std, float* __restrict__ z) {
  constexpr int T = H / 32, KS2 = H / 16;
  __shared__ u32x4 w2s[T * KS2 * 64];
  __shared__ __align__(16) float b1s[H];
  __shared__ __align__(16) float b2s[H];
  __shared__ u32x4 w3s[KS2 * 64];
  __shared__ __align__(16) float mus[kFusedMaxDp];
  __shared__ __align__(16) float iss[kFusedMaxDp];
  for (int i = threadIdx.x; i < T * KS2 * 64; i += blockDim.x)
    w2s[i] = reinterpret_cast<const u32x4*>(w2p)[i];
  for (int i = threadIdx.x; i < H; i += blockDim.x) {
    b1s[i] = b1[i];
    b2s[i] = b2[i];
  }
  for (int i = threadIdx.x; i < KS2 * 64; i += blockDim.x)
    w3s[i] = reinterpret_cast<const u32x4*>(w3p)[i];
  for (int i = threadIdx.x; i < dp; i += blockDim.x) {
    mus[i] = i < d ? mean[i] : 0.f;
    iss[i] = i < d ? inv_std[i] : 0.f;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, g = lane >> 5, r32 = lane & 31;
  const int ks1 = dp >> 4;
  const u32x4* w1v = reinterpret_cast<const u32x4*>(w1p) + lane;
  const int64_t n_tiles = (rows + kFusedRowsPerWave - 1) / kFusedRowsPerWave;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kFusedWaves;
  const float4 f4zero = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kFusedWaves + (threadIdx.x >> 6);
       t < n_tiles; t += stride) {
    const int64_t row = t * kFusedRowsPerWave + r32;
    const bool live = row < rows;
    const float* xr = x + (live ? row : 0) * static_cast<int64_t>(d) + 8 * g;
    const int fmax = live ? d - 8 * g : 0;  // inputs at or past this lane's first that exist
    floatx16 acc[T];
#pragma unroll
    for (int o = 0; o < T; ++o)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[o][i] = 0.f;

    // ---- layer 1: acc = W1^T x^T, x standardised and rounded to BF16 ----
    float4 xa_lo = f4zero, xa_hi = f4zero, xb_lo = f4zero, xb_hi = f4zero;
    if (0 < fmax) xa_lo = *reinterpret_cast<const float4*>(xr);
    if (4 < fmax) xa_hi = *reinterpret_cast<const float4*>(xr + 4);
    if (ks1 > 1) {
      if (16 < fmax) xb_lo = *reinterpret_cast<const float4*>(xr + 16);
      if (20 < fmax) xb_hi = *reinterpret_cast<const float4*>(xr + 20);
    }
    // W1 fragments of the next step, read through the caches (every wave of
    // the chip reads the same dp H values)
    u32x4 wa[T];
#pragma unroll
    for (int o = 0; o < T; ++o) wa[o] = w1v[o * ks1 * 64];
    for (int s = 0; s < ks1; ++s) {
      const int f = 16 * s + 8 * g;
      const float4 m_lo = *reinterpret_cast<const float4*>(&mus[f]);
      const float4 m_hi = *reinterpret_cast<const float4*>(&mus[f + 4]);
      const float4 i_lo = *reinterpret_cast<const float4*>(&iss[f]);
      const float4 i_hi = *reinterpret_cast<const float4*>(&iss[f + 4]);
      u32x4 xv;
      xv[0] = static_cast<uint32_t>(to_bf16((xa_lo.x - m_lo.x) * i_lo.x)) |
              (static_cast<uint32_t>(to_bf16((xa_lo.y - m_lo.y) * i_lo.y)) << 16);
      xv[1] = static_cast<uint32_t>(to_bf16((xa_lo.z - m_lo.z) * i_lo.z)) |
              (static_cast<uint32_t>(to_bf16((xa_lo.w - m_lo.w) * i_lo.w)) << 16);
      xv[2] = static_cast<uint32_t>(to_bf16((xa_hi.x - m_hi.x) * i_hi.x)) |
              (static_cast<uint32_t>(to_bf16((xa_hi.y - m_hi.y) * i_hi.y)) << 16);
      xv[3] = static_cast<uint32_t>(to_bf16((xa_hi.z - m_hi.z) * i_hi.z)) |
              (static_cast<uint32_t>(to_bf16((xa_hi.w - m_hi.w) * i_hi.w)) << 16);
      // inputs two steps ahead (rows are read once, from HBM)
      xa_lo = xb_lo;
      xa_hi = xb_hi;
      xb_lo = f4zero;
      xb_hi = f4zero;
      const int f2 = 16 * (s + 2);
      if (s + 2 < ks1) {
        if (f2 < fmax) xb_lo = *reinterpret_cast<const float4*>(xr + f2);
        if (f2 + 4 < fmax) xb_hi = *reinterpret_cast<const float4*>(xr + f2 + 4);
      }
      const bf16x8 xf = __builtin_bit_cast(bf16x8, xv);
#pragma unroll
      for (int o = 0; o < T; ++o)
        acc[o] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, wa[o]), xf,
                                                         acc[o], 0, 0, 0);
      if (s + 1 < ks1) {
#pragma unroll
        for (int o = 0; o < T; ++o) wa[o] = w1v[(o * ks1 + s + 1) * 64];
      }
    }

    // ---- H1 = BF16(relu(acc + b1)), kept as layer 2's B operand ----
    u32x4 h1[T][2];
#pragma unroll
    for (int o = 0; o < T; ++o) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 bb = *reinterpret_cast<const float4*>(&b1s[32 * o + 8 * q + 4 * g]);
        const uint32_t lo = static_cast<uint32_t>(to_bf16(fmaxf(acc[o][4 * q + 0] + bb.x, 0.f))) |
                            (static_cast<uint32_t>(to_bf16(fmaxf(acc[o][4 * q + 1] + bb.y, 0.f))) << 16);
        const uint32_t hi = static_cast<uint32_t>(to_bf16(fmaxf(acc[o][4 * q + 2] + bb.z, 0.f))) |
                            (static_cast<uint32_t>(to_bf16(fmaxf(acc[o][4 * q + 3] + bb.w, 0.f))) << 16);
        h1[o][q >> 1][2 * (q & 1)] = lo;
        h1[o][q >> 1][2 * (q & 1) + 1] = hi;
      }
    }

    // ---- layer 2 (W2 from LDS), two output tiles at a time, each folded
    // into z as soon as it is done: H2 = BF16(relu(acc + b2)), z = H2 W3 ----
    floatx16 a3;
#pragma unroll
    for (int i = 0; i < 16; ++i) a3[i] = 0.f;
#pragma unroll
    for (int op = 0; op < T; op += 2) {
      floatx16 a2[2];
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int i = 0; i < 16; ++i) a2[u][i] = 0.f;
      // W2 fragments one step ahead; the scheduling barrier keeps the
      // compiler from hoisting every step's LDS reads (register spills)
      u32x4 wf[2] = {w2s[(op * KS2) * 64 + lane], w2s[((op + 1) * KS2) * 64 + lane]};
#pragma unroll
      for (int s2 = 0; s2 < KS2; ++s2) {
        const bf16x8 hb = __builtin_bit_cast(bf16x8, h1[s2 >> 1][s2 & 1]);
        u32x4 wc[2] = {wf[0], wf[1]};
        if (s2 + 1 < KS2) {
          wf[0] = w2s[(op * KS2 + s2 + 1) * 64 + lane];
          wf[1] = w2s[((op + 1) * KS2 + s2 + 1) * 64 + lane];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
          a2[u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, wc[u]), hb,
                                                          a2[u], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
      // H2 of these two tiles in layer 2's output footprint is again a B
      // operand: the out layer z^T = W3^T H2^T runs as one more MFMA per
      // 16 features, with W3^T's rows past dim zero
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        u32x4 h2[2];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float4 bb = *reinterpret_cast<const float4*>(&b2s[32 * (op + u) + 8 * q + 4 * g]);
          h2[q >> 1][2 * (q & 1)] =
              static_cast<uint32_t>(to_bf16(fmaxf(a2[u][4 * q + 0] + bb.x, 0.f))) |
              (static_cast<uint32_t>(to_bf16(fmaxf(a2[u][4 * q + 1] + bb.y, 0.f))) << 16);
          h2[q >> 1][2 * (q & 1) + 1] =
              static_cast<uint32_t>(to_bf16(fmaxf(a2[u][4 * q + 2] + bb.z, 0.f))) |
              (static_cast<uint32_t>(to_bf16(fmaxf(a2[u][4 * q + 3] + bb.w, 0.f))) << 16);
        }
#pragma unroll
        for (int hh = 0; hh < 2; ++hh)
          a3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
              __builtin_bit_cast(bf16x8, w3s[(2 * (op + u) + hh) * 64 + lane]),
              __builtin_bit_cast(bf16x8, h2[hh]), a3, 0, 0, 0);
      }
    }
    // lane (g, row) holds z[row][4 g + i] in a3[i], i < 4
    if (live) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = 4 * g + i;
        if (c < dim) z[row * dim + c] = a3[i] + b3[c];
      }
    }
  }
}

bool fused_eligible(const MlpShape& sh, const float* x) {
  return (sh.h == 64 || sh.h == 128 || sh.h == 192 || sh.h == 256) && sh.d % 4 == 0 &&
         sh.dp <= kFusedMaxDp && sh.dim <= kMaxDim &&
         (reinterpret_cast<uintptr_t>(x) & 15u) == 0;
}

// Packs a trained model's BF16 weights for the fused kernel.
void pack_fused(const MlpShape& sh, const uint16_t* wh, DevBuf<uint16_t>& w1p,
                DevBuf<uint16_t>& w2p, DevBuf<uint16_t>& w3p, hipStream_t s) {
  const int64_t n1 = static_cast<int64_t>(sh.dp) * sh.h, n2 = static_cast<int64_t>(sh.h) * sh.h;
  const int64_t n3 = static_cast<int64_t>(sh.h) * 32;
  w1p.ensure(n1);
  w2p.ensure(n2);
  w3p.ensure(n3);
  hipLaunchKernelGGL(mlp_pack_kernel, dim3(blocks_for(n1 + n2 + n3, 256)), dim3(256), 0, s, wh, sh,
                     w1p.get(), w2p.get(), w3p.get());
  KERNEL_CHECK();
}

void run_fused(const MlpShape& sh, const float* params, const uint16_t* w1p, const uint16_t* w2p,
               const uint16_t* w3p, const float* mean, const float* inv_std, const float* x,
               int64_t rows, float* out, hipStream_t s) {
  if (rows <= 0) return;
  int dev = 0, cus = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t tiles = (rows + kFusedRowsPerWave - 1) / kFusedRowsPerWave;
  const unsigned blocks = static_cast<unsigned>(
      std::max<int64_t>(1, std::min<int64_t>((tiles + kFusedWaves - 1) / kFusedWaves, cus)));
  const float* b1 = params + sh.off_b1;
  const float* b2 = params + sh.off_b2;
  const float* b3 = params + sh.off_b3;
#define UMAPX_FUSED(HH)                                                                          \
  hipLaunchKernelGGL(mlp_infer_kernel<HH>, dim3(blocks), dim3(kFusedWaves * 64), 0, s, x, rows, \
                     sh.d, sh.dp, sh.dim, w1p, w2p, b1, b2, w3p, b3, mean, inv_std, out)
  switch (sh.h) {
    case 64: UMAPX_FUSED(64); break;
    case 128: UMAPX_FUSED(128); break;
    case 192: UMAPX_FUSED(192); break;
    case 256: UMAPX_FUSED(256); break;
    default: throw_invalid("the fused encoder supports mlp_hidden of 64, 128, 192 or 256");
  }
#undef UMAPX_FUSED
  KERNEL_CHECK();
}

// The unfused chain (the training forward) over row chunks.
void forward_chain(const MlpShape& sh, const float* params, const uint16_t* wh, const float* mean,
                   const float* inv_std, const float* x, int64_t rows, float* out, hipStream_t s) {
  constexpr int64_t kChunk = 65536;
  Pass ps;
  ps.ensure(sh, std::min<int64_t>(pad64(rows), kChunk));
  for (int64_t r0 = 0; r0 < rows; r0 += kChunk) {
    const int64_t cnt = std::min(kChunk, rows - r0);
    const int64_t cpad = pad64(cnt);
    hipLaunchKernelGGL(gather_rows_kernel, dim3(blocks_for(cpad * sh.dp, 256)), dim3(256), 0, s, x,
                       sh.d, sh.dp, static_cast<const int32_t*>(nullptr), r0, cnt, mean, inv_std,
                       ps.xb.get(), static_cast<const float*>(nullptr), sh.dim,
                       static_cast<float*>(nullptr));
    KERNEL_CHECK();
    if (cpad > cnt)
      HIP_CHECK(hipMemsetAsync(ps.xb.get() + cnt * sh.dp, 0, sizeof(uint16_t) * (cpad - cnt) * sh.dp, s));
    forward(sh, params, wh, ps, cpad, s);
    HIP_CHECK(hipMemcpyAsync(out + r0 * sh.dim, ps.z.get(), sizeof(float) * cnt * sh.dim,
                             hipMemcpyDeviceToDevice, s));
  }
}

}  // namespace

void pack_model(MlpModel& m, hipStream_t s) {
  m.packed = false;
  if (m.shape.h != 64 && m.shape.h != 128 && m.shape.h != 192 && m.shape.h != 256) return;
  if (m.shape.dp > kFusedMaxDp) return;
  pack_fused(m.shape, m.wh.get(), m.w1p, m.w2p, m.w3p, s);
  m.packed = true;
}

void mlp_forward_rows(const MlpModel& m, const float* x, int64_t rows, float* out, hipStream_t s) {
  if (!m.trained) throw Error(kNotFitted, "the learned transform has not been trained");
  const MlpShape& sh = m.shape;
  if (m.packed && fused_eligible(sh, x)) {
    run_fused(sh, m.params.get(), m.w1p.get(), m.w2p.get(), m.w3p.get(), m.mean.get(),
              m.inv_std.get(), x, rows, out, s);
    return;
  }
  forward_chain(sh, m.params.get(), m.wh.get(), m.mean.get(), m.inv_std.get(), x, rows, out, s);
}

void mlp_forward_host(int d, int h, int dim, const float* params, const float* mean,
                      const float* inv_std, const float* x, int64_t rows, bool fused, float* z) {
  check_shape(d, h, dim);
  if (rows < 1) throw_invalid("rows must be >= 1");
  const MlpShape sh = MlpShape::make(d, h, dim);
  hipStream_t s = nullptr;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  DevBuf<float> dp(sh.total), dmean(d), dinv(d), dx(rows * d), dz(rows * dim);
  DevBuf<uint16_t> dwh(sh.total);
  HIP_CHECK(hipMemcpyAsync(dp.get(), params, 4 * sh.total, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dmean.get(), mean, 4 * d, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dinv.get(), inv_std, 4 * d, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dx.get(), x, 4 * rows * d, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(sh.total, 256)), dim3(256), 0, s, dp.get(),
                     sh.total, dwh.get());
  KERNEL_CHECK();
  if (fused) {
    if (!fused_eligible(sh, dx.get()))
      throw_invalid("the fused encoder needs mlp_hidden in {64, 128, 192, 256}, d % 4 == 0 and d <= 1024");
    DevBuf<uint16_t> w1p, w2p;
    DevBuf<uint16_t> w3p;
    pack_fused(sh, dwh.get(), w1p, w2p, w3p, s);
    run_fused(sh, dp.get(), w1p.get(), w2p.get(), w3p.get(), dmean.get(), dinv.get(), dx.get(), rows,
              dz.get(), s);
    HIP_CHECK(hipMemcpyAsync(z, dz.get(), 4 * rows * dim, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } else {
    forward_chain(sh, dp.get(), dwh.get(), dmean.get(), dinv.get(), dx.get(), rows, dz.get(), s);
    HIP_CHECK(hipMemcpyAsync(z, dz.get(), 4 * rows * dim, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
}

void mlp_batch_grad(int d, int h, int dim, const float* params, const float* mean,
                    const float* inv_std, const float* x_rows, const float* y_rows, int n_edges,
                    int n_neg, float a, float b, float mse_weight, float* grads, float* loss,
                    float* z_out) {
  check_shape(d, h, dim);
  if (n_edges < 1 || n_neg < 0) throw_invalid("n_edges must be >= 1 and n_neg >= 0");
  const MlpShape sh = MlpShape::make(d, h, dim);
  const int64_t s_real = static_cast<int64_t>(n_edges) * (2 + n_neg);
  const int64_t s_pad = pad64(s_real);
  hipStream_t s = nullptr;
  HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  DevBuf<float> dp(sh.total), dg(sh.total), dmean(d), dinv(d), dx(s_real * d), dy(s_real * dim);
  DevBuf<uint16_t> dwh(sh.total);
  HIP_CHECK(hipMemcpyAsync(dp.get(), params, 4 * sh.total, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dmean.get(), mean, 4 * d, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dinv.get(), inv_std, 4 * d, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dx.get(), x_rows, 4 * s_real * d, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(dy.get(), y_rows, 4 * s_real * dim, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(sh.total, 256)), dim3(256), 0, s, dp.get(),
                     sh.total, dwh.get());
  Pass ps;
  ps.ensure(sh, s_pad);
  DevBuf<int32_t> rows(s_pad);
  {
    std::vector<int32_t> r(s_pad, -1);
    for (int64_t i = 0; i < s_real; ++i) r[i] = static_cast<int32_t>(i);
    HIP_CHECK(hipMemcpyAsync(rows.get(), r.data(), 4 * s_pad, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(gather_rows_kernel, dim3(blocks_for(s_pad * sh.dp, 256)), dim3(256), 0, s,
                       dx.get(), d, sh.dp, rows.get(), int64_t{0}, s_pad, dmean.get(), dinv.get(),
                       ps.xb.get(), dy.get(), dim, ps.yt.get());
    KERNEL_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));
  }
  HIP_CHECK(hipMemsetAsync(ps.dz.get(), 0, sizeof(float) * s_pad * dim, s));
  HIP_CHECK(hipMemsetAsync(ps.loss.get(), 0, sizeof(double), s));
  forward(sh, dp.get(), dwh.get(), ps, s_pad, s);
  hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks_for(n_edges, 256)), dim3(256), 0, s, ps.z.get(),
                     ps.yt.get(), n_edges, n_neg, dim, static_cast<int>(s_real), a, b, mse_weight,
                     ps.dz.get(), ps.loss.get());
  KERNEL_CHECK();
  backward(sh, dwh.get(), ps, s_pad, dg.get(), s);
  double l = 0.0;
  HIP_CHECK(hipMemcpyAsync(grads, dg.get(), 4 * sh.total, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(&l, ps.loss.get(), sizeof(double), hipMemcpyDeviceToHost, s));
  if (z_out) HIP_CHECK(hipMemcpyAsync(z_out, ps.z.get(), 4 * s_real * dim, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  *loss = static_cast<float>(l);
}

}  // namespace umapx
