// Spectral initialisation on gfx950.
//
// The reference runs, per connected component, 256 sequential f64 block
// iterations Y <- orth((I + D^-1/2 W D^-1/2) Y) with 18 columns and a
// twice-iterated modified Gram-Schmidt (rust_umap/src/lib.rs:4161-4326), then a
// Ritz eigen-solve.  Here all components of at least 128 rows iterate together
// on the GPU:
//   - the block lives in f32 (the product's traffic is one neighbour-row
//     gather per graph entry, so storage width is its cost); Gram matrices,
//     Cholesky factors, triangular solves and the Ritz problem are f64;
//   - the shifted normalised adjacency product is a CSR row gather;
//   - orthonormalisation is CholeskyQR per component (Gram matrix by a
//     segmented reduction over the component's rows, an 18x18 Cholesky and
//     its inverse, then every row times R^-1).  In exact arithmetic it
//     returns the same Q as Gram-Schmidt (the QR factor with a positive
//     diagonal is unique) and its pivots are the residual norms Gram-Schmidt
//     checks against 1e-12;
//   - the 18x18 Ritz problems are solved on the host.
// Smaller components take the reference's dense Laplacian path on the host,
// and tiny ones its random layout, exactly as lib.rs:4670-4764 decides.
#include "spectral.hpp"
#include "knn.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <cstring>
#include <numeric>
#include <thread>
#include <type_traits>

#include "hostmath.hpp"

namespace umapx {
namespace {

// The iterated block has KS = (n_components + 1) * 6 columns
// (lib.rs:4231-4234): the kernels below are instantiated for n_components
// 1 .. 8, i.e. KS = 12, 18, .., 54.
constexpr int kMaxDim = 8;
constexpr int kIters = 256;         // SPECTRAL_ITERATIVE_MAX_ITERS
// The block is re-orthonormalised after every kOrthEvery products instead of
// after each one.  For an upper-triangular R with a positive diagonal,
// orth(A R^-1) = orth(A), so in exact arithmetic the iterates equal the
// reference's; with the operator's spectrum in [0, 2] four unnormalised
// products raise the block's condition number kappa by at most
// (2 / lambda_18)^4, far inside what CholeskyQR handles in f64, and storing
// the block in f32 moves its weakest direction by about kappa * 2^-24.
constexpr int kOrthEvery = 4;
constexpr int kItemRows = 128;      // rows per work item of the iteration kernels
constexpr double kOrthoEps = 1e-12; // SPECTRAL_ORTHO_EPS

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t key) {
  return static_cast<double>(mix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------- components ---
__global__ void init_labels_kernel(int32_t* label, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) label[i] = static_cast<int32_t>(i);
}

__global__ void hook_kernel(const int32_t* __restrict__ head, const int32_t* __restrict__ tail,
                            int64_t ne, int32_t* __restrict__ label, int32_t* __restrict__ changed) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool hooked = false;
  if (e < ne) {
    const int32_t lu = label[head[e]];
    const int32_t lv = label[tail[e]];
    if (lu != lv) {
      const int32_t hi = lu > lv ? lu : lv;
      const int32_t lo = lu > lv ? lv : lu;
      atomicMin(&label[hi], lo);
      hooked = true;
    }
  }
  // one store per wave, not one per hooked edge
  if (__ballot(hooked) != 0ULL && (threadIdx.x & 63) == 0) *changed = 1;
}

// The same hooking over a symmetric edge list: each undirected edge is
// stored both ways, so only the direction whose head carries the larger label
// hooks.  Edges sorted by head put a head's run in consecutive lanes; the
// wave takes one atomic per distinct target (the head's label) with the
// smallest label offered for it, instead of one atomic per edge.
__global__ void hook_sym_kernel(const int32_t* __restrict__ head, const int32_t* __restrict__ tail,
                                int64_t ne, int32_t* __restrict__ label,
                                int32_t* __restrict__ changed) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int32_t hi = 0, lo = 0;
  bool act = false;
  if (e < ne) {
    hi = label[head[e]];
    lo = label[tail[e]];
    act = hi > lo;
  }
  uint64_t m = __ballot(act);
  if (m != 0ULL && lane == 0) *changed = 1;
  while (m != 0ULL) {
    const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
    const int32_t target = __shfl(hi, leader);
    const bool mine = act && hi == target;
    int32_t v = mine ? lo : INT32_MAX;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    if (lane == leader) atomicMin(&label[target], v);
    act = act && !mine;
    m = __ballot(act);
  }
}

__global__ void jump_kernel(int32_t* __restrict__ label, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t l = label[i];
  int32_t ll = label[l];
  while (l != ll) {
    l = ll;
    ll = label[l];
  }
  label[i] = l;
}

// --------------------------------------------------------------- degree ---
__global__ void degree_kernel(const int32_t* __restrict__ row_ptr, const float* __restrict__ w,
                              int64_t n, double* __restrict__ deg, double* __restrict__ inv) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) s += static_cast<double>(fmaxf(w[e], 0.f));
  deg[i] = s;
  inv[i] = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
}

// ------------------------------------------------ permuted block layout ---
// Rows of the components solved on the GPU are renumbered so that every
// component occupies a contiguous range of positions (members ascending).
// The block Y, the CSR graph and the normalised weights all live in this
// order, so a component's neighbour gathers stay inside its own range and
// hit L2.
__global__ void pos_of_kernel(const int32_t* __restrict__ perm, int64_t rows,
                              int32_t* __restrict__ pos_of) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < rows) pos_of[perm[p]] = static_cast<int32_t>(p);
}

__global__ void prow_count_kernel(const int32_t* __restrict__ perm, int64_t rows,
                                  const int32_t* __restrict__ row_ptr, int32_t* __restrict__ cnt) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < rows) cnt[p] = row_ptr[perm[p] + 1] - row_ptr[perm[p]];
}

// Normalised weights w_ij / sqrt(d_i d_j) in f64 (the value
// shifted_normalized_adjacency_mul forms, lib.rs:4211).
// One thread per graph entry; entries of vertices outside the GPU
// components (pos_of < 0) are skipped.
__global__ void prow_fill_kernel(const int32_t* __restrict__ head, int64_t ne,
                                 const int32_t* __restrict__ row_ptr,
                                 const int32_t* __restrict__ tail, const float* __restrict__ w,
                                 const double* __restrict__ inv, const int32_t* __restrict__ pos_of,
                                 const int32_t* __restrict__ prow_ptr, int32_t* __restrict__ pcol,
                                 float* __restrict__ pnw) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int32_t v = head[e];
  const int32_t p = pos_of[v];
  if (p < 0) return;
  const int32_t o = prow_ptr[p] + static_cast<int32_t>(e - row_ptr[v]);
  const int32_t t = tail[e];
  pcol[o] = pos_of[t];
  pnw[o] = (t == v) ? 0.f : static_cast<float>(static_cast<double>(fmaxf(w[e], 0.f)) * inv[v] * inv[t]);
}

__global__ void fill_i32_kernel(int32_t* __restrict__ a, int64_t n, int32_t v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Start block: column 0 = sqrt(degree), the rest U(-1, 1) (lib.rs:4252-4260;
// a counter-based stream replaces SmallRng).
template <int KS>
__global__ void start_block_kernel(const int32_t* __restrict__ perm, int64_t rows,
                                   const int32_t* __restrict__ slot_of_pos,
                                   const int32_t* __restrict__ local_of_pos,
                                   const uint64_t* __restrict__ seeds,
                                   const double* __restrict__ deg, float* __restrict__ yp) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= rows) return;
  const int32_t v = perm[p];
  const uint64_t key = seeds[slot_of_pos[p]] ^ 0x7A5B3D91CE42A117ULL;
  const uint64_t lr = static_cast<uint64_t>(local_of_pos[p]);
  float* row = yp + p * KS;
  row[0] = static_cast<float>(sqrt(deg[v]));
  for (int c = 1; c < KS; ++c)
    row[c] = static_cast<float>(
        2.0 * uniform01(key + lr * KS + static_cast<uint64_t>(c) * 0x9E3779B97F4A7C15ULL) - 1.0);
}

// Work items are laid out component by component.  Hardware hands block b to
// XCD b % 8, so remap blocks so that each XCD walks a contiguous run of items:
// an XCD then gathers from only a few components, whose rows fit in its L2.
__device__ __forceinline__ int xcd_item(int b, int nblocks) {
  const int q = nblocks / 8, r = nblocks % 8;
  const int x = b % 8;
  return x * q + (x < r ? x : r) + b / 8;
}

template <int KS>
__device__ __forceinline__ void pair_of(int tid, int& pi, int& pj) {
  int t = tid, i = 0;
  while (i < KS && t >= KS - i) {
    t -= KS - i;
    ++i;
  }
  pi = i;
  pj = i + t;
}

// z = y + D^-1/2 W D^-1/2 y, one launch per product.  A wave holds
// SpmmShape<KS>::kGroups lane groups of KS / 2 lanes (lanes past the last
// group idle), lane j of a group carrying columns 2j, 2j+1 (for 2-D output,
// KS = 18: 7 groups of 9).  Rows of at most kSeg entries are packed one per
// group.  A longer row (hub rows run to hundreds of entries after the
// union) takes a whole wave: its kSeg-entry pieces are dealt to the groups
// in turn, and the groups' partial sums are added in group order through
// LDS, so the result does not depend on scheduling.  For every piece the
// group first stages its column indices and weights in LDS with coalesced
// loads, then walks them (LDS broadcast reads) with 8 neighbour-row gathers
// in flight per lane: the vector memory pipeline is left to the neighbour
// rows, which bound the kernel.
//
// The block is stored in f32 (72 B a row): the product is a gather of one
// neighbour row per graph entry, so the row width sets the traffic.  The
// normalised weights are formed in f64 and rounded once; the Gram matrices
// and the Cholesky factors below stay in f64.
constexpr int kSeg = 64;
constexpr int kSpmmWaves = 4;
template <int KS>
struct SpmmShape {
  static constexpr int kLanes = KS / 2;        // lanes of one group
  static constexpr int kGroups = 64 / kLanes;  // groups of one wave
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Waves 0 .. n_long-1 take the long rows, the rest kGroups short rows each.
template <int KS>
__global__ __launch_bounds__(64 * kSpmmWaves) void spmm_kernel(
    int64_t n_long, int64_t n_short, const int32_t* __restrict__ long_list,
    const int32_t* __restrict__ short_list, const int32_t* __restrict__ prow_ptr,
    const int32_t* __restrict__ pcol, const float* __restrict__ pnw,
    const float* __restrict__ yp, float* __restrict__ zp) {
  constexpr int kLanes = SpmmShape<KS>::kLanes, kGroups = SpmmShape<KS>::kGroups;
  __shared__ int2 staged[kSpmmWaves][kGroups][kSeg];
  __shared__ float2 partial[kSpmmWaves][kGroups][kLanes];
  const int wv = threadIdx.x >> 6;
  const int wl = threadIdx.x & 63;
  const int g = wl / kLanes;
  const int lj = wl - kLanes * g;
  const int64_t w = static_cast<int64_t>(blockIdx.x) * kSpmmWaves + wv;
  const float2* y2 = reinterpret_cast<const float2*>(yp);
  float2* z2 = reinterpret_cast<float2*>(zp);
  int2* st = staged[wv][g < kGroups ? g : 0];

  // sum of w * y[col] over entries [e0, e0 + cnt) of one piece (cnt <= kSeg)
  auto piece = [&](int32_t e0, int32_t cnt, float2 acc) {
    for (int k = lj; k < cnt; k += kLanes) st[k] = make_int2(pcol[e0 + k], __float_as_int(pnw[e0 + k]));
    wave_sync();
    int k = 0;
    for (; k + 8 <= cnt; k += 8) {
      int2 c[8];
      float2 q[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) c[j] = st[k + j];
#pragma unroll
      for (int j = 0; j < 8; ++j) q[j] = y2[static_cast<int64_t>(c[j].x) * kLanes + lj];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float wt = __int_as_float(c[j].y);
        acc.x = fmaf(wt, q[j].x, acc.x);
        acc.y = fmaf(wt, q[j].y, acc.y);
      }
    }
    for (; k < cnt; ++k) {
      const int2 c = st[k];
      const float2 q = y2[static_cast<int64_t>(c.x) * kLanes + lj];
      const float wt = __int_as_float(c.y);
      acc.x = fmaf(wt, q.x, acc.x);
      acc.y = fmaf(wt, q.y, acc.y);
    }
    wave_sync();  // the next piece overwrites the staged entries
    return acc;
  };

  if (w < n_long) {
    const int32_t p = long_list[w];
    const int32_t r0 = prow_ptr[p], r1 = prow_ptr[p + 1];
    float2 acc = make_float2(0.f, 0.f);
    if (g < kGroups)
      for (int32_t e0 = r0 + g * kSeg; e0 < r1; e0 += kGroups * kSeg)
        acc = piece(e0, min(kSeg, r1 - e0), acc);
    if (g < kGroups) partial[wv][g][lj] = acc;
    wave_sync();
    if (g == 0) {
      float2 sum = y2[static_cast<int64_t>(p) * kLanes + lj];
#pragma unroll
      for (int q = 0; q < kGroups; ++q) {
        sum.x += partial[wv][q][lj].x;
        sum.y += partial[wv][q][lj].y;
      }
      z2[static_cast<int64_t>(p) * kLanes + lj] = sum;
    }
  } else {
    const int64_t i = (w - n_long) * kGroups + g;
    if (g >= kGroups || i >= n_short) return;
    const int32_t p = short_list[i];
    const int32_t r0 = prow_ptr[p], r1 = prow_ptr[p + 1];
    const float2 acc = piece(r0, r1 - r0, y2[static_cast<int64_t>(p) * kLanes + lj]);
    z2[static_cast<int64_t>(p) * kLanes + lj] = acc;
  }
}

// Rows of the slots still iterated (active[slot] != 0) only.
__global__ void long_rows_kernel(int64_t rows, const int32_t* __restrict__ prow_ptr,
                                 const int32_t* __restrict__ slot_of_pos,
                                 const int32_t* __restrict__ active, int32_t* __restrict__ list,
                                 int32_t* __restrict__ count) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < rows && active[slot_of_pos[p]] && prow_ptr[p + 1] - prow_ptr[p] > kSeg)
    list[atomicAdd(count, 1)] = static_cast<int32_t>(p);
}

__global__ void short_flag_kernel(int64_t rows, const int32_t* __restrict__ prow_ptr,
                                  const int32_t* __restrict__ slot_of_pos,
                                  const int32_t* __restrict__ active, int32_t* __restrict__ flag) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < rows)
    flag[p] = (active[slot_of_pos[p]] && prow_ptr[p + 1] - prow_ptr[p] <= kSeg) ? 1 : 0;
}

__global__ void short_fill_kernel(int64_t rows, const int32_t* __restrict__ flag,
                                  const int32_t* __restrict__ off, int32_t* __restrict__ list) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < rows && flag[p]) list[off[p]] = static_cast<int32_t>(p);
}

// The item's share of Z^T Z (or Y^T Z when CROSS) added to gram[slot], with
// f64 products of the stored f32 values; the rows are already computed.
template <bool CROSS, int KS>
__global__ __launch_bounds__(256) void gram_kernel(const int32_t* __restrict__ items,
                                                   const float* __restrict__ yp,
                                                   const float* __restrict__ zp,
                                                   double* __restrict__ gram) {
  __shared__ float sz[kItemRows][KS + 1];
  __shared__ float sy[CROSS ? kItemRows : 1][KS + 1];
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int32_t slot = items[item * 3 + 0];
  const int32_t p0 = items[item * 3 + 1];
  const int32_t cnt = items[item * 3 + 2] - p0;
  const int tid = threadIdx.x;
  for (int q = tid; q < cnt * KS; q += 256) {
    const int r = q / KS, c = q % KS;
    sz[r][c] = zp[(static_cast<int64_t>(p0) + r) * KS + c];
    if (CROSS) sy[r][c] = yp[(static_cast<int64_t>(p0) + r) * KS + c];
  }
  __syncthreads();
  for (int pair = tid; pair < KS * (KS + 1) / 2; pair += 256) {
    int pi, pj;
    pair_of<KS>(pair, pi, pj);
    double acc = 0.0;
    if (CROSS) {
      for (int r = 0; r < cnt; ++r)
        acc = fma(static_cast<double>(sy[r][pi]), static_cast<double>(sz[r][pj]), acc);
    } else {
      for (int r = 0; r < cnt; ++r)
        acc = fma(static_cast<double>(sz[r][pi]), static_cast<double>(sz[r][pj]), acc);
    }
    atomicAdd(&gram[static_cast<int64_t>(slot) * KS * KS + pi * KS + pj], acc);
  }
}

// v held by `lane`, as a wave-uniform value
__device__ __forceinline__ double lane_value(double v, int lane) {
  const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

// Cholesky factor of every slot's Gram matrix (upper R, G = R^T R) and its
// inverse, one wave per slot; gram[slot] is cleared for the next
// accumulation.  A pivot <= 1e-12 (or non-finite) flags the slot as failed:
// it equals the residual norm Gram-Schmidt checks in lib.rs:4186-4190.
//
// Lane c holds column c of the matrix in registers (rows 0..c; the rest stay
// zero) and the factorisation runs right-looking: step j takes the pivot from
// lane j, scales row j and subtracts its outer product from the trailing rows,
// with row j's entries broadcast by readlane.  The subtractions reach each
// entry in the order k = 0, 1, .., j-1, i.e. G[j][c] - sum_k R[k][j] R[k][c]
// in index order, and R^-1 is formed column by column by back substitution.
template <int KS>
__global__ __launch_bounds__(64) void chol_kernel(double* __restrict__ gram,
                                                  double* __restrict__ rinv,
                                                  int32_t* __restrict__ fail,
                                                  const int32_t* __restrict__ active) {
  if (!active[blockIdx.x]) return;  // a frozen slot (block-uniform)
  const int c = threadIdx.x;
  double* g = gram + static_cast<int64_t>(blockIdx.x) * KS * KS;
  double a[KS];
#pragma unroll
  for (int i = 0; i < KS; ++i) a[i] = (c < KS && i <= c) ? g[i * KS + c] : 0.0;
  bool bad = false;
#pragma unroll
  for (int j = 0; j < KS; ++j) {
    double piv = lane_value(a[j], j);
    if (!(piv > kOrthoEps) || !isfinite(piv)) {
      bad = true;
      piv = 1.0;
    }
    const double rjj = sqrt(piv);
    a[j] = c == j ? rjj : (c > j ? a[j] / rjj : a[j]);  // row j of R
#pragma unroll
    for (int i = j + 1; i < KS; ++i) {
      const double rji = lane_value(a[j], i);
      if (c >= i) a[i] = fma(-rji, a[j], a[i]);
    }
  }
  // column c of R^-1: x[i] = -(sum_{k=i+1..c} R[i][k] x[k]) / R[i][i], x[c] = 1 / R[c][c]
  double x[KS];
#pragma unroll
  for (int i = KS - 1; i >= 0; --i) {
    double v = 0.0;
#pragma unroll
    for (int k = i + 1; k < KS; ++k) v = fma(lane_value(a[i], k), x[k], v);
    const double rii = lane_value(a[i], i);
    x[i] = c == i ? 1.0 / rii : (c > i ? -v / rii : 0.0);
  }
  if (c < KS) {
    double* out = rinv + static_cast<int64_t>(blockIdx.x) * KS * KS;
#pragma unroll
    for (int i = 0; i < KS; ++i) out[i * KS + c] = x[i];
  }
  for (int q = c; q < KS * KS; q += 64) g[q] = 0.0;
  if (c == 0 && bad) fail[blockIdx.x] = 1;
}

// dst = src R^-1 for the item's rows (f64 products, stored f32).  One
// CholeskyQR pass suffices here: the Gram matrix is formed in f64 from the
// stored f32 rows, so the factor is accurate to about kappa^2 * 2^-53, far
// below the 2^-24 of storing the result; a second pass could not take the
// block's orthogonality past that storage rounding.
template <int KS>
__global__ __launch_bounds__(kItemRows) void rotate_kernel(const int32_t* __restrict__ items,
                                                           const double* __restrict__ rinv,
                                                           const float* __restrict__ src,
                                                           float* __restrict__ dst) {
  __shared__ double Ri[KS][KS];
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int32_t slot = items[item * 3 + 0];
  const int32_t p0 = items[item * 3 + 1];
  const int32_t cnt = items[item * 3 + 2] - p0;
  const int tid = threadIdx.x;
  const double* r = rinv + static_cast<int64_t>(slot) * KS * KS;
  for (int q = tid; q < KS * KS; q += kItemRows) Ri[q / KS][q % KS] = r[q];
  __syncthreads();
  if (tid >= cnt) return;
  const int64_t p = p0 + tid;
  const float* row = src + p * KS;
  double x[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) x[k] = row[k];
  float* out = dst + p * KS;
#pragma unroll
  for (int j = 0; j < KS; ++j) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k <= j; ++k) v = fma(x[k], Ri[k][j], v);
    out[j] = static_cast<float>(v);
  }
}

// coords = Y V_slot (KS x dim) in f64, stored f32; per-slot max |coord|
// (reduced over the lanes of a wave that share a slot, then one atomic).
template <int KS>
__global__ __launch_bounds__(256) void project_kernel(const int32_t* __restrict__ perm, int64_t rows,
                                                      const int32_t* __restrict__ slot_of_pos,
                                                      const float* __restrict__ y,
                                                      const double* __restrict__ vmat, int dim,
                                                      float* __restrict__ out,
                                                      float* __restrict__ maxabs) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int32_t slot = -1;
  float mx = 0.f;
  if (p < rows) {
    const int32_t v = perm[p];
    slot = slot_of_pos[p];
    const float* row = y + p * KS;
    const double* vm = vmat + static_cast<int64_t>(slot) * KS * dim;
    for (int c = 0; c < dim; ++c) {
      double s = 0.0;
      for (int b = 0; b < KS; ++b) s += static_cast<double>(row[b]) * vm[b * dim + c];
      const float f = static_cast<float>(s);
      out[static_cast<int64_t>(v) * dim + c] = f;
      mx = fmaxf(mx, fabsf(f));
    }
  }
  // positions are laid out slot by slot, so equal slots are adjacent lanes
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const float o = __shfl_down(mx, off);
    const int32_t os = __shfl_down(slot, off);
    if (lane + off < 64 && os == slot) mx = fmaxf(mx, o);
  }
  const int32_t prev = __shfl_up(slot, 1);
  if (slot >= 0 && (lane == 0 || prev != slot))
    atomicMax(reinterpret_cast<unsigned int*>(&maxabs[slot]), __float_as_uint(mx));
}

// Rows of the slots flagged in take[] copied from src to dst.
template <int KS>
__global__ void take_rows_kernel(int64_t rows, const int32_t* __restrict__ slot_of_pos,
                                 const int32_t* __restrict__ take, const float* __restrict__ src,
                                 float* __restrict__ dst) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= rows || !take[slot_of_pos[p]]) return;
  const int64_t o = p * KS;
  for (int b = 0; b < KS; ++b) dst[o + b] = src[o + b];
}

// scale[slot] = target[slot] / max|coord| of the slot, 1 when that is 0
// (noisy_scale_coords, lib.rs:4766-4779).
__global__ void scale_kernel(const float* __restrict__ maxabs, const float* __restrict__ target,
                             int n_slots, float* __restrict__ scale) {
  const int sl = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (sl >= n_slots) return;
  const float mx = maxabs[sl];
  scale[sl] = mx > 0.f ? fdiv(target[sl], mx) : 1.f;
}

// y = coords * scale[slot] + N(0, 1e-4) + anchor[slot]  (noisy_scale_coords /
// add_noise + anchor, lib.rs:4745-4791; counter-based normals).
__global__ void finish_kernel(const int32_t* __restrict__ perm, int64_t rows,
                              const int32_t* __restrict__ slot_of_pos,
                              const int32_t* __restrict__ local_of_pos,
                              const uint64_t* __restrict__ noise_seeds,
                              const float* __restrict__ scale, const float* __restrict__ anchor,
                              int dim, const float* __restrict__ coords, float* __restrict__ y) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= rows) return;
  const int32_t v = perm[p];
  const int32_t slot = slot_of_pos[p];
  const uint64_t key = noise_seeds[slot];
  for (int c = 0; c < dim; ++c) {
    const uint64_t ctr = static_cast<uint64_t>(local_of_pos[p]) * dim + c;
    double u1 = uniform01(key + 2 * ctr * 0x9E3779B97F4A7C15ULL);
    const double u2 = uniform01(key + (2 * ctr + 1) * 0x9E3779B97F4A7C15ULL);
    if (u1 <= 0.0) u1 = 1e-300;
    const double nz = 1e-4 * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    const float val = coords[static_cast<int64_t>(v) * dim + c] * scale[slot] + static_cast<float>(nz);
    y[static_cast<int64_t>(v) * dim + c] = val + anchor[slot * dim + c];
  }
}

__global__ void centroid_kernel(const int32_t* __restrict__ items, const int32_t* __restrict__ perm,
                                const float* __restrict__ x, int d, double* __restrict__ cent) {
  const int32_t slot = items[blockIdx.x * 3 + 0];
  const int32_t p0 = items[blockIdx.x * 3 + 1];
  const int32_t p1 = items[blockIdx.x * 3 + 2];
  for (int f = threadIdx.x; f < d; f += blockDim.x) {
    double s = 0.0;
    for (int32_t p = p0; p < p1; ++p) s += static_cast<double>(x[static_cast<int64_t>(perm[p]) * d + f]);
    atomicAdd(&cent[static_cast<int64_t>(slot) * d + f], s);
  }
}

// component_centroids_sparse (lib.rs:4487-4516): the item's rows' nonzeros
// added into their columns (zeros add nothing, so the sums equal the dense
// ones).
__global__ void centroid_csr_kernel(const int32_t* __restrict__ items,
                                    const int32_t* __restrict__ perm,
                                    const int64_t* __restrict__ ptr,
                                    const int32_t* __restrict__ ind,
                                    const float* __restrict__ val, int64_t d,
                                    double* __restrict__ cent) {
  const int32_t slot = items[blockIdx.x * 3 + 0];
  const int32_t p0 = items[blockIdx.x * 3 + 1];
  const int32_t p1 = items[blockIdx.x * 3 + 2];
  for (int32_t p = p0 + static_cast<int32_t>(threadIdx.x); p < p1; p += blockDim.x) {
    const int64_t row = perm[p];
    for (int64_t e = ptr[row]; e < ptr[row + 1]; ++e)
      atomicAdd(&cent[static_cast<int64_t>(slot) * d + ind[e]], static_cast<double>(val[e]));
  }
}

__global__ void scatter_rows_kernel(const int32_t* __restrict__ rows_idx, int64_t count, int dim,
                                    const float* __restrict__ vals, float* __restrict__ y) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count * dim) return;
  const int64_t r = i / dim;
  const int c = static_cast<int>(i % dim);
  y[static_cast<int64_t>(rows_idx[r]) * dim + c] = vals[i];
}

// ------------------------------------------------------------ host side ---
struct Comp {
  std::vector<int32_t> members;  // ascending
};

// lib.rs:3966-4023 normalized_laplacian_from_edges + lib.rs:3885-3926 with
// drop_first: dense host path for small components.
bool dense_embedding_from_edges(int64_t n, int dim,
                                const std::vector<std::tuple<int32_t, int32_t, float>>& edges,
                                std::vector<float>& coords) {
  if (n == 0 || edges.empty()) return false;
  std::vector<double> lap(static_cast<size_t>(n) * n, 0.0);
  for (int64_t i = 0; i < n; ++i) lap[i * n + i] = 1.0;
  std::vector<double> deg(n, 0.0);
  std::vector<std::pair<int64_t, int64_t>> pairs;
  for (const auto& e : edges) {
    const int64_t h = std::get<0>(e), t = std::get<1>(e);
    if (h == t) continue;
    const double w = static_cast<double>(std::max(std::get<2>(e), 0.0f));
    if (w <= 0.0) continue;
    const int64_t i = std::min(h, t), j = std::max(h, t);
    const double prev = lap[i * n + j];
    if (w > prev) {
      if (prev == 0.0) pairs.push_back({i, j});
      const double delta = w - prev;
      lap[i * n + j] = w;
      lap[j * n + i] = w;
      deg[i] += delta;
      deg[j] += delta;
    }
  }
  bool any = false;
  for (double g : deg) any |= g > 0.0;
  if (!any) return false;
  std::vector<double> inv(n);
  for (int64_t i = 0; i < n; ++i) inv[i] = deg[i] > 0.0 ? 1.0 / std::sqrt(deg[i]) : 0.0;
  for (const auto& pr : pairs) {
    const double v = lap[pr.first * n + pr.second] * inv[pr.first] * inv[pr.second];
    lap[pr.first * n + pr.second] = -v;
    lap[pr.second * n + pr.first] = -v;
  }
  for (int64_t i = 0; i < n; ++i) lap[i * n + i] = 1.0;
  std::vector<double> ev, vec;
  symmetric_eigen(lap, static_cast<int>(n), ev, vec);
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ev[a] < ev[b]; });
  if (static_cast<int64_t>(order.size()) <= dim + 1) return false;
  coords.assign(static_cast<size_t>(n) * dim, 0.f);
  for (int c = 0; c < dim; ++c)
    for (int64_t r = 0; r < n; ++r) coords[r * dim + c] = static_cast<float>(vec[r * n + order[c + 1]]);
  for (float v : coords)
    if (!std::isfinite(v)) return false;
  return true;
}

// lib.rs:3928-3964 + 4518-4550: meta layout of the component centroids.
bool meta_layout(const std::vector<std::vector<double>>& cent, int dim, uint64_t seed,
                 std::vector<float>& layout) {
  const size_t nc = cent.size();
  if (nc == 0) return false;
  if (nc == 1) {
    layout.assign(dim, 0.f);
    return true;
  }
  if (nc <= static_cast<size_t>(2 * dim)) {
    const size_t k = (nc + 1) / 2;
    layout.assign(nc * dim, 0.f);
    for (size_t i = 0; i < nc; ++i) layout[i * dim + (i % k)] = i < k ? 1.f : -1.f;
    return true;
  }
  std::vector<double> aff(nc * nc, 0.0);
  for (size_t i = 0; i < nc; ++i) {
    aff[i * nc + i] = 1.0;
    for (size_t j = i + 1; j < nc; ++j) {
      double d2 = 0.0;
      for (size_t f = 0; f < cent[i].size(); ++f) {
        const double df = cent[i][f] - cent[j][f];
        d2 += df * df;
      }
      aff[i * nc + j] = aff[j * nc + i] = std::exp(-d2);
    }
  }
  std::vector<double> deg(nc, 0.0);
  for (size_t i = 0; i < nc; ++i)
    for (size_t j = 0; j < nc; ++j) deg[i] += aff[i * nc + j];
  std::vector<double> inv(nc);
  for (size_t i = 0; i < nc; ++i) inv[i] = deg[i] > 0.0 ? 1.0 / std::sqrt(deg[i]) : 0.0;
  std::vector<double> lap(nc * nc, 0.0);
  for (size_t i = 0; i < nc; ++i) lap[i * nc + i] = 1.0;
  for (size_t i = 0; i < nc; ++i) {
    if (deg[i] <= 0.0) continue;
    for (size_t j = 0; j < nc; ++j) {
      if (deg[j] <= 0.0) continue;
      const double w = aff[i * nc + j];
      if (w > 0.0) lap[i * nc + j] -= w * inv[i] * inv[j];
    }
  }
  std::vector<double> ev, vec;
  symmetric_eigen(lap, static_cast<int>(nc), ev, vec);
  std::vector<int> order(nc);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ev[a] < ev[b]; });
  if (order.size() <= static_cast<size_t>(dim) + 1) return false;
  layout.assign(nc * dim, 0.f);
  for (int c = 0; c < dim; ++c)
    for (size_t r = 0; r < nc; ++r) layout[r * dim + c] = static_cast<float>(vec[r * nc + order[c + 1]]);
  // noisy_scale_coords(seed ^ 0xC85E7D6ADA318F21, 10, 1e-4)
  float mx = 0.f;
  for (float v : layout) mx = std::max(mx, std::fabs(v));
  const float expansion = mx > 0.f ? 10.0f / mx : 1.0f;
  SmallRng rng(seed ^ 0xC85E7D6ADA318F21ULL);
  for (auto& v : layout) v = v * expansion + static_cast<float>(rng.normal(1e-4));
  for (float v : layout)
    if (!std::isfinite(v)) return false;
  return true;
}

}  // namespace

int64_t connected_components_device(const Graph& g, std::vector<int32_t>& labels_host,
                                    SpectralWorkspace& ws, hipStream_t s) {
  const int64_t n = g.n;
  ws.label.ensure(n);
  ws.flag.ensure(1);
  hipLaunchKernelGGL(init_labels_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, ws.label.get(), n);
  KERNEL_CHECK();
  for (int it = 0; it < 10000; ++it) {
    HIP_CHECK(hipMemsetAsync(ws.flag.get(), 0, sizeof(int32_t), s));
    if (g.n_edges > 0 && g.symmetric)
      hipLaunchKernelGGL(hook_sym_kernel, dim3(blocks_for(g.n_edges, 256)), dim3(256), 0, s,
                         g.head.get(), g.tail.get(), g.n_edges, ws.label.get(), ws.flag.get());
    else if (g.n_edges > 0)
      hipLaunchKernelGGL(hook_kernel, dim3(blocks_for(g.n_edges, 256)), dim3(256), 0, s,
                         g.head.get(), g.tail.get(), g.n_edges, ws.label.get(), ws.flag.get());
    hipLaunchKernelGGL(jump_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, ws.label.get(), n);
    KERNEL_CHECK();
    int32_t changed = 0;
    HIP_CHECK(hipMemcpyAsync(&changed, ws.flag.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (!changed) break;
  }
  labels_host.resize(n);
  HIP_CHECK(hipMemcpyAsync(labels_host.data(), ws.label.get(), sizeof(int32_t) * n,
                           hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  int64_t nc = 0;
  for (int64_t i = 0; i < n; ++i)
    if (labels_host[i] == i) ++nc;
  return nc;
}

bool spectral_init_device(const float* x, int64_t n, int64_t d, int dim, const Graph& g,
                          uint64_t seed, float* y, SpectralWorkspace& ws, hipStream_t s,
                          SpectralReport* report, const CsrDevice* csr) {
  maybe_inject_fault("init");
  SpectralReport rep;
  if (n <= dim + 1 || g.n_edges == 0) return false;
  if (dim < 1 || dim > kMaxDim)
    throw_invalid("spectral init on the GPU supports n_components 1 to 8");
  // components, ordered by smallest member, members ascending
  std::vector<int32_t> labels;
  connected_components_device(g, labels, ws, s);
  // two passes: component of each row (labels[v] is overwritten by it) and
  // sizes, then members filled into reserved lists
  std::vector<int32_t> comp_index(n, -1);
  std::vector<int32_t> sizes;
  for (int64_t v = 0; v < n; ++v) {
    int32_t& c = comp_index[labels[v]];
    if (c < 0) {
      c = static_cast<int32_t>(sizes.size());
      sizes.push_back(0);
    }
    labels[v] = c;
    ++sizes[c];
  }
  std::vector<Comp> comps(sizes.size());
  for (size_t c = 0; c < sizes.size(); ++c) comps[c].members.reserve(sizes[c]);
  for (int64_t v = 0; v < n; ++v) comps[labels[v]].members.push_back(static_cast<int32_t>(v));
  const int64_t nc = static_cast<int64_t>(comps.size());
  rep.n_components = nc;
  const bool multi = nc > 1;

  // per-component plan: 0 random, 1 dense host, 2 iterative GPU
  std::vector<int> plan(nc, 2);
  std::vector<uint64_t> comp_seed(nc, seed);
  std::vector<float> layout;  // anchors
  std::vector<float> range(nc, 1.f);
  if (multi) {
    // centroids on the GPU (f64), meta layout on the host
    std::vector<int32_t> perm_all;
    perm_all.reserve(n);
    std::vector<int32_t> items;
    for (int64_t c = 0; c < nc; ++c) {
      const int32_t start = static_cast<int32_t>(perm_all.size());
      perm_all.insert(perm_all.end(), comps[c].members.begin(), comps[c].members.end());
      const int32_t end = static_cast<int32_t>(perm_all.size());
      for (int32_t p = start; p < end; p += kItemRows) {
        items.push_back(static_cast<int32_t>(c));
        items.push_back(p);
        items.push_back(std::min(end, p + kItemRows));
      }
    }
    ws.perm.ensure(perm_all.size());
    ws.items.ensure(items.size());
    ws.cent.ensure(static_cast<size_t>(nc) * d);
    HIP_CHECK(hipMemcpyAsync(ws.perm.get(), perm_all.data(), perm_all.size() * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(ws.items.get(), items.data(), items.size() * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemsetAsync(ws.cent.get(), 0, sizeof(double) * nc * d, s));
    if (csr)
      hipLaunchKernelGGL(centroid_csr_kernel, dim3(static_cast<unsigned>(items.size() / 3)),
                         dim3(256), 0, s, ws.items.get(), ws.perm.get(), csr->ptr.get(),
                         csr->ind.get(), csr->val.get(), d, ws.cent.get());
    else
      hipLaunchKernelGGL(centroid_kernel, dim3(static_cast<unsigned>(items.size() / 3)), dim3(256),
                         0, s, ws.items.get(), ws.perm.get(), x, static_cast<int>(d),
                         ws.cent.get());
    KERNEL_CHECK();
    std::vector<double> csum(static_cast<size_t>(nc) * d);
    HIP_CHECK(hipMemcpyAsync(csum.data(), ws.cent.get(), sizeof(double) * nc * d, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<std::vector<double>> cent(nc, std::vector<double>(d));
    for (int64_t c = 0; c < nc; ++c) {
      const double sc = 1.0 / static_cast<double>(comps[c].members.size());
      for (int64_t f = 0; f < d; ++f) cent[c][f] = csum[c * d + f] * sc;
    }
    if (!meta_layout(cent, dim, seed ^ 0xA13F52A92D4CB801ULL, layout)) return false;
    for (int64_t c = 0; c < nc; ++c) {
      float r = INFINITY;
      for (int64_t o = 0; o < nc; ++o) {
        if (o == c) continue;
        float sq = 0.f;
        for (int k = 0; k < dim; ++k) {
          const float df = layout[c * dim + k] - layout[o * dim + k];
          sq += df * df;
        }
        const float dd = std::sqrt(sq);
        if (dd > 0.f) r = std::min(r, dd / 2.0f);
      }
      if (!std::isfinite(r) || r <= 0.f) r = 1.f;
      range[c] = r;
      const size_t cs = comps[c].members.size();
      if (cs < static_cast<size_t>(2 * dim) || cs <= static_cast<size_t>(dim) + 1) {
        plan[c] = 0;
      } else {
        comp_seed[c] = seed ^ ((static_cast<uint64_t>(c) + 1) * 0xBF58476D1CE4E5B9ULL);
        plan[c] = cs >= 128 ? 2 : 1;
      }
    }
  } else {
    layout.assign(dim, 0.f);
    const int64_t thr = d <= 8 ? 2049 : 0;
    plan[0] = n >= thr ? 2 : 1;
  }

  std::vector<float> y_host_rows;     // rows produced on the host
  std::vector<int32_t> y_host_index;  // their global row ids
  std::vector<std::tuple<int32_t, int32_t, float>> all_edges;
  bool edges_loaded = false;
  auto load_edges = [&]() {
    if (edges_loaded) return;
    std::vector<int32_t> hh(g.n_edges), tt(g.n_edges);
    std::vector<float> ww(g.n_edges);
    HIP_CHECK(hipMemcpyAsync(hh.data(), g.head.get(), 4 * g.n_edges, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(tt.data(), g.tail.get(), 4 * g.n_edges, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(ww.data(), g.weight.get(), 4 * g.n_edges, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    all_edges.resize(g.n_edges);
    for (int64_t e = 0; e < g.n_edges; ++e) all_edges[e] = {hh[e], tt[e], ww[e]};
    edges_loaded = true;
  };
  auto dense_component = [&](int64_t c, std::vector<float>& coords) {
    load_edges();
    const auto& mem = comps[c].members;
    std::vector<int32_t> local(n, -1);
    for (size_t i = 0; i < mem.size(); ++i) local[mem[i]] = static_cast<int32_t>(i);
    std::vector<std::tuple<int32_t, int32_t, float>> ce;
    for (const auto& e : all_edges) {
      const int32_t h = std::get<0>(e), t = std::get<1>(e);
      if (local[h] >= 0 && local[t] >= 0) ce.emplace_back(local[h], local[t], std::get<2>(e));
    }
    return dense_embedding_from_edges(static_cast<int64_t>(mem.size()), dim, ce, coords);
  };

  // host-side components (random / dense)
  for (int64_t c = 0; c < nc; ++c) {
    if (plan[c] == 2) continue;
    const auto& mem = comps[c].members;
    std::vector<float> lc;
    if (plan[c] == 0) {
      lc = random_init(static_cast<int64_t>(mem.size()), dim,
                       seed ^ ((static_cast<uint64_t>(c) + 1) * 0x9E3779B9ULL), -range[c], range[c]);
      ++rep.n_random;
    } else {
      if (!dense_component(c, lc)) return false;
      ++rep.n_dense;
      float mx = 0.f;
      for (float v : lc) mx = std::max(mx, std::fabs(v));
      if (multi) {
        const float expansion = mx > 0.f ? range[c] / mx : 1.f;
        for (auto& v : lc) v *= expansion;
        SmallRng rng(seed ^ ((static_cast<uint64_t>(c) + 1) * 0x94D049BBULL));
        for (auto& v : lc) v += static_cast<float>(rng.normal(1e-4));
      } else {
        const float expansion = mx > 0.f ? 10.0f / mx : 1.f;
        SmallRng rng(seed ^ 0x5DEECE66DULL);
        for (auto& v : lc) v = v * expansion + static_cast<float>(rng.normal(1e-4));
      }
    }
    for (size_t i = 0; i < mem.size(); ++i) {
      y_host_index.push_back(mem[i]);
      for (int k = 0; k < dim; ++k) y_host_rows.push_back(lc[i * dim + k] + layout[c * dim + k]);
    }
  }

  // GPU iterative components
  std::vector<int64_t> gpu_comps;
  for (int64_t c = 0; c < nc; ++c)
    if (plan[c] == 2) gpu_comps.push_back(c);
  const int n_slots = static_cast<int>(gpu_comps.size());
  // the iterated components, for the block width KS = (dim + 1) * 6
  auto iterate = [&](auto ks_tag) -> bool {
    constexpr int KS = decltype(ks_tag)::value;
    constexpr int kGroups = SpmmShape<KS>::kGroups;
    std::vector<int32_t> perm, slot_of_pos, local_of_pos, items;
    std::vector<uint64_t> seeds(n_slots), noise_seeds(n_slots);
    size_t total_rows = 0;
    for (int sl = 0; sl < n_slots; ++sl) total_rows += comps[gpu_comps[sl]].members.size();
    perm.resize(total_rows);
    slot_of_pos.resize(total_rows);
    local_of_pos.resize(total_rows);
    items.reserve(3 * (total_rows / kItemRows + n_slots));
    size_t at = 0;
    for (int sl = 0; sl < n_slots; ++sl) {
      const int64_t c = gpu_comps[sl];
      const int32_t start = static_cast<int32_t>(at);
      const auto& mem = comps[c].members;
      if (static_cast<int64_t>(mem.size()) <= KS) return false;
      std::copy(mem.begin(), mem.end(), perm.begin() + at);
      std::fill(slot_of_pos.begin() + at, slot_of_pos.begin() + at + mem.size(), sl);
      std::iota(local_of_pos.begin() + at, local_of_pos.begin() + at + mem.size(), 0);
      at += mem.size();
      const int32_t end = static_cast<int32_t>(at);
      for (int32_t p = start; p < end; p += kItemRows) {
        items.push_back(sl);
        items.push_back(p);
        items.push_back(std::min(end, p + kItemRows));
      }
      seeds[sl] = comp_seed[c];
      noise_seeds[sl] = multi ? (seed ^ ((static_cast<uint64_t>(c) + 1) * 0x94D049BBULL))
                              : (seed ^ 0x5DEECE66DULL);
    }
    const int64_t rows = static_cast<int64_t>(perm.size());
    const int n_items = static_cast<int>(items.size() / 3);
    ws.perm.ensure(rows);
    ws.slot_of.ensure(2 * rows);
    ws.items.ensure(items.size());
    ws.seeds.ensure(2 * n_slots);
    HIP_CHECK(hipMemcpyAsync(ws.perm.get(), perm.data(), 4 * rows, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(ws.slot_of.get(), slot_of_pos.data(), 4 * rows, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(ws.slot_of.get() + rows, local_of_pos.data(), 4 * rows, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(ws.items.get(), items.data(), 4 * items.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(ws.seeds.get(), seeds.data(), 8 * n_slots, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(ws.seeds.get() + n_slots, noise_seeds.data(), 8 * n_slots, hipMemcpyHostToDevice, s));
    const int32_t* d_perm = ws.perm.get();
    const int32_t* d_slot = ws.slot_of.get();
    const int32_t* d_local = ws.slot_of.get() + rows;

    ws.deg.ensure(n);
    ws.inv.ensure(n);
    ws.y.ensure(static_cast<size_t>(rows) * KS);
    ws.z.ensure(static_cast<size_t>(rows) * KS);
    ws.gram.ensure(static_cast<size_t>(n_slots) * KS * KS);
    ws.fail.ensure(n_slots);
    ws.label.ensure(n);  // reused as pos_of
    hipLaunchKernelGGL(degree_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, g.row_ptr.get(),
                       g.weight.get(), n, ws.deg.get(), ws.inv.get());
    hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, ws.label.get(), n,
                       -1);
    hipLaunchKernelGGL(pos_of_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, s, d_perm, rows,
                       ws.label.get());
    KERNEL_CHECK();
    // permuted CSR with normalised f64 weights
    DevBuf<int32_t> pcnt(rows), prow_ptr(rows + 1);
    hipLaunchKernelGGL(prow_count_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, s, d_perm, rows,
                       g.row_ptr.get(), pcnt.get());
    KERNEL_CHECK();
    ScanWorkspace scan_ws;
    const int64_t pe = exclusive_scan(pcnt.get(), prow_ptr.get(), rows, scan_ws, s);
    const int32_t pe32 = static_cast<int32_t>(pe);
    HIP_CHECK(hipMemcpyAsync(prow_ptr.get() + rows, &pe32, 4, hipMemcpyHostToDevice, s));
    DevBuf<int32_t> pcol(std::max<int64_t>(pe, 1));
    DevBuf<float> pnw(std::max<int64_t>(pe, 1));
    hipLaunchKernelGGL(prow_fill_kernel, dim3(blocks_for(g.n_edges, 256)), dim3(256), 0, s,
                       g.head.get(), g.n_edges, g.row_ptr.get(), g.tail.get(), g.weight.get(),
                       ws.inv.get(), ws.label.get(), prow_ptr.get(), pcol.get(), pnw.get());
    hipLaunchKernelGGL(start_block_kernel<KS>, dim3(blocks_for(rows, 256)), dim3(256), 0, s, d_perm,
                       rows, d_slot, d_local, ws.seeds.get(), ws.deg.get(), ws.y.get());
    KERNEL_CHECK();
    const int glen = n_slots * KS * KS;
    double* g1 = ws.gram.get();
    HIP_CHECK(hipMemsetAsync(ws.fail.get(), 0, 4 * n_slots, s));
    HIP_CHECK(hipMemsetAsync(ws.gram.get(), 0, sizeof(double) * glen, s));
    float* yb = ws.y.get();
    float* zb = ws.z.get();
    const int32_t* it_items = ws.items.get();
    int n_act_items = n_items;
    ws.rmat.ensure(static_cast<size_t>(glen));
    // slots still iterated: a slot whose Ritz vectors have stopped moving is
    // frozen (its rows and Ritz vectors kept, its rows left out of the
    // products); components share no edges, so the others' products are
    // unchanged
    std::vector<int32_t> active(n_slots, 1);
    DevBuf<int32_t> d_active(n_slots), d_take(n_slots);
    HIP_CHECK(hipMemcpyAsync(d_active.get(), active.data(), 4 * n_slots, hipMemcpyHostToDevice, s));
    // Orthonormalise Z into Y (CholeskyQR): g1 holds Z^T Z on entry and is
    // cleared for the next accumulation.
    auto orth_into_y = [&]() {
      hipLaunchKernelGGL(chol_kernel<KS>, dim3(n_slots), dim3(64), 0, s, g1, ws.rmat.get(),
                         ws.fail.get(), d_active.get());
      hipLaunchKernelGGL(rotate_kernel<KS>, dim3(n_act_items), dim3(kItemRows), 0, s, it_items,
                         ws.rmat.get(), zb, yb);
      KERNEL_CHECK();
    };
    // start block: Z = Y, g1 = Y^T Y
    HIP_CHECK(hipMemcpyAsync(zb, yb, sizeof(float) * rows * KS, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL((gram_kernel<false, KS>), dim3(n_items), dim3(256), 0, s, it_items, yb, zb, g1);
    KERNEL_CHECK();
    orth_into_y();
    // rows of more than kSeg entries (one wave each) and the others in order,
    // of the active slots
    DevBuf<int32_t> long_list(rows), long_count(1), short_flag(rows), short_off(rows),
        short_list(rows);
    int32_t n_long = 0;
    int64_t n_short = 0;
    auto build_lists = [&]() {
      HIP_CHECK(hipMemsetAsync(long_count.get(), 0, 4, s));
      hipLaunchKernelGGL(long_rows_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows,
                         prow_ptr.get(), d_slot, d_active.get(), long_list.get(), long_count.get());
      hipLaunchKernelGGL(short_flag_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows,
                         prow_ptr.get(), d_slot, d_active.get(), short_flag.get());
      KERNEL_CHECK();
      n_short = exclusive_scan(short_flag.get(), short_off.get(), rows, scan_ws, s);
      hipLaunchKernelGGL(short_fill_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows,
                         short_flag.get(), short_off.get(), short_list.get());
      KERNEL_CHECK();
      HIP_CHECK(hipMemcpyAsync(&n_long, long_count.get(), 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
    };
    build_lists();
    auto product = [&](const float* src, float* dst) {
      const int64_t spmm_waves = n_long + (n_short + kGroups - 1) / kGroups;
      if (spmm_waves == 0) return;
      hipLaunchKernelGGL(spmm_kernel<KS>, dim3(blocks_for(spmm_waves, kSpmmWaves)),
                         dim3(64 * kSpmmWaves), 0, s, static_cast<int64_t>(n_long), n_short,
                         long_list.get(), short_list.get(), prow_ptr.get(), pcol.get(), pnw.get(),
                         src, dst);
      KERNEL_CHECK();
    };
    std::vector<double> vmat(static_cast<size_t>(n_slots) * KS * dim, 0.0);
    std::vector<float> anchor(static_cast<size_t>(n_slots) * dim);
    std::vector<double> ritz(static_cast<size_t>(n_slots) * KS * KS);
    std::vector<int32_t> fail(n_slots);
    // Ritz vectors of slot sl from its Ritz matrix Y^T (I + N) Y (upper
    // triangle in ritz): the eigenvectors of the 2nd .. (dim+1)-th largest
    // eigenvalues, into out (KS x dim)
    auto ritz_vectors = [&](int sl, double* out) {
      std::vector<double> a(ritz.begin() + static_cast<size_t>(sl) * KS * KS,
                            ritz.begin() + static_cast<size_t>(sl + 1) * KS * KS);
      for (int i = 0; i < KS; ++i)
        for (int j = 0; j < i; ++j) a[i * KS + j] = a[j * KS + i];
      std::vector<double> ev, vec;
      symmetric_eigen(a, KS, ev, vec);
      std::vector<int> order(KS);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return ev[q] < ev[p]; });
      for (int b = 0; b < KS; ++b)
        for (int k = 0; k < dim; ++k) out[static_cast<size_t>(b) * dim + k] = vec[b * KS + order[k + 1]];
    };
    // independent per slot: with many slots they are spread over host threads
    auto for_slots = [&](const std::vector<int>& list, auto fn) {
      const int cnt = static_cast<int>(list.size());
      const int n_threads = std::min(8, cnt / 4);
      if (n_threads <= 1) {
        for (int sl : list) fn(sl);
        return;
      }
      std::atomic<int> next{0};
      std::vector<std::exception_ptr> errors(n_threads);
      std::vector<std::thread> pool;
      for (int t = 0; t < n_threads; ++t)
        pool.emplace_back([&, t] {
          try {
            for (int i = next++; i < cnt; i = next++) fn(list[i]);
          } catch (...) {
            errors[t] = std::current_exception();
            next = cnt;  // the other threads stop at their next slot
          }
        });
      for (auto& th : pool) th.join();
      for (auto& err : errors)
        if (err) std::rethrow_exception(err);
    };
    // Ritz matrices of the active slots (Z = (I + N) Y computed here) into ritz
    auto ritz_matrices = [&]() {
      product(yb, zb);
      hipLaunchKernelGGL((gram_kernel<true, KS>), dim3(n_act_items), dim3(256), 0, s, it_items, yb,
                         zb, g1);
      KERNEL_CHECK();
      HIP_CHECK(hipMemcpyAsync(ritz.data(), g1, 8 * ritz.size(), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipMemcpyAsync(fail.data(), ws.fail.get(), 4 * n_slots, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
    };
    // the frozen and final rows of Y
    DevBuf<float> yfin(static_cast<size_t>(rows) * KS);
    std::vector<int32_t> take(n_slots);
    // Components that the meta layout packs into a box of half-width
    // range <= kTinyRange (components placed within ~1e-4 of each other,
    // e.g. 30 of config 3's 32) are scaled into that box, where any layout
    // differs from any other by at most 2 * range, the size of the N(0, 1e-4)
    // noise the reference adds to every coordinate (lib.rs:4706-4740).  Their
    // iteration stops at kTinyIters products: their rows leave the products
    // (components share no edges, so the other components' iterates are
    // unchanged) and their Ritz vectors are taken there.
    constexpr float kTinyRange = 1e-4f;
    constexpr int kTinyIters = 16;
    std::vector<int32_t> tiny(n_slots, 0);
    int n_tiny = 0;
    if (multi)
      for (int sl = 0; sl < n_slots; ++sl)
        if (range[gpu_comps[sl]] <= kTinyRange) {
          tiny[sl] = 1;
          ++n_tiny;
        }
    // freezes the tiny slots: Ritz vectors into vmat, rows into yfin, the
    // active lists rebuilt; false when no slot is left to iterate
    auto freeze_tiny = [&]() -> bool {
      ritz_matrices();
      std::vector<int> list;
      std::fill(take.begin(), take.end(), 0);
      for (int sl = 0; sl < n_slots; ++sl)
        if (tiny[sl] && !fail[sl]) {
          list.push_back(sl);
          take[sl] = 1;
          active[sl] = 0;
        }
      for_slots(list, [&](int sl) { ritz_vectors(sl, vmat.data() + static_cast<size_t>(sl) * KS * dim); });
      HIP_CHECK(hipMemsetAsync(g1, 0, sizeof(double) * glen, s));
      if (list.empty()) return true;
      HIP_CHECK(hipMemcpyAsync(d_take.get(), take.data(), 4 * n_slots, hipMemcpyHostToDevice, s));
      HIP_CHECK(hipMemcpyAsync(d_active.get(), active.data(), 4 * n_slots, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(take_rows_kernel<KS>, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows,
                         d_slot, d_take.get(), yb, yfin.get());
      KERNEL_CHECK();
      std::vector<int32_t> act_items;
      for (size_t i = 0; i < items.size(); i += 3)
        if (active[items[i]]) act_items.insert(act_items.end(), items.begin() + i, items.begin() + i + 3);
      n_act_items = static_cast<int>(act_items.size() / 3);
      if (n_act_items > 0)
        HIP_CHECK(hipMemcpyAsync(ws.items.get(), act_items.data(), 4 * act_items.size(),
                                 hipMemcpyHostToDevice, s));
      build_lists();  // synchronises (the host item list is freed on return)
      return n_act_items > 0;
    };
    for (int it = 0; it < kIters; ++it) {
      product(yb, zb);
      if ((it + 1) % kOrthEvery == 0 || it + 1 == kIters) {
        hipLaunchKernelGGL((gram_kernel<false, KS>), dim3(n_act_items), dim3(256), 0, s, it_items,
                           yb, zb, g1);
        KERNEL_CHECK();
        orth_into_y();
      } else {
        std::swap(yb, zb);
      }
      static_assert(kTinyIters % kOrthEvery == 0, "the tiny slots stop at an orthonormalised block");
      if (n_tiny > 0 && it + 1 == kTinyIters && !freeze_tiny()) break;  // every slot frozen
    }
    // the Ritz vectors of the slots iterated to the end (lib.rs:4267-4300)
    std::vector<int> ok_slots, last;
    if (n_act_items > 0) ritz_matrices();
    for (int sl = 0; sl < n_slots; ++sl) {
      const int64_t c = gpu_comps[sl];
      if (!active[sl]) {
        ok_slots.push_back(sl);  // frozen: its Ritz vectors are in vmat
        for (int k = 0; k < dim; ++k) anchor[sl * dim + k] = layout[c * dim + k];
        continue;
      }
      if (fail[sl]) {
        // lib.rs:4339-4348: a failed subspace iteration falls back to the
        // dense Laplacian only up to 4096 rows
        const auto& mem = comps[c].members;
        if (mem.size() > 4096) return false;
        std::vector<float> lc;
        if (!dense_component(c, lc)) return false;
        float mxa = 0.f;
        for (float v : lc) mxa = std::max(mxa, std::fabs(v));
        if (multi) {
          const float expansion = mxa > 0.f ? range[c] / mxa : 1.f;
          SmallRng rng(seed ^ ((static_cast<uint64_t>(c) + 1) * 0x94D049BBULL));
          for (auto& v : lc) v = v * expansion + static_cast<float>(rng.normal(1e-4));
        } else {
          const float expansion = mxa > 0.f ? 10.0f / mxa : 1.f;
          SmallRng rng(seed ^ 0x5DEECE66DULL);
          for (auto& v : lc) v = v * expansion + static_cast<float>(rng.normal(1e-4));
        }
        for (size_t i = 0; i < mem.size(); ++i) {
          y_host_index.push_back(mem[i]);
          for (int k = 0; k < dim; ++k) y_host_rows.push_back(lc[i * dim + k] + layout[c * dim + k]);
        }
        ++rep.n_dense;
        for (int b = 0; b < KS; ++b)
          for (int k = 0; k < dim; ++k) vmat[(static_cast<size_t>(sl) * KS + b) * dim + k] = 0.0;
        continue;
      }
      ok_slots.push_back(sl);
      last.push_back(sl);
      for (int k = 0; k < dim; ++k) anchor[sl * dim + k] = layout[c * dim + k];
    }
    for_slots(last, [&](int sl) { ritz_vectors(sl, vmat.data() + static_cast<size_t>(sl) * KS * dim); });
    if (!last.empty()) {
      std::fill(take.begin(), take.end(), 0);
      for (int sl : last) take[sl] = 1;
      HIP_CHECK(hipMemcpyAsync(d_take.get(), take.data(), 4 * n_slots, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(take_rows_kernel<KS>, dim3(blocks_for(rows, 256)), dim3(256), 0, s, rows,
                         d_slot, d_take.get(), yb, yfin.get());
      KERNEL_CHECK();
    }
    yb = yfin.get();
    const int n_ok = static_cast<int>(ok_slots.size());
    rep.n_iterative += n_ok;
    ws.vmat.ensure(vmat.size());
    ws.maxabs.ensure(n_slots);
    ws.scale.ensure(n_slots);
    // anchors then the per-slot target extent in one upload; the scale
    // target / max|coord| is formed on the device (no round trip)
    for (int sl = 0; sl < n_slots; ++sl) anchor.push_back(multi ? range[gpu_comps[sl]] : 10.0f);
    ws.anchor.ensure(anchor.size());
    DevBuf<float> coords(static_cast<size_t>(n) * dim);
    HIP_CHECK(hipMemcpyAsync(ws.vmat.get(), vmat.data(), 8 * vmat.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(ws.anchor.get(), anchor.data(), 4 * anchor.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemsetAsync(ws.maxabs.get(), 0, 4 * n_slots, s));
    hipLaunchKernelGGL(project_kernel<KS>, dim3(blocks_for(rows, 256)), dim3(256), 0, s, d_perm, rows,
                       d_slot, yb, ws.vmat.get(), dim, coords.get(), ws.maxabs.get());
    hipLaunchKernelGGL(scale_kernel, dim3(blocks_for(n_slots, 64)), dim3(64), 0, s,
                       ws.maxabs.get(), ws.anchor.get() + static_cast<size_t>(n_slots) * dim, n_slots,
                       ws.scale.get());
    KERNEL_CHECK();
    hipLaunchKernelGGL(finish_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, s, d_perm, rows, d_slot,
                       d_local, ws.seeds.get() + n_slots, ws.scale.get(), ws.anchor.get(), dim,
                       coords.get(), y);
    KERNEL_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));
    return true;
  };
  if (n_slots > 0) {
    bool ok = false;
    switch (dim) {
      case 1: ok = iterate(std::integral_constant<int, 12>{}); break;
      case 2: ok = iterate(std::integral_constant<int, 18>{}); break;
      case 3: ok = iterate(std::integral_constant<int, 24>{}); break;
      case 4: ok = iterate(std::integral_constant<int, 30>{}); break;
      case 5: ok = iterate(std::integral_constant<int, 36>{}); break;
      case 6: ok = iterate(std::integral_constant<int, 42>{}); break;
      case 7: ok = iterate(std::integral_constant<int, 48>{}); break;
      case 8: ok = iterate(std::integral_constant<int, 54>{}); break;
      default: throw_invalid("spectral init on the GPU supports n_components 1 to 8");
    }
    if (!ok) return false;
  }
  if (!y_host_index.empty()) {
    DevBuf<int32_t> ridx(y_host_index.size());
    DevBuf<float> rv(y_host_rows.size());
    HIP_CHECK(hipMemcpyAsync(ridx.get(), y_host_index.data(), 4 * y_host_index.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(rv.get(), y_host_rows.data(), 4 * y_host_rows.size(), hipMemcpyHostToDevice, s));
    const int64_t cnt = static_cast<int64_t>(y_host_index.size());
    hipLaunchKernelGGL(scatter_rows_kernel, dim3(blocks_for(cnt * dim, 256)), dim3(256), 0, s, ridx.get(),
                       cnt, dim, rv.get(), y);
    KERNEL_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));
  }
  rep.used_spectral = true;
  if (report) *report = rep;
  return true;
}

}  // namespace umapx
