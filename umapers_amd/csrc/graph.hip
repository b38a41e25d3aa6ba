// Fuzzy simplicial set construction on gfx950.
//
// The reference builds the graph with a global sort of (min, max) pairs
// (symmetrize_fuzzy_graph, rust_umap/src/lib.rs:3597-3649).  Here the union is
// a CSR transpose instead: every directed kNN edge i->j is scattered into j's
// incoming list, then one thread per row merges its (index-sorted) outgoing and
// incoming lists.  The result is the same head-sorted, tail-sorted edge list
// with w = mix*(a + b - ab) + (1 - mix)*ab, without any global sort.
#include "graph.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstring>

namespace umapx {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// Running maximum of non-negative weights (as bits, whose unsigned order is
// the float order).  The maximum is reached early (weights run to 1), so a
// read of the running value spares nearly every row its atomic on one
// address; a stale read only costs an atomic.
__device__ __forceinline__ void update_max(uint32_t* bits, float v) {
  const uint32_t u = __float_as_uint(v);
  if (u > __hip_atomic_load(bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(bits, u);
}

__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* lds, int32_t* total) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  int32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane == 63) lds[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t run = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) {
      const int32_t t = lds[w];
      lds[w] = run;
      run += t;
    }
    lds[kScanThreads / 64] = run;
  }
  __syncthreads();
  const int32_t ex = inc - v + lds[wave];
  *total = lds[kScanThreads / 64];
  return ex;
}

__global__ __launch_bounds__(kScanThreads) void scan_tiles_kernel(const int32_t* __restrict__ in,
                                                                   int32_t* __restrict__ out,
                                                                   int64_t n,
                                                                   int32_t* __restrict__ tile_sums) {
  __shared__ int32_t lds[kScanThreads / 64 + 1];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    s += v[i];
  }
  int32_t total;
  int32_t ex = block_exclusive_scan(s, lds, &total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = ex;
    ex += v[i];
  }
  if (threadIdx.x == 0 && tile_sums) tile_sums[blockIdx.x] = total;
}

__global__ void scan_add_kernel(int32_t* __restrict__ out, int64_t n,
                                const int32_t* __restrict__ tile_off) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_off[i / kScanTile];
}

__global__ void write_total_kernel(const int32_t* in, const int32_t* ex, int64_t n, int32_t* total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *total = (n > 0) ? ex[n - 1] + in[n - 1] : 0;
}

// ------------------------------------------------------------ smooth_knn ---
__global__ void include_self_kernel(const int32_t* __restrict__ idx, const float* __restrict__ dist,
                                    int64_t n, int k, int32_t* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (idx[i * k] != i || !(fabsf(dist[i * k]) <= 1e-5f)) atomicAnd(flag, 0);
}

__global__ void sum_nonself_kernel(const float* __restrict__ dist, int64_t n, int k, int start,
                                   double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    for (int c = start; c < k; ++c) s += static_cast<double>(dist[i * k + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// One thread per row; same iteration structure as lib.rs:3455-3519.
template <bool FP64>
__global__ void smooth_knn_kernel(const float* __restrict__ dist, int64_t n, int k, float target,
                                  float local_connectivity, int start,
                                  const double* __restrict__ total, int64_t total_count,
                                  float* __restrict__ sigmas, float* __restrict__ rhos) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* act = dist + i * k + start;
  const int na = k - start;
  // mean of this row (sequential f32, as the reference)
  float mean_ith = 0.f;
  if (na > 0) {
    float s = 0.f;
    for (int c = 0; c < na; ++c) s = fadd(s, act[c]);
    mean_ith = fdiv(s, static_cast<float>(na));
  }
  // rho from the non-zero distances
  int nnz = 0;
  float mx = 0.f;
  for (int c = 0; c < na; ++c)
    if (act[c] > 0.f) {
      mx = nnz == 0 ? act[c] : fmaxf(mx, act[c]);
      ++nnz;
    }
  auto nth_nonzero = [&](int want) {
    int seen = 0;
    for (int c = 0; c < na; ++c)
      if (act[c] > 0.f) {
        if (seen == want) return act[c];
        ++seen;
      }
    return 0.f;
  };
  float rho = 0.f;
  if (static_cast<float>(nnz) >= local_connectivity) {
    const int index = static_cast<int>(floorf(local_connectivity));
    const float interp = fsub(local_connectivity, static_cast<float>(index));
    if (index > 0) {
      rho = nth_nonzero(index - 1);
      if (interp > 1e-5f && index < nnz)
        rho = fadd(rho, fmul(interp, fsub(nth_nonzero(index), nth_nonzero(index - 1))));
    } else if (nnz > 0) {
      rho = fmul(interp, nth_nonzero(0));
    }
  } else if (nnz > 0) {
    rho = mx;
  }
  float lo = 0.f, hi = INFINITY, mid = 1.f;
  for (int it = 0; it < 64; ++it) {
    bool stop, above;
    if constexpr (FP64) {
      double psum = 0.0;
      for (int c = 0; c < na; ++c) {
        const float dd = fsub(act[c], rho);
        psum += dd > 0.f ? exp(-static_cast<double>(fdiv(dd, mid))) : 1.0;
      }
      stop = fabs(psum - static_cast<double>(target)) < 1e-5;
      above = psum > static_cast<double>(target);
    } else {
      float psum = 0.f;
      for (int c = 0; c < na; ++c) {
        const float dd = fsub(act[c], rho);
        psum = fadd(psum, dd > 0.f ? expf(-fdiv(dd, mid)) : 1.f);
      }
      stop = fabsf(fsub(psum, target)) < 1e-5f;
      above = psum > target;
    }
    if (stop) break;
    if (above) {
      hi = mid;
      mid = fdiv(fadd(lo, hi), 2.f);
    } else {
      lo = mid;
      if (isinf(hi))
        mid = fmul(mid, 2.f);
      else
        mid = fdiv(fadd(lo, hi), 2.f);
    }
  }
  const float mean_all =
      total_count > 0 ? static_cast<float>(*total / static_cast<double>(total_count)) : 0.f;
  const float min_scale = rho > 0.f ? fmul(1e-3f, mean_ith) : fmul(1e-3f, mean_all);
  sigmas[i] = fmaxf(mid, min_scale);
  rhos[i] = rho;
}

// ----------------------------------------------------------- membership ---
// lib.rs:3525-3560 (bipartite=0) and lib.rs:3562-3595 (bipartite=1).
__global__ void membership_kernel(const int32_t* __restrict__ idx, const float* __restrict__ dist,
                                  const float* __restrict__ sig, const float* __restrict__ rho,
                                  int64_t n, int k, int bipartite, float* __restrict__ wdir,
                                  int32_t* __restrict__ indeg) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k;
  const int32_t nb = idx[e];
  const float d = dist[e];
  float val;
  if (!bipartite && nb == i) {
    val = 0.f;
  } else if (fsub(d, rho[i]) <= 0.f || sig[i] == 0.f) {
    val = 1.f;
  } else {
    val = expf(fdiv(-fsub(d, rho[i]), sig[i]));
  }
  wdir[e] = val;
  if (indeg && val > 0.f && nb != i) atomicAdd(&indeg[nb], 1);
}

__global__ void scatter_incoming_kernel(const int32_t* __restrict__ idx,
                                        const float* __restrict__ wdir, int64_t n, int k,
                                        const int32_t* __restrict__ in_off,
                                        int32_t* __restrict__ fill, int32_t* __restrict__ in_src,
                                        float* __restrict__ in_w) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const int32_t i = static_cast<int32_t>(e / k);
  const float w = wdir[e];
  const int32_t j = idx[e];
  if (!(w > 0.f) || j == i) return;
  const int32_t p = in_off[j] + atomicAdd(&fill[j], 1);
  in_src[p] = i;
  in_w[p] = w;
}

// Serial merge of one row (used for rows whose outgoing + incoming lists do
// not fit in one wavefront): sort both lists by neighbour index in place and
// merge them with the arithmetic of lib.rs:3611-3646.
__device__ void union_row_serial(const int32_t* __restrict__ idx, const float* __restrict__ wdir,
                                 int64_t i, int k, int32_t b0, int32_t b1,
                                 int32_t* __restrict__ in_src, float* __restrict__ in_w,
                                 float mix, int32_t* __restrict__ tmp_col,
                                 float* __restrict__ tmp_w, int32_t* __restrict__ ucount,
                                 uint32_t* __restrict__ wmax_bits, int32_t* __restrict__ out_c,
                                 float* __restrict__ out_w) {
  // outgoing entries go to the scratch area out_c/out_w (k slots per row)
  int no = 0;
  for (int c = 0; c < k; ++c) {
    const float w = wdir[i * k + c];
    const int32_t j = idx[i * k + c];
    if (!(w > 0.f) || j == i) continue;
    int p = no++;
    while (p > 0 && out_c[p - 1] > j) {
      out_c[p] = out_c[p - 1];
      out_w[p] = out_w[p - 1];
      --p;
    }
    out_c[p] = j;
    out_w[p] = w;
  }
  for (int32_t a = b0 + 1; a < b1; ++a) {
    const int32_t sv = in_src[a];
    const float wv = in_w[a];
    int32_t p = a;
    while (p > b0 && in_src[p - 1] > sv) {
      in_src[p] = in_src[p - 1];
      in_w[p] = in_w[p - 1];
      --p;
    }
    in_src[p] = sv;
    in_w[p] = wv;
  }
  const int64_t base = i * k + b0;
  int cnt = 0, po = 0;
  int32_t pi = b0;
  float local_max = 0.f;
  while (po < no || pi < b1) {
    int32_t j;
    float wij = 0.f, wji = 0.f;
    if (pi >= b1 || (po < no && out_c[po] < in_src[pi])) {
      j = out_c[po];
      wij = out_w[po++];
    } else if (po >= no || in_src[pi] < out_c[po]) {
      j = in_src[pi];
      wji = in_w[pi++];
    } else {
      j = out_c[po];
      wij = out_w[po++];
      wji = in_w[pi++];
    }
    const float prod = fmul(wij, wji);
    const float sym = fadd(fmul(mix, fsub(fadd(wij, wji), prod)), fmul(fsub(1.f, mix), prod));
    if (sym > 0.f) {
      tmp_col[base + cnt] = j;
      tmp_w[base + cnt] = sym;
      ++cnt;
      local_max = fmaxf(local_max, sym);
    }
  }
  ucount[i] = cnt;
  if (cnt > 0) update_max(wmax_bits, local_max);
}

// One wavefront per row: the outgoing (k) and incoming entries of the row are
// loaded one per lane with key 2*col + (incoming), sorted by a register
// bitonic network, and adjacent equal columns are combined with the
// arithmetic of lib.rs:3611-3646; surviving entries are compacted in column
// order.  Rows with more than 64 entries are listed for union_long_kernel
// (up to kUnionBig entries) or take union_row_serial.
constexpr int kUnionMid = 1024;   // long rows handled with a small LDS footprint
constexpr int kUnionBig = 16384;  // longest row sorted in LDS
__global__ __launch_bounds__(256) void union_rows_kernel(
    const int32_t* __restrict__ idx, const float* __restrict__ wdir, int64_t n, int k,
    const int32_t* __restrict__ in_off, int32_t* __restrict__ in_src, float* __restrict__ in_w,
    float mix, int32_t* __restrict__ tmp_col, float* __restrict__ tmp_w,
    int32_t* __restrict__ ucount, uint32_t* __restrict__ wmax_bits,
    int32_t* __restrict__ serial_c, float* __restrict__ serial_w, int32_t* __restrict__ mid_list,
    int32_t* __restrict__ big_list, int32_t* __restrict__ long_count) {
  const int lane = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n) return;
  const int32_t b0 = in_off[i], b1 = in_off[i + 1];
  const int m_in = b1 - b0;
  if (k + m_in > 64) {
    if (lane == 0) {
      if (k + m_in <= kUnionMid)
        mid_list[atomicAdd(&long_count[0], 1)] = static_cast<int32_t>(i);
      else if (k + m_in <= kUnionBig)
        big_list[atomicAdd(&long_count[1], 1)] = static_cast<int32_t>(i);
      else
        union_row_serial(idx, wdir, i, k, b0, b1, in_src, in_w, mix, tmp_col, tmp_w, ucount,
                         wmax_bits, serial_c + i * k, serial_w + i * k);
    }
    return;
  }
  int key = INT_MAX;
  float w = 0.f;
  if (lane < k) {
    const float wv = wdir[i * k + lane];
    const int32_t j = idx[i * k + lane];
    if (wv > 0.f && j != i) {
      key = 2 * j;
      w = wv;
    }
  } else if (lane < k + m_in) {
    key = 2 * in_src[b0 + lane - k] + 1;
    w = in_w[b0 + lane - k];
  }
  // bitonic sort ascending by key (keys are unique except INT_MAX padding)
#pragma unroll
  for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int ok = __shfl_xor(key, stride);
      const float ow = __shfl_xor(w, stride);
      const bool up = (lane & size) == 0;
      const bool lower = (lane & stride) == 0;
      const bool take = (lower == up) ? (ok < key) : (ok > key);
      if (take) {
        key = ok;
        w = ow;
      }
    }
  }
  const int col = key >> 1;
  const int next_key = __shfl_down(key, 1);
  const float next_w = __shfl_down(w, 1);
  const int prev_key = __shfl_up(key, 1);
  const bool valid = key != INT_MAX;
  const bool first = valid && (lane == 0 || (prev_key >> 1) != col || prev_key == INT_MAX);
  float wij = 0.f, wji = 0.f;
  if ((key & 1) == 0) {
    wij = w;
    if (lane < 63 && next_key != INT_MAX && (next_key >> 1) == col) wji = next_w;
  } else {
    wji = w;
  }
  const float prod = fmul(wij, wji);
  const float sym = fadd(fmul(mix, fsub(fadd(wij, wji), prod)), fmul(fsub(1.f, mix), prod));
  const bool keep = first && sym > 0.f;
  const unsigned long long bal = __ballot(keep);
  const int pos = __popcll(bal & ((1ULL << lane) - 1ULL));
  const int64_t base = i * k + b0;
  if (keep) {
    tmp_col[base + pos] = col;
    tmp_w[base + pos] = sym;
  }
  float mx = keep ? sym : 0.f;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  if (lane == 0) {
    const int cnt = __popcll(bal);
    ucount[i] = cnt;
    if (cnt > 0) update_max(wmax_bits, mx);
  }
}

// Block-per-row union for rows with more than 64 entries: the same keys as
// union_rows_kernel sorted by a bitonic network in LDS over the next power of
// two of the row's length (at most cap), equal columns combined, survivors
// compacted by a block prefix sum.  Dynamic LDS: cap * 8 bytes + 1 KiB.
__global__ __launch_bounds__(256) void union_long_kernel(
    const int32_t* __restrict__ list, int cap, const int32_t* __restrict__ idx,
    const float* __restrict__ wdir, int k, const int32_t* __restrict__ in_off,
    const int32_t* __restrict__ in_src, const float* __restrict__ in_w, float mix,
    int32_t* __restrict__ tmp_col, float* __restrict__ tmp_w, int32_t* __restrict__ ucount,
    uint32_t* __restrict__ wmax_bits) {
  extern __shared__ __align__(16) unsigned char lds_u[];
  int* skey = reinterpret_cast<int*>(lds_u);
  float* sw = reinterpret_cast<float*>(lds_u + 4 * cap);
  int* part = reinterpret_cast<int*>(lds_u + 8 * cap);  // 256 partial counts
  const int tid = threadIdx.x;
  const int64_t i = list[blockIdx.x];
  const int32_t b0 = in_off[i], b1 = in_off[i + 1];
  const int len = k + (b1 - b0);
  int lp = 64;
  while (lp < len) lp <<= 1;
  for (int q = tid; q < lp; q += 256) {
    int key = INT_MAX;
    float w = 0.f;
    if (q < k) {
      const float wv = wdir[i * k + q];
      const int32_t j = idx[i * k + q];
      if (wv > 0.f && j != i) {
        key = 2 * j;
        w = wv;
      }
    } else if (q < len) {
      key = 2 * in_src[b0 + q - k] + 1;
      w = in_w[b0 + q - k];
    }
    skey[q] = key;
    sw[q] = w;
  }
  __syncthreads();
  for (int size = 2; size <= lp; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = tid; t < lp / 2; t += 256) {
        const int lo = 2 * t - (t & (stride - 1));
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const int kl = skey[lo], kh = skey[hi];
        if ((kl > kh) == up) {
          skey[lo] = kh;
          skey[hi] = kl;
          const float wt = sw[lo];
          sw[lo] = sw[hi];
          sw[hi] = wt;
        }
      }
      __syncthreads();
    }
  }
  // survivors: first entry of each column with a positive symmetric weight;
  // thread tid scans the contiguous run [tid * per, (tid + 1) * per)
  const int per = lp / 256 > 0 ? lp / 256 : 1;
  const int q0 = tid * per;
  auto sym_at = [&](int q, float& out) {
    const int key = skey[q];
    if (key == INT_MAX) return false;
    const int col = key >> 1;
    if (q > 0 && skey[q - 1] != INT_MAX && (skey[q - 1] >> 1) == col) return false;
    float wij = 0.f, wji = 0.f;
    if ((key & 1) == 0) {
      wij = sw[q];
      if (q + 1 < lp && skey[q + 1] != INT_MAX && (skey[q + 1] >> 1) == col) wji = sw[q + 1];
    } else {
      wji = sw[q];
    }
    const float prod = fmul(wij, wji);
    out = fadd(fmul(mix, fsub(fadd(wij, wji), prod)), fmul(fsub(1.f, mix), prod));
    return out > 0.f;
  };
  int mine = 0;
  float mx = 0.f;
  for (int q = q0; q < q0 + per && q < lp; ++q) {
    float v;
    if (sym_at(q, v)) {
      ++mine;
      mx = fmaxf(mx, v);
    }
  }
  part[tid] = mine;
  __syncthreads();
  // exclusive scan of the 256 partial counts (thread 0; 256 adds)
  __shared__ int total;
  if (tid == 0) {
    int run = 0;
    for (int t = 0; t < 256; ++t) {
      const int c = part[t];
      part[t] = run;
      run += c;
    }
    total = run;
  }
  __syncthreads();
  int pos = part[tid];
  const int64_t base = i * k + b0;
  for (int q = q0; q < q0 + per && q < lp; ++q) {
    float v;
    if (sym_at(q, v)) {
      tmp_col[base + pos] = skey[q] >> 1;
      tmp_w[base + pos] = v;
      ++pos;
    }
  }
  // row maximum
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  __shared__ float wmx[4];
  if ((tid & 63) == 0) wmx[tid >> 6] = mx;
  __syncthreads();
  if (tid == 0) {
    ucount[i] = total;
    const float m4 = fmaxf(fmaxf(wmx[0], wmx[1]), fmaxf(wmx[2], wmx[3]));
    if (total > 0) update_max(wmax_bits, m4);
  }
}

// Rows of a bipartite graph keep their kNN order (lib.rs:3562-3595).
__global__ void bipartite_rows_kernel(const int32_t* __restrict__ idx,
                                      const float* __restrict__ wdir, int64_t m, int k,
                                      int32_t* __restrict__ tmp_col, float* __restrict__ tmp_w,
                                      int32_t* __restrict__ ucount,
                                      uint32_t* __restrict__ wmax_bits) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int cnt = 0;
  float local_max = 0.f;
  for (int c = 0; c < k; ++c) {
    const float w = wdir[i * k + c];
    if (w > 0.f) {
      tmp_col[i * k + cnt] = idx[i * k + c];
      tmp_w[i * k + cnt] = w;
      ++cnt;
      local_max = fmaxf(local_max, w);
    }
  }
  ucount[i] = cnt;
  if (cnt > 0) update_max(wmax_bits, local_max);
}

// prune_edges (lib.rs:3804-3822): keep w >= max/denom and w > 0.
// prune_edges (lib.rs:3804-3822) over the per-row union lists: one wave per
// row, lanes over the row's entries, so reads and writes are coalesced and a
// hub row does not hold one lane for a thousand iterations.  kcount[i] = the
// entries kept; the write pass keeps their order (ballot prefix counts).
constexpr int kPruneWaves = 4;

__device__ __forceinline__ bool prune_keep(float w, float thr, int do_prune) {
  return (!do_prune || w >= thr) && w > 0.f;
}

__global__ __launch_bounds__(64 * kPruneWaves) void prune_count_kernel(
    const int32_t* __restrict__ tmp_off, const float* __restrict__ tmp_w,
    const int32_t* __restrict__ ucount, int64_t n, const uint32_t* __restrict__ wmax_bits,
    float denom, int do_prune, int32_t* __restrict__ kcount) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPruneWaves + (threadIdx.x >> 6);
  if (i >= n) return;  // the whole wave
  const int lane = threadIdx.x & 63;
  const float thr = do_prune ? fdiv(__uint_as_float(*wmax_bits), denom) : 0.f;
  const int32_t base = tmp_off[i], cnt = ucount[i];
  int kept = 0;
  for (int c0 = 0; c0 < cnt; c0 += 64) {
    const int c = c0 + lane;
    const bool keep = c < cnt && prune_keep(tmp_w[base + c], thr, do_prune);
    kept += __popcll(__ballot(keep));
  }
  if (lane == 0) kcount[i] = kept;
}

__global__ __launch_bounds__(64 * kPruneWaves) void prune_write_kernel(
    const int32_t* __restrict__ tmp_off, const int32_t* __restrict__ tmp_col,
    const float* __restrict__ tmp_w, const int32_t* __restrict__ ucount, int64_t n,
    const uint32_t* __restrict__ wmax_bits, float denom, int do_prune,
    const int32_t* __restrict__ row_ptr, int32_t* __restrict__ head, int32_t* __restrict__ tail,
    float* __restrict__ weight) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kPruneWaves + (threadIdx.x >> 6);
  if (i >= n) return;  // the whole wave
  const int lane = threadIdx.x & 63;
  const float thr = do_prune ? fdiv(__uint_as_float(*wmax_bits), denom) : 0.f;
  const int32_t base = tmp_off[i], cnt = ucount[i];
  int32_t o = row_ptr[i];
  const uint64_t below = (uint64_t{1} << lane) - 1;
  for (int c0 = 0; c0 < cnt; c0 += 64) {
    const int c = c0 + lane;
    const float w = c < cnt ? tmp_w[base + c] : 0.f;
    const bool keep = c < cnt && prune_keep(w, thr, do_prune);
    const uint64_t mask = __ballot(keep);
    if (keep) {
      const int32_t at = o + __popcll(mask & below);
      head[at] = static_cast<int32_t>(i);
      tail[at] = tmp_col[base + c];
      weight[at] = w;
    }
    o += __popcll(mask);
  }
}

__global__ void tmp_offsets_kernel(const int32_t* __restrict__ in_off, int64_t n, int k,
                                   int32_t* __restrict__ tmp_off) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  tmp_off[i] = static_cast<int32_t>(i * k + (in_off ? in_off[i] : 0));
}

// lib.rs:4937-4960: eps = T / (T * (w / max)).
__global__ void eps_kernel(const float* __restrict__ w, int64_t ne, float wmax, float ef,
                           float* __restrict__ eps) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const float ns = fmul(ef, fdiv(w[e], wmax));
  eps[e] = ns > 0.f ? fdiv(ef, ns) : INFINITY;
}

__global__ void local_radii_kernel(const int32_t* __restrict__ idx, const float* __restrict__ dist,
                                   int64_t n, int k, float vs, float* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  int c = 0;
  for (int j = 0; j < k; ++j) {
    const int32_t nb = idx[i * k + j];
    const float d = dist[i * k + j];
    if (nb == i && fabsf(d) <= 1e-5f) continue;
    s = fadd(s, logf(fadd(fmaxf(d, 0.f), vs)));
    ++c;
  }
  out[i] = c == 0 ? logf(vs) : fdiv(s, static_cast<float>(c));
}

__global__ void embedding_radii_kernel(const float* __restrict__ y, int64_t n, int dim,
                                       const int32_t* __restrict__ idx, int k, float vs,
                                       float* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  int c = 0;
  for (int j = 0; j < k; ++j) {
    const int32_t nb = idx[i * k + j];
    if (nb == i) continue;
    float sq = 0.f;
    for (int t = 0; t < dim; ++t) {
      const float df = fsub(y[i * dim + t], y[static_cast<int64_t>(nb) * dim + t]);
      sq = fadd(sq, fmul(df, df));
    }
    s = fadd(s, logf(fadd(fmaxf(fsqrt(sq), 0.f), vs)));
    ++c;
  }
  out[i] = c == 0 ? logf(vs) : fdiv(s, static_cast<float>(c));
}

}  // namespace

int64_t exclusive_scan(const int32_t* in, int32_t* out, int64_t n, ScanWorkspace& ws,
                       hipStream_t s) {
  ws.total.ensure(1);
  if (n == 0) return 0;
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(1), dim3(kScanThreads), 0, s, in, out, n,
                       ws.total.get());
    KERNEL_CHECK();
  } else {
    DevBuf<int32_t> sums(static_cast<size_t>(tiles)), offs(static_cast<size_t>(tiles));
    hipLaunchKernelGGL(scan_tiles_kernel, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0,
                       s, in, out, n, sums.get());
    KERNEL_CHECK();
    ScanWorkspace inner;
    exclusive_scan(sums.get(), offs.get(), tiles, inner, s);
    hipLaunchKernelGGL(scan_add_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, out, n,
                       offs.get());
    KERNEL_CHECK();
    hipLaunchKernelGGL(write_total_kernel, dim3(1), dim3(64), 0, s, in, out, n, ws.total.get());
    KERNEL_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));
  }
  int32_t total = 0;
  HIP_CHECK(hipMemcpyAsync(&total, ws.total.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return total;
}

bool knn_include_self_device(const int32_t* idx, const float* dist, int64_t n, int k,
                             GraphWorkspace& ws, hipStream_t s) {
  ws.flag.ensure(1);
  const int32_t one = 1;
  HIP_CHECK(hipMemcpyAsync(ws.flag.get(), &one, sizeof(one), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(include_self_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, idx, dist, n,
                     k, ws.flag.get());
  KERNEL_CHECK();
  int32_t flag = 0;
  HIP_CHECK(hipMemcpyAsync(&flag, ws.flag.get(), sizeof(flag), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return flag != 0 && k > 0;
}

void smooth_knn_dist_device(const float* dist, int64_t n, int k, float k_value,
                            float local_connectivity, float bandwidth, bool include_self,
                            bool fp64_sums, float* sigmas, float* rhos, GraphWorkspace& ws,
                            hipStream_t s) {
  maybe_inject_fault("smooth_knn_dist");
  if (n == 0) return;
  const int start = (include_self && k > 0) ? 1 : 0;
  ws.dsum.ensure(1);
  HIP_CHECK(hipMemsetAsync(ws.dsum.get(), 0, sizeof(double), s));
  hipLaunchKernelGGL(sum_nonself_kernel, dim3(std::min<int64_t>(blocks_for(n, 256), 512)),
                     dim3(256), 0, s, dist, n, k, start, ws.dsum.get());
  KERNEL_CHECK();
  const float target = log2f(k_value) * bandwidth;
  const int64_t count = n * static_cast<int64_t>(k - start);
  if (fp64_sums)
    hipLaunchKernelGGL(smooth_knn_kernel<true>, dim3(blocks_for(n, 128)), dim3(128), 0, s, dist, n,
                       k, target, local_connectivity, start, ws.dsum.get(), count, sigmas, rhos);
  else
    hipLaunchKernelGGL(smooth_knn_kernel<false>, dim3(blocks_for(n, 128)), dim3(128), 0, s, dist,
                       n, k, target, local_connectivity, start, ws.dsum.get(), count, sigmas, rhos);
  KERNEL_CHECK();
}

static void finish_graph(int64_t n, int k, int prune_epochs, const int32_t* tmp_off, Graph& g,
                         GraphWorkspace& ws, hipStream_t s) {
  g.symmetric = false;
  const float denom = prune_epochs > 10 ? static_cast<float>(prune_epochs) : 500.0f;
  const int do_prune = prune_epochs > 0 ? 1 : 0;
  ws.kcount.ensure(static_cast<size_t>(n));
  hipLaunchKernelGGL(prune_count_kernel, dim3(blocks_for(n, kPruneWaves)), dim3(64 * kPruneWaves), 0, s, tmp_off,
                     ws.tmp_w.get(), ws.ucount.get(), n, ws.umax.get(), denom, do_prune,
                     ws.kcount.get());
  KERNEL_CHECK();
  g.n = n;
  g.row_ptr.ensure(static_cast<size_t>(n) + 1);
  const int64_t ne = exclusive_scan(ws.kcount.get(), g.row_ptr.get(), n, ws.scan, s);
  const int32_t ne32 = static_cast<int32_t>(ne);
  HIP_CHECK(hipMemcpyAsync(g.row_ptr.get() + n, &ne32, sizeof(int32_t), hipMemcpyHostToDevice, s));
  g.n_edges = ne;
  g.head.ensure(std::max<int64_t>(ne, 1));
  g.tail.ensure(std::max<int64_t>(ne, 1));
  g.weight.ensure(std::max<int64_t>(ne, 1));
  hipLaunchKernelGGL(prune_write_kernel, dim3(blocks_for(n, kPruneWaves)), dim3(64 * kPruneWaves), 0, s, tmp_off,
                     ws.tmp_col.get(), ws.tmp_w.get(), ws.ucount.get(), n, ws.umax.get(), denom,
                     do_prune, g.row_ptr.get(), g.head.get(), g.tail.get(), g.weight.get());
  KERNEL_CHECK();
  uint32_t bits = 0;
  HIP_CHECK(hipMemcpyAsync(&bits, ws.umax.get(), sizeof(bits), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float mx;
  std::memcpy(&mx, &bits, 4);
  g.max_weight = mx;
  (void)k;
}

void fuzzy_graph_device(const int32_t* idx, const float* dist, const float* sigmas,
                        const float* rhos, int64_t n, int k, float mix, int prune_epochs, Graph& g,
                        GraphWorkspace& ws, hipStream_t s) {
  maybe_inject_fault("fuzzy_graph");
  if (static_cast<double>(n) * k * 2.0 >= 2147483647.0)
    throw_invalid("graph too large for 32-bit edge offsets");
  const int64_t nk = n * k;
  ws.wdir.ensure(nk);
  ws.indeg.ensure(n);
  ws.fill.ensure(n);
  ws.in_off.ensure(n + 1);
  HIP_CHECK(hipMemsetAsync(ws.indeg.get(), 0, sizeof(int32_t) * n, s));
  HIP_CHECK(hipMemsetAsync(ws.fill.get(), 0, sizeof(int32_t) * n, s));
  hipLaunchKernelGGL(membership_kernel, dim3(blocks_for(nk, 256)), dim3(256), 0, s, idx, dist,
                     sigmas, rhos, n, k, 0, ws.wdir.get(), ws.indeg.get());
  KERNEL_CHECK();
  const int64_t n_in = exclusive_scan(ws.indeg.get(), ws.in_off.get(), n, ws.scan, s);
  const int32_t nin32 = static_cast<int32_t>(n_in);
  HIP_CHECK(hipMemcpyAsync(ws.in_off.get() + n, &nin32, sizeof(int32_t), hipMemcpyHostToDevice, s));
  ws.in_src.ensure(std::max<int64_t>(n_in, 1));
  ws.in_w.ensure(std::max<int64_t>(n_in, 1));
  hipLaunchKernelGGL(scatter_incoming_kernel, dim3(blocks_for(nk, 256)), dim3(256), 0, s, idx,
                     ws.wdir.get(), n, k, ws.in_off.get(), ws.fill.get(), ws.in_src.get(),
                     ws.in_w.get());
  KERNEL_CHECK();
  const int64_t cap = nk + n_in;
  ws.tmp_col.ensure(std::max<int64_t>(cap, 1));
  ws.tmp_w.ensure(std::max<int64_t>(cap, 1));
  ws.ucount.ensure(n);
  ws.tmp_off.ensure(n);
  ws.umax.ensure(1);
  HIP_CHECK(hipMemsetAsync(ws.umax.get(), 0, sizeof(uint32_t), s));
  if (k > 64) throw_invalid("n_neighbors must be <= 64 on the GPU path");
  ws.serial_c.ensure(nk);
  ws.serial_w.ensure(nk);
  ws.long_rows.ensure(2 * n + 2);
  int32_t* mid_list = ws.long_rows.get();
  int32_t* big_list = ws.long_rows.get() + n;
  int32_t* long_count = ws.long_rows.get() + 2 * n;
  HIP_CHECK(hipMemsetAsync(long_count, 0, 8, s));
  hipLaunchKernelGGL(union_rows_kernel, dim3(blocks_for(n, 4)), dim3(256), 0, s, idx,
                     ws.wdir.get(), n, k, ws.in_off.get(), ws.in_src.get(), ws.in_w.get(), mix,
                     ws.tmp_col.get(), ws.tmp_w.get(), ws.ucount.get(), ws.umax.get(),
                     ws.serial_c.get(), ws.serial_w.get(), mid_list, big_list, long_count);
  KERNEL_CHECK();
  int32_t n_long[2] = {0, 0};
  HIP_CHECK(hipMemcpyAsync(n_long, long_count, 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  auto long_rows = [&](const int32_t* list, int count, int cap) {
    if (count == 0) return;
    const size_t bytes = static_cast<size_t>(cap) * 8 + 1024;
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(union_long_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(bytes)));
    hipLaunchKernelGGL(union_long_kernel, dim3(count), dim3(256), bytes, s, list, cap, idx,
                       ws.wdir.get(), k, ws.in_off.get(), ws.in_src.get(), ws.in_w.get(), mix,
                       ws.tmp_col.get(), ws.tmp_w.get(), ws.ucount.get(), ws.umax.get());
    KERNEL_CHECK();
  };
  long_rows(mid_list, n_long[0], kUnionMid);
  long_rows(big_list, n_long[1], kUnionBig);
  hipLaunchKernelGGL(tmp_offsets_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                     ws.in_off.get(), n, k, ws.tmp_off.get());
  KERNEL_CHECK();
  finish_graph(n, k, prune_epochs, ws.tmp_off.get(), g, ws, s);
  g.symmetric = true;  // the union holds every edge in both directions
}

void bipartite_graph_device(const int32_t* idx, const float* dist, const float* sigmas,
                            const float* rhos, int64_t m, int k, int prune_epochs, Graph& g,
                            GraphWorkspace& ws, hipStream_t s) {
  maybe_inject_fault("fuzzy_graph");
  const int64_t mk = m * k;
  ws.wdir.ensure(mk);
  hipLaunchKernelGGL(membership_kernel, dim3(blocks_for(mk, 256)), dim3(256), 0, s, idx, dist,
                     sigmas, rhos, m, k, 1, ws.wdir.get(), static_cast<int32_t*>(nullptr));
  KERNEL_CHECK();
  ws.tmp_col.ensure(mk);
  ws.tmp_w.ensure(mk);
  ws.ucount.ensure(m);
  ws.tmp_off.ensure(m);
  ws.umax.ensure(1);
  HIP_CHECK(hipMemsetAsync(ws.umax.get(), 0, sizeof(uint32_t), s));
  hipLaunchKernelGGL(bipartite_rows_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, s, idx,
                     ws.wdir.get(), m, k, ws.tmp_col.get(), ws.tmp_w.get(), ws.ucount.get(),
                     ws.umax.get());
  KERNEL_CHECK();
  hipLaunchKernelGGL(tmp_offsets_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, s,
                     static_cast<const int32_t*>(nullptr), m, k, ws.tmp_off.get());
  KERNEL_CHECK();
  finish_graph(m, k, prune_epochs, ws.tmp_off.get(), g, ws, s);
}

void epochs_per_sample_device(const float* w, int64_t ne, float max_weight, int n_epochs,
                              float* eps, hipStream_t s) {
  if (ne == 0) return;
  const float wmax = std::max(max_weight, 1e-12f);
  hipLaunchKernelGGL(eps_kernel, dim3(blocks_for(ne, 256)), dim3(256), 0, s, w, ne, wmax,
                     static_cast<float>(n_epochs), eps);
  KERNEL_CHECK();
}

void local_radii_device(const int32_t* idx, const float* dist, int64_t n, int k, float vs,
                        float* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(local_radii_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, idx, dist, n,
                     k, vs, out);
  KERNEL_CHECK();
}

void embedding_radii_device(const float* y, int64_t n, int dim, const int32_t* idx, int k,
                            float vs, float* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(embedding_radii_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, y, n, dim,
                     idx, k, vs, out);
  KERNEL_CHECK();
}

}  // namespace umapx
