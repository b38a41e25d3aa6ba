// Exact k-nearest neighbours on gfx950.
//
// Replaces exact_nearest_neighbors(_euclidean/_cosine) and
// exact_nearest_neighbors_to_reference (rust_umap/src/lib.rs:2863-2928,
// 3283-3399).  Two passes:
//   1. candidates: ||y||^2 - 2 x.y score tiles on MFMA (FP32 32x32x2, or BF16
//      32x32x16 with FP32 accumulation) over mean-centred rows; every wave
//      owns 32 query rows and keeps a sorted top-KC (score, index) list per
//      row in LDS, fed by a register-level threshold filter;
//   2. re-rank: the KC candidates of each row get the reference's exact f32
//      distance (direct difference, 8-lane partial sums for d >= 64, summed in
//      lane order, then the tail; cosine per lib.rs:2740-2765), are sorted by
//      (distance total order, index) and the first k are emitted.
// With the true top-k inside the candidates, indices and distances are bit
// identical to the reference's CPU path.
#include "knn.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

namespace umapx {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kRowsPerWave = 32; // query rows owned by a wave
constexpr int kWaves = 4;        // waves per block

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// ---------------------------------------------------------------- prep ----
// Column sums in f64 (for mean-centring).  grid-stride over rows.
__global__ void column_sums_kernel(const float* __restrict__ x, int64_t n, int d,
                                   int normalize, double* __restrict__ colsum) {
  // one thread per feature column within a block-row chunk
  const int f = threadIdx.x;
  if (f >= d) return;
  double s = 0.0;
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    float v = x[r * d + f];
    s += static_cast<double>(v);
  }
  (void)normalize;
  atomicAdd(&colsum[f], s);
}

// Column sums of the L2-normalised rows (cosine).  One wave per row.
__global__ void column_sums_normalized_kernel(const float* __restrict__ x, int64_t n,
                                              int d, double* __restrict__ colsum) {
  // blocks of 256 = 4 waves; each wave walks rows with a grid stride and
  // accumulates per-lane partial column sums in registers is not possible for
  // arbitrary d, so accumulate in LDS per block then flush.
  extern __shared__ double lds_cols[];
  for (int f = threadIdx.x; f < d; f += blockDim.x) lds_cols[f] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * nw + wave; r < n;
       r += static_cast<int64_t>(gridDim.x) * nw) {
    float ss = 0.f;
    for (int f = lane; f < d; f += 64) {
      const float v = x[r * d + f];
      ss += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
    const float inv = ss > 0.f ? 1.0f / sqrtf(ss) : 0.f;
    for (int f = lane; f < d; f += 64) atomicAdd(&lds_cols[f], static_cast<double>(x[r * d + f] * inv));
  }
  __syncthreads();
  for (int f = threadIdx.x; f < d; f += blockDim.x) atomicAdd(&colsum[f], lds_cols[f]);
}

// Writes the padded, centred (and, for cosine, normalised) compute copy of
// each row plus its squared norm in the compute precision.  One wave per row.
template <typename T>
__global__ void prep_rows_kernel(const float* __restrict__ x, int64_t n, int d, int dpad,
                                 int normalize, const double* __restrict__ colsum,
                                 double inv_n, T* __restrict__ out,
                                 float* __restrict__ sqnorm) {
  const int lane = threadIdx.x & 63;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= n) return;
  float inv = 1.0f;
  if (normalize) {
    float ss = 0.f;
    for (int f = lane; f < d; f += 64) {
      const float v = x[r * d + f];
      ss += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
    inv = ss > 0.f ? 1.0f / sqrtf(ss) : 0.f;
  }
  float nrm = 0.f;
  for (int f = lane; f < dpad; f += 64) {
    float v = 0.f;
    if (f < d) v = x[r * d + f] * inv - static_cast<float>(colsum[f] * inv_n);
    if constexpr (sizeof(T) == 2) {
      const uint16_t h = f32_to_bf16_rne(v);
      out[r * dpad + f] = h;
      const float vr = bf16_to_f32(h);
      nrm += vr * vr;
    } else {
      out[r * dpad + f] = v;
      nrm += v * v;
    }
  }
  for (int o = 32; o > 0; o >>= 1) nrm += __shfl_xor(nrm, o);
  if (lane == 0) sqnorm[r] = nrm;
}

// ---------------------------------------------------------- candidates ----
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// One 16-byte fragment vector per lane feeds STEPS MFMA steps.  Fragments
// are passed by value (never through pointers) so they stay in VGPRs.
template <typename T>
struct MfmaOp;

template <>
struct MfmaOp<float> {
  static constexpr int kSteps = 4;  // 4 x (32x32x2) per 16-byte vector
  template <int ST>
  __device__ static inline floatx16 run(u32x4 a, u32x4 b, floatx16 c) {
    const floatx4 av = __builtin_bit_cast(floatx4, a);
    const floatx4 bv = __builtin_bit_cast(floatx4, b);
    return __builtin_amdgcn_mfma_f32_32x32x2f32(av[ST], bv[ST], c, 0, 0, 0);
  }
};

template <>
struct MfmaOp<uint16_t> {
  static constexpr int kSteps = 1;  // 1 x (32x32x16) per 16-byte vector
  template <int ST>
  __device__ static inline floatx16 run(u32x4 a, u32x4 b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  }
};

template <typename T, int ST>
struct MfmaSteps {
  __device__ static inline floatx16 run(u32x4 a, u32x4 b, floatx16 c) {
    c = MfmaSteps<T, ST - 1>::run(a, b, c);
    return MfmaOp<T>::template run<ST - 1>(a, b, c);
  }
};
template <typename T>
struct MfmaSteps<T, 0> {
  __device__ static inline floatx16 run(u32x4, u32x4, floatx16 c) { return c; }
};

// LDS layout (bytes) of one block: two reference tiles (double buffer),
// their norms, and per query row the sorted top-KC list, a pass buffer and
// counters.
template <typename T, int DPAD, int KC, int TILE, int BUF>
struct CandLayout {
  static constexpr int kRowBytes = DPAD * static_cast<int>(sizeof(T)) + 16;
  static constexpr int kTileBytes = TILE * kRowBytes;
  static constexpr int kNormBytes = TILE * 4;
  static constexpr int kRows = kWaves * kRowsPerWave;
  static constexpr int kListBytes = kRows * KC * 8;
  static constexpr int kBufBytes = kRows * BUF * 8;
  static constexpr int kMetaBytes = kRows * 8;
  static constexpr int kTiles = 2 * (kTileBytes + kNormBytes);
  static constexpr int kTotal = kTiles + kListBytes + kBufBytes + kMetaBytes;
};
constexpr int kLdsBudget = 160 * 1024;

// Merge the cnt buffered candidates of one row into its sorted top-KC list
// (wave-wide; the wave owns the row).  Few entries: one insertion each, a
// ballot finds the slot and the tail shifts up one lane.  Otherwise the list
// and the buffer are sorted together by a bitonic network over (score,
// index), two entries per lane.  Returns the row's new threshold.
template <int KC>
__device__ __noinline__ float merge_row(float* __restrict__ ld_row, int32_t* __restrict__ li_row,
                                           const float* __restrict__ bd_row,
                                           const int32_t* __restrict__ bi_row, int cnt, int lane) {
  if (cnt <= 4) {
    float ld = lane < KC ? ld_row[lane] : INFINITY;
    int32_t li = lane < KC ? li_row[lane] : INT_MAX;
    for (int e = 0; e < cnt; ++e) {
      const float cd = bd_row[e];
      const int32_t ci = bi_row[e];
      const bool less = lane < KC && (ld < cd || (ld == cd && li < ci));
      const int pos = __popcll(__ballot(less));
      if (pos < KC) {
        const float ud = __shfl_up(ld, 1);
        const int32_t ui = __shfl_up(li, 1);
        if (lane > pos) {
          ld = ud;
          li = ui;
        }
        if (lane == pos) {
          ld = cd;
          li = ci;
        }
      }
    }
    if (lane < KC) {
      ld_row[lane] = ld;
      li_row[lane] = li;
    }
    return __shfl(ld, KC - 1);
  }
  // 128 slots: element j = 2*lane + q.  List entries occupy j < KC, buffered
  // ones j in [KC, KC + cnt).
  float d[2];
  int32_t id[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int j = 2 * lane + q;
    d[q] = INFINITY;
    id[q] = INT_MAX;
    if (j < KC) {
      d[q] = ld_row[j];
      id[q] = li_row[j];
    } else if (j < KC + cnt) {
      d[q] = bd_row[j - KC];
      id[q] = bi_row[j - KC];
    }
  }
  auto less = [](float ad, int32_t ai, float bd, int32_t bi) {
    return ad < bd || (ad == bd && ai < bi);
  };
#pragma unroll
  for (int size = 2; size <= 128; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride == 1) {
        // partner is the other element of this lane
        const int j = 2 * lane;
        const bool up = (j & size) == 0;
        const bool swap = up ? less(d[1], id[1], d[0], id[0]) : less(d[0], id[0], d[1], id[1]);
        if (swap) {
          const float td = d[0];
          d[0] = d[1];
          d[1] = td;
          const int32_t ti = id[0];
          id[0] = id[1];
          id[1] = ti;
        }
      } else {
        const int ls = stride >> 1;  // lane distance
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int j = 2 * lane + q;
          const float od = __shfl_xor(d[q], ls);
          const int32_t oi = __shfl_xor(id[q], ls);
          const bool up = (j & size) == 0;
          const bool lower = (j & stride) == 0;
          const bool take = (lower == up) ? less(od, oi, d[q], id[q]) : less(d[q], id[q], od, oi);
          if (take) {
            d[q] = od;
            id[q] = oi;
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int j = 2 * lane + q;
    if (j < KC) {
      ld_row[j] = d[q];
      li_row[j] = id[q];
    }
  }
  // element KC - 1 sits in lane (KC - 1) / 2, slot (KC - 1) % 2
  const float last = ((KC - 1) % 2) ? d[1] : d[0];
  return __shfl(last, (KC - 1) / 2);
}

// Each block owns kWaves*kRowsPerWave query rows; each wave owns 32 of them
// and keeps their A fragments in registers for the whole feature range.  The
// feature axis is permuted so a lane's consecutive MFMA steps read
// consecutive 16-byte chunks: at step s, lane half h supplies features
// h*DPAD/2 + s*E .. +E (the dot product is invariant to this permutation).
// A 32x32 accumulator tile gives lane l the scores of rows
// 8*(i/4) + 4*(l/32) + i%4 (i = 0..15) against column l%32, so for a fixed i
// the 32 lanes of a half hold all 32 columns of one row: a ballot places that
// row's passing candidates in its buffer without atomics.  The main loop is
// software pipelined: the MFMAs of one 32-column subtile are issued before
// the threshold filter of the previous subtile runs, so the matrix core works
// while the VALU filters; reference tiles are double buffered in LDS.
template <typename T, int DPAD, int KC, int TILE, int BUF>
__global__ __launch_bounds__(kWaves * 64) void knn_candidates_kernel(
    const T* __restrict__ xq, int64_t m, const T* __restrict__ xr,
    const float* __restrict__ nr, int64_t n, int32_t* __restrict__ cand) {
  using L = CandLayout<T, DPAD, KC, TILE, BUF>;
  constexpr int kRefTile = TILE;
  constexpr int kBuf = BUF;
  constexpr int kThreads = kWaves * 64;
  constexpr int VEC = 16 / static_cast<int>(sizeof(T));  // elements per 16 B
  constexpr int HALF = DPAD / 2;
  constexpr int NV = HALF / VEC;                        // 16-byte vectors per lane
  extern __shared__ __align__(16) unsigned char lds[];
  unsigned char* s_tiles = lds;  // [2][tile + norms]
  float* s_ld = reinterpret_cast<float*>(lds + L::kTiles);
  int32_t* s_li = reinterpret_cast<int32_t*>(s_ld + L::kRows * KC);
  float* s_bd = reinterpret_cast<float*>(lds + L::kTiles + L::kListBytes);
  int32_t* s_bi = reinterpret_cast<int32_t*>(s_bd + L::kRows * kBuf);
  int32_t* s_cnt = reinterpret_cast<int32_t*>(lds + L::kTiles + L::kListBytes + L::kBufBytes);
  float* s_thr = reinterpret_cast<float*>(s_cnt + L::kRows);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lrow = lane & 31;
  const int half = lane >> 5;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * L::kRows + wave * kRowsPerWave;

  float* w_ld = s_ld + wave * kRowsPerWave * KC;
  int32_t* w_li = s_li + wave * kRowsPerWave * KC;
  float* w_bd = s_bd + wave * kRowsPerWave * kBuf;
  int32_t* w_bi = s_bi + wave * kRowsPerWave * kBuf;
  int32_t* w_cnt = s_cnt + wave * kRowsPerWave;
  float* w_thr = s_thr + wave * kRowsPerWave;

  for (int i = lane; i < kRowsPerWave * KC; i += 64) {
    w_ld[i] = INFINITY;
    w_li[i] = INT_MAX;
  }
  if (lane < kRowsPerWave) {
    w_cnt[lane] = 0;
    w_thr[lane] = (q0 + lane < m) ? INFINITY : -INFINITY;  // rows past m take nothing
  }

  u32x4 a[NV];
  {
    const int64_t qr = (q0 + lrow < m) ? (q0 + lrow) : (m - 1);
    const u32x4* src = reinterpret_cast<const u32x4*>(xq + qr * DPAD + half * HALF);
#pragma unroll
    for (int v = 0; v < NV; ++v) a[v] = src[v];
  }

  constexpr int CHUNKS = DPAD * static_cast<int>(sizeof(T)) / 16;
  constexpr int PER_THREAD = (kRefTile * CHUNKS + kThreads - 1) / kThreads;
  const int n_tiles = static_cast<int>((n + kRefTile - 1) / kRefTile);
  u32x4 pre[PER_THREAD];
  float pre_nr = INFINITY;
  auto fetch = [&](int tile) {
    const int64_t base = static_cast<int64_t>(tile) * kRefTile;
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i) {
      const int c = tid + i * kThreads;
      u32x4 v = {0u, 0u, 0u, 0u};
      if (tile < n_tiles && c < kRefTile * CHUNKS && base + c / CHUNKS < n)
        v = reinterpret_cast<const u32x4*>(xr + (base + c / CHUNKS) * DPAD)[c % CHUNKS];
      pre[i] = v;
    }
    if (tid < kRefTile) pre_nr = (tile < n_tiles && base + tid < n) ? nr[base + tid] : INFINITY;
  };
  auto stage = [&](int buf) {
    unsigned char* t = s_tiles + buf * (L::kTileBytes + L::kNormBytes);
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i) {
      const int c = tid + i * kThreads;
      if (c < kRefTile * CHUNKS)
        *reinterpret_cast<u32x4*>(t + (c / CHUNKS) * L::kRowBytes + (c % CHUNKS) * 16) = pre[i];
    }
    if (tid < kRefTile) reinterpret_cast<float*>(t + L::kTileBytes)[tid] = pre_nr;
  };
  fetch(0);
  stage(0);
  fetch(1);
  __syncthreads();

  float thr[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) thr[i] = w_thr[8 * (i >> 2) + 4 * half + (i & 3)];

  auto mma = [&](int buf, int sub) {
    const unsigned char* brow = s_tiles + buf * (L::kTileBytes + L::kNormBytes) +
                                (sub * 32 + lrow) * L::kRowBytes + half * HALF * sizeof(T);
    u32x4 bv[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) bv[v] = *reinterpret_cast<const u32x4*>(brow + v * 16);
    floatx16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
#pragma unroll
    for (int v = 0; v < NV; ++v) acc = MfmaSteps<T, MfmaOp<T>::kSteps>::run(a[v], bv[v], acc);
    return acc;
  };
  // threshold filter of one subtile: passing scores go to their row's buffer
  auto filter = [&](const floatx16& acc, float ncol, int32_t col) {
    bool any = false;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float sc = ncol - 2.0f * acc[i];
      const bool keep = sc < thr[i];
      const unsigned long long kb = __ballot(keep);
      if (kb == 0ULL) continue;
      any = true;
      const int row0 = 8 * (i >> 2) + (i & 3), row1 = row0 + 4;
      const int cnt0 = w_cnt[row0], cnt1 = w_cnt[row1];
      if (keep) {
        const unsigned mine = half ? static_cast<unsigned>(kb >> 32) : static_cast<unsigned>(kb);
        const int row = half ? row1 : row0;
        const int slot = (half ? cnt1 : cnt0) + __popc(mine & ((1u << lrow) - 1u));
        w_bd[row * kBuf + slot] = sc;
        w_bi[row * kBuf + slot] = col;
      }
      if (lane == 0) {
        w_cnt[row0] = cnt0 + __popc(static_cast<unsigned>(kb));
        w_cnt[row1] = cnt1 + __popc(static_cast<unsigned>(kb >> 32));
      }
    }
    return any;
  };
  // merge every row whose buffer could not take another full subtile
  auto flush = [&](int limit) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int my_cnt = lane < kRowsPerWave ? w_cnt[lane] : 0;
    unsigned long long pending = __ballot(my_cnt > limit);
    if (pending == 0ULL) return;
    while (pending) {
      const int row = __builtin_ctzll(pending);
      pending &= pending - 1;
      const int cnt = __shfl(my_cnt, row);
      const float nt = merge_row<KC>(w_ld + row * KC, w_li + row * KC, w_bd + row * kBuf,
                                     w_bi + row * kBuf, cnt, lane);
      if (lane == 0) {
        w_cnt[row] = 0;
        w_thr[row] = nt;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int i = 0; i < 16; ++i) thr[i] = w_thr[8 * (i >> 2) + 4 * half + (i & 3)];
  };

  // pipeline: the MFMAs of subtile g are issued before the filter of g - 1
  floatx16 acc_p;  // scores of the previous subtile
  float ncol_p = INFINITY;
  int32_t col_p = 0;
  bool have_p = false;
  for (int tile = 0; tile < n_tiles; ++tile) {
    const int buf = tile & 1;
    const float* tnr = reinterpret_cast<const float*>(s_tiles + buf * (L::kTileBytes + L::kNormBytes) +
                                                      L::kTileBytes);
#pragma unroll
    for (int sub = 0; sub < kRefTile / 32; ++sub) {
      const floatx16 acc_c = mma(buf, sub);
      if (have_p && filter(acc_p, ncol_p, col_p)) flush(kBuf - 32);
      acc_p = acc_c;
      ncol_p = tnr[sub * 32 + lrow];
      col_p = static_cast<int32_t>(static_cast<int64_t>(tile) * kRefTile + sub * 32 + lrow);
      have_p = true;
    }
    // stage the next tile into the other buffer (its last reads were the
    // previous tile's MFMAs, before the barrier that ended that iteration)
    __syncthreads();
    stage(buf ^ 1);
    fetch(tile + 2);
    __syncthreads();
  }
  if (have_p && filter(acc_p, ncol_p, col_p)) flush(kBuf - 32);
  flush(0);
  // emit candidate indices
  for (int i = lane; i < kRowsPerWave * KC; i += 64) {
    const int row = i / KC;
    if (q0 + row < m) cand[(q0 + row) * KC + (i % KC)] = w_li[i] == INT_MAX ? -1 : w_li[i];
  }
}

// ------------------------------------------------------------- re-rank ----
// lib.rs:2445-2492 squared_distance with the reference's summation order.
__device__ __forceinline__ float exact_sq_euclid(const float* __restrict__ x,
                                                 const float* __restrict__ y, int d) {
  if (d >= 64) {
    float acc[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) acc[l] = 0.f;
    int i = 0;
    for (; i + 8 <= d; i += 8) {
#pragma unroll
      for (int l = 0; l < 8; ++l) {
        const float df = fsub(x[i + l], y[i + l]);
        acc[l] = fadd(acc[l], fmul(df, df));
      }
    }
    float s = 0.f;
#pragma unroll
    for (int l = 0; l < 8; ++l) s = fadd(s, acc[l]);
    for (; i < d; ++i) {
      const float df = fsub(x[i], y[i]);
      s = fadd(s, fmul(df, df));
    }
    return s;
  }
  float s = 0.f;
  for (int i = 0; i < d; ++i) {
    const float df = fsub(x[i], y[i]);
    s = fadd(s, fmul(df, df));
  }
  return s;
}

// lib.rs:2700-2737 dot product order.
__device__ __forceinline__ float exact_dot(const float* __restrict__ x,
                                           const float* __restrict__ y, int d) {
  if (d >= 64) {
    float acc[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) acc[l] = 0.f;
    int i = 0;
    for (; i + 8 <= d; i += 8) {
#pragma unroll
      for (int l = 0; l < 8; ++l) acc[l] = fadd(acc[l], fmul(x[i + l], y[i + l]));
    }
    float s = 0.f;
#pragma unroll
    for (int l = 0; l < 8; ++l) s = fadd(s, acc[l]);
    for (; i < d; ++i) s = fadd(s, fmul(x[i], y[i]));
    return s;
  }
  float s = 0.f;
  for (int i = 0; i < d; ++i) s = fadd(s, fmul(x[i], y[i]));
  return s;
}

// lib.rs:2603-2646 l2_norm, one thread per row.
__global__ void exact_norms_kernel(const float* __restrict__ x, int64_t n, int d,
                                   float* __restrict__ out) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float* row = x + r * d;
  out[r] = fsqrt(exact_dot(row, row, d));
}

// One group of KC lanes per query row: exact distance of each candidate, a
// bitonic sort by (key, index) across the group, then the first k written.
template <int KC>
__global__ __launch_bounds__(256) void knn_rerank_kernel(
    const float* __restrict__ q, int64_t m, const float* __restrict__ r, int64_t n, int d,
    int metric, const float* __restrict__ qnorm, const float* __restrict__ rnorm,
    const int32_t* __restrict__ cand, int k, int32_t* __restrict__ out_idx,
    float* __restrict__ out_dist) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = gid / KC;
  const int j = gid % KC;
  const bool live = row < m;
  const int64_t rr = live ? row : m - 1;
  int32_t c = live ? cand[rr * KC + j] : -1;
  float key = INFINITY;
  if (c >= 0 && c < n) {
    if (metric == 1) {
      const float dt = exact_dot(q + rr * d, r + static_cast<int64_t>(c) * d, d);
      const float nx = qnorm[rr], ny = rnorm[c];
      if (nx == 0.f && ny == 0.f) {
        key = 0.f;
      } else if (nx == 0.f || ny == 0.f) {
        key = 1.f;
      } else {
        float cs = fdiv(dt, fmul(nx, ny));
        cs = fminf(fmaxf(cs, -1.f), 1.f);
        key = fsub(1.f, cs);
      }
    } else {
      key = exact_sq_euclid(q + rr * d, r + static_cast<int64_t>(c) * d, d);
    }
  } else {
    c = INT_MAX;
  }
  // bitonic sort ascending over the KC-lane group by (total_key(key), c)
  int kk = total_key(key);
#pragma unroll
  for (int size = 2; size <= KC; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int ok = __shfl_xor(kk, stride);
      const int oc = __shfl_xor(c, stride);
      const float okey = __shfl_xor(key, stride);
      const bool up = ((j & size) == 0);
      const bool lower = (j & stride) == 0;
      const bool other_less = ok < kk || (ok == kk && oc < c);
      const bool take = lower ? (up ? other_less : !other_less) : (up ? !other_less : other_less);
      // a pair (lower,upper) must swap iff it is out of order for its direction
      if (take && !(ok == kk && oc == c)) {
        kk = ok;
        c = oc;
        key = okey;
      }
    }
  }
  if (live && j < k) {
    out_idx[row * k + j] = c == INT_MAX ? -1 : c;
    out_dist[row * k + j] = metric == 1 ? key : fsqrt(key);
  }
}

template <typename T, int DPAD, int KC, int TILE, int BUF>
void launch_candidates_cfg(const T* xq, int64_t m, const T* xr, const float* nr, int64_t n,
                           int32_t* cand, hipStream_t s) {
  using L = CandLayout<T, DPAD, KC, TILE, BUF>;
  auto kern = knn_candidates_kernel<T, DPAD, KC, TILE, BUF>;
  HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, L::kTotal));
  const unsigned blocks = blocks_for(static_cast<size_t>(m), L::kRows);
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(kWaves * 64), L::kTotal, s, xq, m, xr, nr, n, cand);
  KERNEL_CHECK();
}

// Largest configuration that fits in LDS: 64-row reference tiles with
// 64-deep pass buffers, then 32-row tiles, then 48- and 32-deep buffers.
template <typename T, int DPAD, int KC>
void launch_candidates(const T* xq, int64_t m, const T* xr, const float* nr, int64_t n,
                       int32_t* cand, hipStream_t s) {
  if constexpr (CandLayout<T, DPAD, KC, 64, 64>::kTotal <= kLdsBudget)
    launch_candidates_cfg<T, DPAD, KC, 64, 64>(xq, m, xr, nr, n, cand, s);
  else if constexpr (CandLayout<T, DPAD, KC, 32, 64>::kTotal <= kLdsBudget)
    launch_candidates_cfg<T, DPAD, KC, 32, 64>(xq, m, xr, nr, n, cand, s);
  else if constexpr (CandLayout<T, DPAD, KC, 32, 48>::kTotal <= kLdsBudget)
    launch_candidates_cfg<T, DPAD, KC, 32, 48>(xq, m, xr, nr, n, cand, s);
  else if constexpr (CandLayout<T, DPAD, KC, 32, 32>::kTotal <= kLdsBudget)
    launch_candidates_cfg<T, DPAD, KC, 32, 32>(xq, m, xr, nr, n, cand, s);
  else
    throw_invalid("this feature count / candidate count does not fit the kNN kernel's LDS budget");
}

template <typename T, int KC>
void dispatch_dpad(int dpad, const T* xq, int64_t m, const T* xr, const float* nr, int64_t n,
                   int32_t* cand, hipStream_t s) {
  switch (dpad) {
    case 32: launch_candidates<T, 32, KC>(xq, m, xr, nr, n, cand, s); break;
    case 64: launch_candidates<T, 64, KC>(xq, m, xr, nr, n, cand, s); break;
    case 128: launch_candidates<T, 128, KC>(xq, m, xr, nr, n, cand, s); break;
    case 192: launch_candidates<T, 192, KC>(xq, m, xr, nr, n, cand, s); break;
    case 256: launch_candidates<T, 256, KC>(xq, m, xr, nr, n, cand, s); break;
    default: throw_invalid("unsupported padded feature count");
  }
}

}  // namespace

int knn_padded_dim(int64_t d) {
  for (int p : {32, 64, 128, 192, 256})
    if (d <= p) return p;
  return -1;
}

int knn_candidate_count(int k, int rerank_factor) {
  const int want = std::max(k, k * std::max(rerank_factor, 1));
  if (want <= 32) return 32;
  if (want <= 64) return 64;
  return -1;
}

void knn_device(const float* q, int64_t m, const float* r, int64_t n, int64_t d, int k,
                int metric, int precision, int rerank_factor, int32_t* out_idx,
                float* out_dist, KnnWorkspace& ws, hipStream_t s, KnnStats* stats) {
  maybe_inject_fault("knn");
  if (m == 0) return;
  const int dpad = knn_padded_dim(d);
  if (dpad < 0) throw_invalid("exact GPU kNN supports at most 256 features");
  const int kc = knn_candidate_count(k, rerank_factor);
  if (kc < 0 || k > kc) throw_invalid("n_neighbors * rerank_factor must be <= 64");
  if (k > n) throw_invalid("n_neighbors must not exceed the number of reference rows");
  const bool self = (q == r) && (m == n);
  const bool bf16 = precision == 1;
  const size_t esz = bf16 ? 2 : 4;

  // column means of the (normalised) reference rows for centring
  ws.colsum.ensure(static_cast<size_t>(d));
  HIP_CHECK(hipMemsetAsync(ws.colsum.get(), 0, sizeof(double) * d, s));
  const unsigned cs_blocks = static_cast<unsigned>(std::min<int64_t>(n, 1024));
  if (metric == 1) {
    hipLaunchKernelGGL(column_sums_normalized_kernel, dim3(cs_blocks), dim3(256),
                       sizeof(double) * d, s, r, n, static_cast<int>(d), ws.colsum.get());
  } else {
    hipLaunchKernelGGL(column_sums_kernel, dim3(cs_blocks), dim3(std::max<int64_t>(64, ((d + 63) / 64) * 64)),
                       0, s, r, n, static_cast<int>(d), 0, ws.colsum.get());
  }
  KERNEL_CHECK();
  const double inv_n = 1.0 / static_cast<double>(n);

  ws.rprep.ensure(static_cast<size_t>(n) * dpad * esz);
  ws.rnorm2.ensure(static_cast<size_t>(n));
  if (!self) {
    ws.qprep.ensure(static_cast<size_t>(m) * dpad * esz);
    ws.qnorm2.ensure(static_cast<size_t>(m));
  }
  auto prep = [&](const float* x, int64_t rows, void* out, float* nrm) {
    const unsigned blocks = blocks_for(static_cast<size_t>(rows), 4);
    if (bf16)
      hipLaunchKernelGGL(prep_rows_kernel<uint16_t>, dim3(blocks), dim3(256), 0, s, x, rows,
                         static_cast<int>(d), dpad, metric == 1 ? 1 : 0, ws.colsum.get(), inv_n,
                         static_cast<uint16_t*>(out), nrm);
    else
      hipLaunchKernelGGL(prep_rows_kernel<float>, dim3(blocks), dim3(256), 0, s, x, rows,
                         static_cast<int>(d), dpad, metric == 1 ? 1 : 0, ws.colsum.get(), inv_n,
                         static_cast<float*>(out), nrm);
    KERNEL_CHECK();
  };
  prep(r, n, ws.rprep.get(), ws.rnorm2.get());
  const void* qp = ws.rprep.get();
  if (!self) {
    prep(q, m, ws.qprep.get(), ws.qnorm2.get());
    qp = ws.qprep.get();
  }

  ws.cand.ensure(static_cast<size_t>(m) * kc);
  if (stats) HIP_CHECK(hipEventRecord(stats->cand_start, s));
  if (bf16) {
    if (kc == 32)
      dispatch_dpad<uint16_t, 32>(dpad, static_cast<const uint16_t*>(qp), m,
                                  reinterpret_cast<const uint16_t*>(ws.rprep.get()),
                                  ws.rnorm2.get(), n, ws.cand.get(), s);
    else
      dispatch_dpad<uint16_t, 64>(dpad, static_cast<const uint16_t*>(qp), m,
                                  reinterpret_cast<const uint16_t*>(ws.rprep.get()),
                                  ws.rnorm2.get(), n, ws.cand.get(), s);
  } else {
    if (kc == 32)
      dispatch_dpad<float, 32>(dpad, static_cast<const float*>(qp), m,
                               reinterpret_cast<const float*>(ws.rprep.get()), ws.rnorm2.get(),
                               n, ws.cand.get(), s);
    else
      dispatch_dpad<float, 64>(dpad, static_cast<const float*>(qp), m,
                               reinterpret_cast<const float*>(ws.rprep.get()), ws.rnorm2.get(),
                               n, ws.cand.get(), s);
  }
  if (stats) {
    HIP_CHECK(hipEventRecord(stats->cand_stop, s));
    stats->flops = 2.0 * static_cast<double>(m) * static_cast<double>(n) * static_cast<double>(d);
    stats->recorded = true;
  }

  // exact norms for cosine re-rank
  const float* qn = nullptr;
  const float* rn = nullptr;
  if (metric == 1) {
    ws.rnorm.ensure(static_cast<size_t>(n));
    hipLaunchKernelGGL(exact_norms_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, r, n,
                       static_cast<int>(d), ws.rnorm.get());
    KERNEL_CHECK();
    rn = ws.rnorm.get();
    qn = rn;
    if (!self) {
      ws.qnorm.ensure(static_cast<size_t>(m));
      hipLaunchKernelGGL(exact_norms_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, s, q, m,
                         static_cast<int>(d), ws.qnorm.get());
      KERNEL_CHECK();
      qn = ws.qnorm.get();
    }
  }
  const size_t threads = static_cast<size_t>(m) * kc;
  if (kc == 32)
    hipLaunchKernelGGL(knn_rerank_kernel<32>, dim3(blocks_for(threads, 256)), dim3(256), 0, s, q,
                       m, r, n, static_cast<int>(d), metric, qn, rn, ws.cand.get(), k, out_idx,
                       out_dist);
  else
    hipLaunchKernelGGL(knn_rerank_kernel<64>, dim3(blocks_for(threads, 256)), dim3(256), 0, s, q,
                       m, r, n, static_cast<int>(d), metric, qn, rn, ws.cand.get(), k, out_idx,
                       out_dist);
  KERNEL_CHECK();
}

}  // namespace umapx
