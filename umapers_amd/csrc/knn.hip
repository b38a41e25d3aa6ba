// Exact k-nearest neighbours on gfx950.
//
// Replaces exact_nearest_neighbors(_euclidean/_cosine) and
// exact_nearest_neighbors_to_reference (rust_umap/src/lib.rs:2863-2928,
// 3283-3399).  Two passes:
//   1. candidates: x.y - ||y||^2/2 score tiles on MFMA (FP32 32x32x2, or BF16
//      32x32x16 with FP32 accumulation) over mean-centred rows; every lane
//      owns one query and half of each reference subtile and keeps its best
//      H (score, index) pairs sorted in registers, so each query gets 2*H
//      candidates;
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
#include <type_traits>

namespace umapx {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

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
  // one thread per feature column (or several, past 1024 columns)
  (void)normalize;
  for (int f = threadIdx.x; f < d; f += blockDim.x) {
    double s = 0.0;
    for (int64_t r = blockIdx.x; r < n; r += gridDim.x) s += static_cast<double>(x[r * d + f]);
    atomicAdd(&colsum[f], s);
  }
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
// each row and, when half_norm is given, -||row||^2 / 2 of the rounded row,
// with -inf for the rows [n, norm_rows) that pad the last reference tile.
// One wave per row.
template <typename T>
__global__ void prep_rows_kernel(const float* __restrict__ x, int64_t n, int64_t norm_rows, int d,
                                 int dpad, int normalize, const double* __restrict__ colsum,
                                 double inv_n, T* __restrict__ out,
                                 float* __restrict__ half_norm) {
  const int lane = threadIdx.x & 63;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= n) {
    if (r < norm_rows && lane == 0 && half_norm) half_norm[r] = -INFINITY;
    return;
  }
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
  if (lane == 0 && half_norm) half_norm[r] = -0.5f * nrm;
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

// Candidate pass.  The MFMA takes the reference subtile as its first operand
// and the query group as its second, so in the 32x32 accumulator lane l holds
// query l%32 against the 16 reference rows 8*(i/4) + 4*(l/32) + i%4
// (i = 0..15): every lane owns one (query, half of the references) pair and
// filters its scores against its own threshold with no cross-lane traffic.
// The accumulator starts at -||y||^2/2 and the MFMA adds x.y, giving
// x.y - ||y||^2/2 = -(||x - y||^2 - ||x||^2)/2: larger is nearer.  Each lane
// keeps its best H (value, index) pairs sorted in registers; values above the
// lane's threshold are appended to a lane-private LDS buffer and inserted in
// batches (all 64 lanes at once) when a buffer could not take another
// subtile.  The two halves of a query together give 2*H candidates, a
// superset of the query's top H >= k under the same scores.
template <int H>
struct LaneTop {
  float v[H];      // descending
  int32_t c[H];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int j = 0; j < H; ++j) {
      v[j] = -INFINITY;
      c[j] = -1;
    }
  }
  __device__ __forceinline__ void insert(float x, int32_t ci) {
#pragma unroll
    for (int j = H - 1; j > 0; --j) {
      const bool up = x > v[j - 1];
      const bool here = x > v[j];
      c[j] = up ? c[j - 1] : (here ? ci : c[j]);
      v[j] = up ? v[j - 1] : (here ? x : v[j]);
    }
    c[0] = x > v[0] ? ci : c[0];
    v[0] = x > v[0] ? x : v[0];
  }
};

constexpr int kLdsBudget = 160 * 1024;

// Direct global-to-LDS loads (the wave's LDS address is uniform; lane i writes
// its bytes at dst + i * size).
__device__ __forceinline__ void lds_load16(const void* src, unsigned char* dst) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}
__device__ __forceinline__ void lds_load4(const void* src, unsigned char* dst) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 4, 0, 0);
}
// Waits for this wave's outstanding vector-memory loads (the direct loads).
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
constexpr int kTile = 32;   // reference rows per staged tile (one MFMA subtile)

// Entries of the half-norm array: whole tiles, the padding -inf.
int64_t knn_norm_rows(int64_t n) { return (n + kTile - 1) / kTile * kTile; }

// LDS of one block: a ring of reference tiles with their half norms and the
// waves' pass buffers ([slot][lane] pairs of 8 bytes).  Twelve waves (three
// per SIMD, at most 168 registers each) when the query fragments are at most
// four 16-byte vectors a lane and a direct ring fits; else eight (two per
// SIMD, at most 256 registers each) when the buffers can stay at least 20
// deep in the three-tile ring and the resident query fragments plus the top
// list leave room; else four.  The next tile's fragments are read into
// registers ahead of use when they are at most eight 16-byte vectors per
// lane.  kSuper tiles share one block-wide barrier: the ring holds two groups
// of kSuper tiles (one being multiplied, the next arriving), the largest of 4
// and 2 that fits; when neither fits, the ring holds three tiles with a
// barrier per tile and block-synchronous drains (register staging, below),
// and for rows too wide for that (past 256 features) two one-tile groups.
//
// The groups of kSuper tiles are loaded straight from global memory into LDS
// (global_load_lds_dwordx4: each lane of a wave writes 16 bytes, the wave one
// contiguous kilobyte), a whole group ahead, so no registers hold tiles in
// flight and the load of group g + 1 has all of group g's multiplies to
// arrive.  Rows are then unpadded; row r's 16-byte chunk c sits at chunk
// c ^ swz(r), so the 32 lanes that read one chunk of 32 rows at once hit
// distinct banks.  The three-tile ring keeps register staging and rows
// padded by 16 bytes (kDirect false).
template <typename T, int DPAD, int H>
struct CandShape {
  static constexpr int kNV = DPAD / 2 / (16 / static_cast<int>(sizeof(T)));
  static constexpr bool kPrefetch = kNV <= 8;
  static constexpr int kChunks = DPAD * static_cast<int>(sizeof(T)) / 16;  // per row
  static constexpr int kFlagBytes = 16;
  static constexpr int row_bytes(bool direct) {
    return DPAD * static_cast<int>(sizeof(T)) + (direct ? 0 : 16);
  }
  static constexpr int stage_bytes(bool direct) { return kTile * row_bytes(direct) + kTile * 4; }
  static constexpr int ring(int super, bool direct) { return direct ? 2 * super : 3; }
  static constexpr int bytes(int waves, int super, bool direct, int buf) {
    return ring(super, direct) * stage_bytes(direct) + (direct ? 0 : kFlagBytes) +
           waves * 64 * buf * 8;
  }
  static constexpr int pick_buf(int waves, int super, bool direct) {
    for (int b = 32; b >= 20; --b)
      if (bytes(waves, super, direct, b) <= kLdsBudget) return b;
    return 0;
  }
  static constexpr bool kRegsFit8 = H <= 16 && (kPrefetch || 4 * kNV + 2 * H <= 120);
  // fragments of at most four 16-byte vectors a lane leave room for three
  // waves per SIMD (at most 168 registers each) when the direct ring fits
  static constexpr bool kFit12 = H <= 16 && kNV <= 4 && bytes(12, 4, true, 20) <= kLdsBudget;
  static constexpr int kWaves = kFit12 ? 12 : ((pick_buf(8, 1, false) && kRegsFit8) ? 8 : 4);
  static constexpr int kSuper =
      pick_buf(kWaves, 4, true) ? 4 : (pick_buf(kWaves, 2, true) ? 2 : 1);
  static constexpr bool kDirect = kSuper > 1 || !pick_buf(kWaves, 1, false);
  static constexpr int kRowBytes = row_bytes(kDirect);
  static constexpr int kStageBytes = stage_bytes(kDirect);
  static constexpr int kRing = ring(kSuper, kDirect);
  static constexpr int kTiles = kRing * kStageBytes;
  static constexpr int kBuf = pick_buf(kWaves, kSuper, kDirect);
  static_assert(kBuf >= 20, "kNN candidate tiles do not fit in LDS");
  static constexpr int kBytes = bytes(kWaves, kSuper, kDirect, kBuf);
  // direct loads: kilobyte loads per tile, and the chunk swizzle
  // swz(r) = (r ^ (r >> kSwzShift)) & kSwzMask, kSwzMask + 1 being the largest
  // power of two up to 16 that divides kChunks (the swizzle stays in the row)
  static constexpr int kLoads = kTile * kChunks / 64;
  static constexpr int kSwzMask = (kChunks % 16 == 0 ? 16 : kChunks % 8 == 0 ? 8 : 4) - 1;
  static constexpr int kSwzShift = kSwzMask == 15 ? 4 : kSwzMask == 7 ? 3 : 2;
  static_assert(!kDirect || (kTile * kChunks) % 64 == 0, "direct tile loads are whole kilobytes");
  static_assert(!kDirect || kChunks % 4 == 0, "direct tile rows are whole 64-byte lines");
  static constexpr int kQueries = kWaves * 32;  // queries per block
};

// Blocks below full_groups each take one query group against all reference
// rows.  The remaining groups (the last, partial round of blocks) are split:
// block full_groups + t takes group full_groups + t / tail_split against rows
// [s * tail_len, (s + 1) * tail_len), s = t % tail_split, and leaves
// (score, index) lists for knn_merge_splits_kernel.  The loop is a
// pipeline over 32-row tiles: tile t + kSuper is staged while tile t is
// multiplied, one block-wide barrier closes every group of kSuper tiles, and
// the fragments of tile t + 1 are read from LDS into the registers tile t's
// MFMAs have just consumed (when they fit) before tile t's scores are
// filtered.  A wave whose buffers could not take another tile drains them
// between tiles on its own.  With register staging the ring holds three tiles
// (tile t + 2 staged during tile t) and a wave that needs a drain raises a
// flag; after the per-tile barrier every wave of the block drains, so no wave
// idles at the barrier while another drains.  Either way the drain sits
// outside the tile loop, so the loop itself never modifies the top lists and
// keeps one register assignment.
template <typename T, int DPAD, int H>
__global__ __launch_bounds__((CandShape<T, DPAD, H>::kWaves * 64)) void knn_candidates_kernel(
    const T* __restrict__ xq, int64_t m, const T* __restrict__ xr,
    const float* __restrict__ nr_half, int64_t n, int64_t full_groups, int tail_split,
    int64_t tail_len, int interleave, int share, int32_t* __restrict__ cand_idx,
    uint2* __restrict__ cand_pairs) {
  using C = CandShape<T, DPAD, H>;
  constexpr int WAVES = C::kWaves;
  constexpr int BUF = C::kBuf;
  constexpr int kThreads = WAVES * 64;
  constexpr int VEC = 16 / static_cast<int>(sizeof(T));  // elements per 16 B
  constexpr int HALF = DPAD / 2;
  constexpr int NV = HALF / VEC;                        // 16-byte vectors per lane
  constexpr bool kPrefetch = C::kPrefetch;
  extern __shared__ __align__(16) unsigned char lds[];
  unsigned char* s_tiles = lds;
  uint8_t* s_flag = lds + C::kTiles;  // register staging only

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l32 = lane & 31;
  const int half = lane >> 5;
  const bool split_block = static_cast<int64_t>(blockIdx.x) >= full_groups;
  const int64_t tail_block = static_cast<int64_t>(blockIdx.x) - full_groups;
  const int split = split_block ? static_cast<int>(tail_block % tail_split) : 0;
  const int64_t group = split_block ? full_groups + tail_block / tail_split : blockIdx.x;
  const int64_t q0 = group * C::kQueries + wave * 32;
  const bool live = q0 + l32 < m;
  // a split takes the reference rows [r0, r1) (tail_len rows each), or with
  // interleave > 1 every interleave-th 32-row tile from tile `split` on
  const bool inter = split_block && interleave > 1;
  const int64_t r0 = split_block && !inter ? static_cast<int64_t>(split) * tail_len : 0;
  const int64_t r1 = split_block && !inter ? (r0 + tail_len < n ? r0 + tail_len : n) : n;
  const int64_t all_tiles = (n + kTile - 1) / kTile;
  const int n_tiles = inter ? static_cast<int>((all_tiles - split + interleave - 1) / interleave)
                            : static_cast<int>((r1 - r0 + kTile - 1) / kTile);
  // first reference row of tile t: row0 + t * row_step (block-uniform)
  const int64_t row0 = inter ? static_cast<int64_t>(split) * kTile : r0;
  const int64_t row_step = inter ? static_cast<int64_t>(interleave) * kTile : kTile;
  auto tile_row = [&](int t) -> int64_t { return row0 + static_cast<int64_t>(t) * row_step; };
  uint2* wbuf = reinterpret_cast<uint2*>(lds + C::kTiles + (C::kDirect ? 0 : C::kFlagBytes)) +
                wave * 64 * BUF + lane;

  u32x4 qf[NV];
  {
    const int64_t qr = live ? (q0 + l32) : (m - 1);
    const u32x4* src = reinterpret_cast<const u32x4*>(xq + qr * DPAD + half * HALF);
#pragma unroll
    for (int v = 0; v < NV; ++v) qf[v] = src[v];
  }

  constexpr int CHUNKS = DPAD * static_cast<int>(sizeof(T)) / 16;
  constexpr int PER_THREAD = (kTile * CHUNKS + kThreads - 1) / kThreads;
  // a thread's share of one tile in flight from global memory to LDS
  struct Pre {
    u32x4 v[PER_THREAD];
    float nr;
  };
  auto fetch = [&](int tile, Pre& p) {
    const int64_t base = tile_row(tile);
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i) {
      const int c = tid + i * kThreads;
      u32x4 v = {0u, 0u, 0u, 0u};
      if (tile < n_tiles && c < kTile * CHUNKS && base + c / CHUNKS < r1)
        v = reinterpret_cast<const u32x4*>(xr + (base + c / CHUNKS) * DPAD)[c % CHUNKS];
      p.v[i] = v;
    }
    p.nr = -INFINITY;
    if (tid < kTile && tile < n_tiles && base + tid < r1) p.nr = nr_half[base + tid];
  };
  auto stage = [&](int tile, const Pre& p) {
    unsigned char* t = s_tiles + (tile % C::kRing) * C::kStageBytes;
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i) {
      const int c = tid + i * kThreads;
      if (c < kTile * CHUNKS)
        *reinterpret_cast<u32x4*>(t + (c / CHUNKS) * C::kRowBytes + (c % CHUNKS) * 16) = p.v[i];
    }
    if (tid < kTile) reinterpret_cast<float*>(t + kTile * C::kRowBytes)[tid] = p.nr;
  };
  // this wave's share of tile `tile`'s direct loads: kilobyte j of its rows
  // goes to wave j % WAVES, the half norms (padded with -inf to whole tiles
  // by the caller) to wave kLoads % WAVES; rows past n repeat row n - 1,
  // whose scores the -inf norms discard
  auto issue = [&](int tile) {
    if (tile >= n_tiles) return;
    unsigned char* st = s_tiles + (tile % C::kRing) * C::kStageBytes;
    const int64_t base = tile_row(tile);
    for (int j = wave; j < C::kLoads; j += WAVES) {
      const int pos = j * 64 + lane;  // chunk position in the tile's rows
      const int r = pos / C::kChunks;
      const int c = (pos % C::kChunks) ^ ((r ^ (r >> C::kSwzShift)) & C::kSwzMask);
      const int64_t row = base + r < n ? base + r : n - 1;
      lds_load16(xr + row * DPAD + c * VEC, st + j * 1024);
    }
    if (wave == C::kLoads % WAVES && lane < kTile)
      lds_load4(nr_half + base + lane, st + kTile * C::kRowBytes);
  };
  // byte offset of this lane's fragment v (chunk half * NV + v of row l32)
  const int swz = C::kDirect ? ((l32 ^ (l32 >> C::kSwzShift)) & C::kSwzMask) : 0;
  auto frag_off = [&](int v) { return l32 * C::kRowBytes + ((half * NV + v) ^ swz) * 16; };
  // fragments (first MFMA operand) and accumulator seeds of one tile
  struct Frags {
    u32x4 rf[kPrefetch ? NV : 1];
    floatx16 seed;
  };
  auto load = [&](int tile, Frags& f) {
    const unsigned char* st = s_tiles + (tile % C::kRing) * C::kStageBytes;
    if constexpr (kPrefetch) {
#pragma unroll
      for (int v = 0; v < NV; ++v) f.rf[v] = *reinterpret_cast<const u32x4*>(st + frag_off(v));
    }
    const float* tn = reinterpret_cast<const float*>(st + kTile * C::kRowBytes) + 4 * half;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const floatx4 nv = *reinterpret_cast<const floatx4*>(tn + 8 * g);
      f.seed[4 * g + 0] = nv[0];
      f.seed[4 * g + 1] = nv[1];
      f.seed[4 * g + 2] = nv[2];
      f.seed[4 * g + 3] = nv[3];
    }
  };
  auto mma = [&](int tile, const Frags& f) {
    floatx16 acc = f.seed;
    if constexpr (kPrefetch) {
#pragma unroll
      for (int v = 0; v < NV; ++v) acc = MfmaSteps<T, MfmaOp<T>::kSteps>::run(f.rf[v], qf[v], acc);
    } else {
      const unsigned char* st = s_tiles + (tile % C::kRing) * C::kStageBytes;
#pragma unroll
      for (int v = 0; v < NV; ++v)
        acc = MfmaSteps<T, MfmaOp<T>::kSteps>::run(
            *reinterpret_cast<const u32x4*>(st + frag_off(v)), qf[v], acc);
    }
    return acc;
  };

  LaneTop<H> top;
  top.init();
  float thr = live ? -INFINITY : INFINITY;  // lanes past m take nothing
  int cnt = 0;
  // appends this lane's passing scores to its buffer
  auto filter = [&](const floatx16& acc, int32_t colbase) {
    float mx = acc[0];
#pragma unroll
    for (int i = 1; i < 16; ++i) mx = fmaxf(mx, acc[i]);
    if (__ballot(mx > thr) == 0ULL) return;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (acc[i] > thr) {
        wbuf[cnt * 64] = make_uint2(__float_as_uint(acc[i]),
                                    static_cast<uint32_t>(colbase + 8 * (i >> 2) + (i & 3)));
        ++cnt;
      }
    }
  };
  auto flush = [&](auto with_union) {
    for (int j = 0; __ballot(j < cnt) != 0ULL; ++j) {
      if (j < cnt) {
        const uint2 e = wbuf[j * 64];
        const float x = __uint_as_float(e.x);
        if (x > thr) {
          top.insert(x, static_cast<int32_t>(e.y));
          thr = fmaxf(thr, top.v[H - 1]);
        }
      }
    }
    cnt = 0;
    if constexpr (!decltype(with_union)::value) return;
    if (!share) return;  // each half keeps its own top H (kernel-uniform)
    // The two halves of a query only need the H best of their union: a
    // score below the H-th best of both lists can never make that set.  That
    // H-th best is the max over i + j = H of min(i-th best here, j-th best
    // there); every term is a lower bound of it, so the even and the two
    // quarter splits, and both ends, give a valid (nearly tight) threshold.
    // exchanged through this lane's first two buffer slots, empty now
    wbuf[0] = make_uint2(__float_as_uint(top.v[H - 1]), __float_as_uint(top.v[H / 2 - 1]));
    wbuf[64] = make_uint2(__float_as_uint(top.v[3 * H / 4 - 1]), __float_as_uint(top.v[H / 4 - 1]));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint2 o1 = wbuf[(lane ^ 32) - lane];
    const uint2 o2 = wbuf[64 + (lane ^ 32) - lane];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const float b1 = __uint_as_float(o1.x), b2 = __uint_as_float(o1.y);
    const float b3 = __uint_as_float(o2.x), b4 = __uint_as_float(o2.y);
    const float kth = fmaxf(fmaxf(fmaxf(top.v[H - 1], b1), fminf(top.v[H / 2 - 1], b2)),
                            fmaxf(fminf(top.v[H / 4 - 1], b3), fminf(top.v[3 * H / 4 - 1], b4)));
    thr = fmaxf(thr, kth);
  };

  Pre pre;
  Frags cur;
  int tile = 0;
  if constexpr (!C::kDirect) {
    fetch(0, pre);
    stage(0, pre);
    fetch(1, pre);
    stage(1, pre);
    fetch(2, pre);
    if (tid < C::kFlagBytes) s_flag[tid] = 0;
    __syncthreads();
    load(0, cur);
    while (tile < n_tiles) {
      bool drain = false;
      for (; tile < n_tiles && !drain; ++tile) {
        const floatx16 acc = mma(tile, cur);
        // the next tile's fragments reuse the registers the MFMAs just consumed
        __builtin_amdgcn_sched_barrier(0);
        if (tile + 1 < n_tiles) load(tile + 1, cur);
        // slot (tile + 2) % 3 last held tile - 1, whose reads all ended before
        // the barrier that closed the previous iteration
        stage(tile + 2, pre);
        fetch(tile + 3, pre);
        filter(acc, static_cast<int32_t>(tile_row(tile) + 4 * half));
        const bool need = __ballot(cnt > BUF - 16) != 0ULL;
        if (lane == 0) s_flag[(tile & 1) * 8 + wave] = need ? 1 : 0;
        __syncthreads();
        // block-uniform; readfirstlane lets the compiler see it as uniform
        drain = __builtin_amdgcn_readfirstlane(
                    *reinterpret_cast<const uint64_t*>(s_flag + (tile & 1) * 8) != 0ULL ? 1 : 0) != 0;
      }
      if (drain) flush(std::true_type{});
    }
  } else {
    constexpr int S = C::kSuper;
    // group 0 now, group 1 while group 0 is multiplied
#pragma unroll
    for (int t = 0; t < S; ++t) issue(t);
    wait_vmem();
    __syncthreads();
#pragma unroll
    for (int t = S; t < 2 * S; ++t) issue(t);
    load(0, cur);
    while (tile < n_tiles) {
      bool need = false;
      for (; tile < n_tiles && !need; ++tile) {
        const floatx16 acc = mma(tile, cur);
        __builtin_amdgcn_sched_barrier(0);
        if (tile % S == S - 1) {
          // the next group has landed (this wave's loads; the barrier
          // publishes every wave's), and every wave has read this group's
          // slots, which now take the group after the next
          wait_vmem();
          __syncthreads();
#pragma unroll
          for (int t = 1; t <= S; ++t) issue(tile + S + t);
        }
        // within a group the next tile is already visible; across groups it
        // became visible at the barrier above
        if (tile + 1 < n_tiles) load(tile + 1, cur);
        filter(acc, static_cast<int32_t>(tile_row(tile) + 4 * half));
        need = __ballot(cnt > BUF - 16) != 0ULL;
      }
      if (need) flush(std::true_type{});
    }
  }
  flush(std::false_type{});
  if (!live) return;
  const int64_t q = q0 + l32;
  if (!split_block) {
    int32_t* dst = cand_idx + q * (2 * H) + half * H;
#pragma unroll
    for (int j = 0; j < H; ++j) dst[j] = top.c[j];
  } else {
    const int64_t qt = q - full_groups * C::kQueries;
    uint2* dst = cand_pairs + ((qt * tail_split + split) * 2 + half) * H;
#pragma unroll
    for (int j = 0; j < H; ++j) dst[j] = make_uint2(__float_as_uint(top.v[j]), static_cast<uint32_t>(top.c[j]));
  }
}

// Best H of the n_split lists of each (query, half) of the split query rows
// [q_begin, m): the per-split top-H lists together hold the overall top H.
// One thread per (query, half).
template <int H>
__global__ void knn_merge_splits_kernel(const uint2* __restrict__ pairs, int64_t q_begin,
                                        int64_t m, int n_split, int32_t* __restrict__ cand_idx) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= 2 * (m - q_begin)) return;
  const int64_t qt = t >> 1;
  const int64_t q = q_begin + qt;
  const int half = static_cast<int>(t & 1);
  LaneTop<H> top;
  top.init();
  for (int s = 0; s < n_split; ++s) {
    const uint2* src = pairs + ((qt * n_split + s) * 2 + half) * H;
    for (int j = 0; j < H; ++j) {
      const uint2 e = src[j];
      const float x = __uint_as_float(e.x);
      if (static_cast<int32_t>(e.y) >= 0 && x > top.v[H - 1]) top.insert(x, static_cast<int32_t>(e.y));
    }
  }
  int32_t* dst = cand_idx + q * (2 * H) + half * H;
#pragma unroll
  for (int j = 0; j < H; ++j) dst[j] = top.c[j];
}

// 17 .. 32 neighbours from the 16-entry lists: each query's reference rows
// are split two ways by interleaved 32-row tiles and each split's two halves
// keep their own top 16 (no shared threshold), so every query gets four
// top-16 lists of disjoint row sets, 64 candidates in all.  They hold the
// query's 32 best scores unless some list was full with its 16th best at or
// above the 32nd best of the union (it may have dropped one of them); such
// queries are listed for the 32-entry kernel.  One thread per query.
__global__ void knn_merge_check_kernel(const uint2* __restrict__ pairs, int64_t m,
                                       int32_t* __restrict__ cand, int32_t* __restrict__ redo,
                                       int32_t* __restrict__ redo_count) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const uint2* src = pairs + q * 64;
  LaneTop<32> top;
  top.init();
  for (int j = 0; j < 64; ++j) {
    const uint2 e = src[j];
    const int32_t c = static_cast<int32_t>(e.y);
    cand[q * 64 + j] = c;
    const float x = __uint_as_float(e.x);
    if (c >= 0 && x > top.v[31]) top.insert(x, c);
  }
  const float s32 = top.v[31];  // -inf with fewer than 32 entries in all
  bool safe = true;
  for (int l = 0; l < 4; ++l) {
    const uint2 last = src[l * 16 + 15];
    if (static_cast<int32_t>(last.y) >= 0 && !(__uint_as_float(last.x) < s32)) safe = false;
  }
  if (!safe) redo[atomicAdd(redo_count, 1)] = static_cast<int32_t>(q);
}

// rows[i] = src[list[i]] for rows of `width` elements of T (16-byte vectors)
template <typename T>
__global__ void gather_rows16_kernel(const T* __restrict__ src, const int32_t* __restrict__ list,
                                     int64_t count, int width, T* __restrict__ dst) {
  const int vecs = width * static_cast<int>(sizeof(T)) / 16;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count * vecs) return;
  const int64_t i = t / vecs;
  const int v = static_cast<int>(t % vecs);
  reinterpret_cast<u32x4*>(dst + i * width)[v] =
      reinterpret_cast<const u32x4*>(src + static_cast<int64_t>(list[i]) * width)[v];
}

// cand[list[i]] = sub[i], 64 candidates a row
__global__ void scatter_cand64_kernel(const int32_t* __restrict__ sub,
                                      const int32_t* __restrict__ list, int64_t count,
                                      int32_t* __restrict__ cand) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= count * 64) return;
  const int64_t i = t >> 6;
  cand[static_cast<int64_t>(list[i]) * 64 + (t & 63)] = sub[t];
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
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t row = gid / KC;
  const int j = static_cast<int>(gid % KC);
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

// Whole rounds of query groups (one block per CU) scan all reference rows;
// the groups of a last, partial round are split over the reference rows so
// that round ends sooner.  Returns the split count for those groups: the
// count minimising the round's length, with a small charge per split for the
// extra threshold warm-up, keeping at least eight tiles per split.
int choose_tail_split(int64_t tail_groups, int64_t n, int cus) {
  if (tail_groups <= 0) return 1;
  int best = 1;
  double best_cost = 1e300;
  for (int s = 1; s <= 32; ++s) {
    if (s > 1 && n < static_cast<int64_t>(s) * 8 * kTile) break;
    const int64_t rounds = (tail_groups * s + cus - 1) / cus;
    const double cost = static_cast<double>(rounds) / s + 0.01 * s;
    if (cost < best_cost) {
      best_cost = cost;
      best = s;
    }
  }
  return best;
}

int device_cus() {
  int dev = 0, cus = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus > 0 ? cus : 1;
}

// How a launch of query groups of `queries` rows covers m queries: `full`
// groups scan every reference row; the rest are split `tail_split` ways over
// slices of `tail_len` reference rows, leaving `tail_rows` query rows whose
// split lists (2 * H pairs per row and split) are merged afterwards.
struct TailPlan {
  int64_t groups = 0, full = 0, tail_len = 0, tail_rows = 0;
  int tail_split = 1;
};

TailPlan plan_tail(int64_t m, int64_t n, int queries, int cus) {
  TailPlan t;
  t.groups = (m + queries - 1) / queries;
  t.full = t.groups / cus * cus;
  int tail_split = choose_tail_split(t.groups - t.full, n, cus);
  auto split_rows = [&](int sp) { return ((n + sp - 1) / sp + kTile - 1) / kTile * kTile; };
  while (tail_split > 1 && (tail_split - 1) * split_rows(tail_split) >= n) --tail_split;  // no empty split
  if (tail_split == 1) t.full = t.groups;
  t.tail_split = tail_split;
  t.tail_len = split_rows(tail_split);
  t.tail_rows = m - t.full * queries;
  return t;
}

template <typename T, int DPAD, int H>
void launch_candidates(const T* xq, int64_t m, const T* xr, const float* nr_half, int64_t n,
                       int32_t* cand, DevBuf<uint2>& pairs, hipStream_t s) {
  using C = CandShape<T, DPAD, H>;
  auto kern = knn_candidates_kernel<T, DPAD, H>;
  HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, C::kBytes));
  const TailPlan tp = plan_tail(m, n, C::kQueries, device_cus());
  const int64_t groups = tp.groups, full = tp.full, tail_len = tp.tail_len,
                tail_rows = tp.tail_rows;
  const int tail_split = tp.tail_split;
  uint2* pp = nullptr;
  if (tail_rows > 0) {
    pairs.ensure(static_cast<size_t>(tail_rows) * tail_split * 2 * H);
    pp = pairs.get();
  }
  const int64_t blocks = full + (groups - full) * tail_split;
  hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(blocks)), dim3(C::kWaves * 64), C::kBytes, s,
                     xq, m, xr, nr_half, n, full, tail_split, tail_len, 0, 1, cand, pp);
  KERNEL_CHECK();
  if (tail_rows > 0) {
    hipLaunchKernelGGL(knn_merge_splits_kernel<H>,
                       dim3(blocks_for(static_cast<size_t>(2 * tail_rows), 256)), dim3(256), 0, s,
                       pp, full * C::kQueries, m, tail_split, cand);
    KERNEL_CHECK();
  }
}

// 64 candidates a query for 17 .. 32 neighbours from the 16-entry kernel,
// which runs two waves per SIMD where the 32-entry lists hold it to one: the
// reference rows are split two ways by interleaved tiles (so neighbours that
// sit together in the row order spread over both splits), every query gets
// four top-16 lists of disjoint rows, and the queries whose lists may have
// dropped one of their 32 best scores (knn_merge_check_kernel; a list holds
// a quarter of them on average) are redone with the 32-entry kernel.  The 64
// then hold the 32 best scores, as the two top-32 lists of the 32-entry
// kernel do.
template <typename T, int DPAD>
void launch_candidates_k32(const T* xq, int64_t m, const T* xr, const float* nr_half, int64_t n,
                           int32_t* cand, DevBuf<uint2>& pairs, hipStream_t s) {
  using C = CandShape<T, DPAD, 16>;
  auto kern = knn_candidates_kernel<T, DPAD, 16>;
  HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, C::kBytes));
  const int64_t groups = (m + C::kQueries - 1) / C::kQueries;
  pairs.ensure(static_cast<size_t>(m) * 64);
  hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups * 2)), dim3(C::kWaves * 64), C::kBytes,
                     s, xq, m, xr, nr_half, n, int64_t{0}, 2, int64_t{0}, 2, 0, cand, pairs.get());
  KERNEL_CHECK();
  DevBuf<int32_t> redo(m + 1);
  HIP_CHECK(hipMemsetAsync(redo.get() + m, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(knn_merge_check_kernel, dim3(blocks_for(static_cast<size_t>(m), 256)), dim3(256),
                     0, s, pairs.get(), m, cand, redo.get(), redo.get() + m);
  KERNEL_CHECK();
  int32_t f = 0;
  HIP_CHECK(hipMemcpyAsync(&f, redo.get() + m, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (f == 0) return;
  DevBuf<T> qsub(static_cast<size_t>(f) * DPAD);
  DevBuf<int32_t> csub(static_cast<size_t>(f) * 64);
  DevBuf<uint2> pairs2;
  const int vecs = DPAD * static_cast<int>(sizeof(T)) / 16;
  hipLaunchKernelGGL(gather_rows16_kernel<T>, dim3(blocks_for(static_cast<size_t>(f) * vecs, 256)),
                     dim3(256), 0, s, xq, redo.get(), static_cast<int64_t>(f), DPAD, qsub.get());
  KERNEL_CHECK();
  launch_candidates<T, DPAD, 32>(qsub.get(), f, xr, nr_half, n, csub.get(), pairs2, s);
  hipLaunchKernelGGL(scatter_cand64_kernel, dim3(blocks_for(static_cast<size_t>(f) * 64, 256)),
                     dim3(256), 0, s, csub.get(), redo.get(), static_cast<int64_t>(f), cand);
  KERNEL_CHECK();
}

// Query rows per block of the candidate kernel for one shape.
template <typename T, int H>
int queries_per_block(int dpad) {
  switch (dpad) {
    case 32: return CandShape<T, 32, H>::kQueries;
    case 64: return CandShape<T, 64, H>::kQueries;
    case 128: return CandShape<T, 128, H>::kQueries;
    case 192: return CandShape<T, 192, H>::kQueries;
    case 256: return CandShape<T, 256, H>::kQueries;
    default: throw_invalid("unsupported padded feature count");
  }
}

__global__ void zero_f64_kernel(double* __restrict__ p, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

template <typename T, int H>
void dispatch_dpad(int dpad, const T* xq, int64_t m, const T* xr, const float* nr_half, int64_t n,
                   int32_t* cand, DevBuf<uint2>& pairs, hipStream_t s) {
  switch (dpad) {
    case 32: launch_candidates<T, 32, H>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 64: launch_candidates<T, 64, H>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 128: launch_candidates<T, 128, H>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 192: launch_candidates<T, 192, H>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 256: launch_candidates<T, 256, H>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    default: throw_invalid("unsupported padded feature count");
  }
}

template <typename T>
void dispatch_dpad_k32(int dpad, const T* xq, int64_t m, const T* xr, const float* nr_half,
                       int64_t n, int32_t* cand, DevBuf<uint2>& pairs, hipStream_t s) {
  switch (dpad) {
    case 32: launch_candidates_k32<T, 32>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 64: launch_candidates_k32<T, 64>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 128: launch_candidates_k32<T, 128>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 192: launch_candidates_k32<T, 192>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 256: launch_candidates_k32<T, 256>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    default: throw_invalid("unsupported padded feature count");
  }
}

// BF16 rows of 257 .. 896 features: the candidate kernel with the queries'
// fragments held in registers (four waves, one per SIMD, so up to 512 of
// them), two one-tile groups of the reference in LDS.  The padded feature
// count rounds up to a multiple of 128 from 384: rows of whole 256-byte
// lines keep the 16-way chunk swizzle (a 200000 x 784 pass takes 130 ms at
// 896 features against 138 ms at 832; narrower shapes run eight waves and
// would spill); 0 past 896 (score blocks instead).
int wide_candidate_pad(int64_t d) {
  return d <= 384 ? 384 : (d <= 896 ? static_cast<int>((d + 127) / 128 * 128) : 0);
}

template <int H, int DPAD>
void launch_wide(const uint16_t* xq, int64_t m, const uint16_t* xr, const float* nr_half,
                 int64_t n, int32_t* cand, DevBuf<uint2>& pairs, hipStream_t s) {
  if constexpr (H == 32 && DPAD > 448)
    throw_invalid("32-entry lists take the score blocks past 448 features");
  else
    launch_candidates<uint16_t, DPAD, H>(xq, m, xr, nr_half, n, cand, pairs, s);
}

template <int H>
void dispatch_wide(int dpad, const uint16_t* xq, int64_t m, const uint16_t* xr,
                   const float* nr_half, int64_t n, int32_t* cand, DevBuf<uint2>& pairs,
                   hipStream_t s) {
  switch (dpad) {
    case 384: launch_wide<H, 384>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 512: launch_wide<H, 512>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 640: launch_wide<H, 640>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 768: launch_wide<H, 768>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    case 896: launch_wide<H, 896>(xq, m, xr, nr_half, n, cand, pairs, s); break;
    default: throw_invalid("unsupported padded feature count");
  }
}

}  // namespace

int knn_padded_dim(int64_t d) {
  for (int p : {32, 64, 128, 192, 256})
    if (d <= p) return p;
  return -1;
}

// Candidates per query: two halves of H each, H the smallest of {16, 32}
// holding both k and half of k * rerank_factor.
int knn_candidate_count(int k, int rerank_factor) {
  const int want = std::max(k, (k * std::max(rerank_factor, 1) + 1) / 2);
  if (want <= 16) return 32;
  if (want <= 32) return 64;
  return -1;
}

KnnScratchNeed knn_scratch_need(int64_t m, int64_t n, int64_t d, int k, int metric,
                                int precision, int rerank_factor, bool self) {
  KnnScratchNeed need;
  const int dpad = knn_padded_dim(d);
  const int kc = knn_candidate_count(k, rerank_factor);
  if (metric > 1 || dpad < 0 || kc < 0 || m == 0) return need;
  const size_t esz = precision == 1 ? 2 : 4;
  const int h = kc / 2;
  const int queries = precision == 1
                          ? (h == 16 ? queries_per_block<uint16_t, 16>(dpad)
                                     : queries_per_block<uint16_t, 32>(dpad))
                          : (h == 16 ? queries_per_block<float, 16>(dpad)
                                     : queries_per_block<float, 32>(dpad));
  const TailPlan tp = plan_tail(m, n, queries, device_cus());
  need.colsum = static_cast<size_t>(d);
  need.rprep = static_cast<size_t>(n) * dpad * esz;
  need.rnorm2 = static_cast<size_t>(knn_norm_rows(n));
  need.qprep = self ? 0 : static_cast<size_t>(m) * dpad * esz;
  need.cand = static_cast<size_t>(m) * kc;
  need.pairs = tp.tail_rows > 0 ? static_cast<size_t>(tp.tail_rows) * tp.tail_split * 2 * h : 0;
  need.rnorm = metric == 1 ? static_cast<size_t>(n) : 0;
  need.qnorm = metric == 1 && !self ? static_cast<size_t>(m) : 0;
  return need;
}

void knn_device(const float* q, int64_t m, const float* r, int64_t n, int64_t d, int k,
                int metric, int precision, int rerank_factor, int32_t* out_idx,
                float* out_dist, KnnWorkspace& ws, hipStream_t s, KnnStats* stats,
                float metric_p) {
  maybe_inject_fault("knn");
  if (m == 0) return;
  // other metrics, and more neighbours than the candidate lists hold (32):
  // the exact distances on the vector ALUs, ranked directly
  if (metric >= 2 || k > 32) {
    knn_metric_device(q, m, r, n, d, k, metric, metric_p, out_idx, out_dist, ws, s);
    return;
  }
  int dpad = knn_padded_dim(d);
  const bool wide = dpad < 0;  // past 256 features
  const int kc = knn_candidate_count(k, rerank_factor);
  // wide BF16 rows up to 896 features: the candidate kernel; otherwise GEMM
  // score blocks + selection
  // (32-entry lists only up to 448: past that they spill registers)
  const int wide_try = wide && precision == 1 && (kc == 32 || kc == 64) ? wide_candidate_pad(d) : 0;
  const int wide_pad = kc == 64 && wide_try > 448 ? 0 : wide_try;
  if (wide) {
    // the cosine column sums keep one f64 per feature in LDS
    if (d > 16384) throw_invalid("exact GPU kNN supports at most 16384 features");
    dpad = wide_pad > 0 ? wide_pad : static_cast<int>((d + 63) / 64 * 64);
  }
  if (kc < 0) throw_invalid("n_neighbors * rerank_factor must be <= 64 for n_neighbors <= 32");
  if (k > n) throw_invalid("n_neighbors must not exceed the number of reference rows");
  const bool self = (q == r) && (m == n);
  const bool bf16 = precision == 1;
  const size_t esz = bf16 ? 2 : 4;

  // column means of the (normalised) reference rows for centring
  ws.colsum.ensure(static_cast<size_t>(d));
  // a fill kernel rather than a memset node, so the call can be captured
  hipLaunchKernelGGL(zero_f64_kernel, dim3(blocks_for(static_cast<size_t>(d), 256)), dim3(256), 0,
                     s, ws.colsum.get(), d);
  const unsigned cs_blocks = static_cast<unsigned>(std::min<int64_t>(n, 1024));
  if (metric == 1) {
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(column_sums_normalized_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(sizeof(double) * d)));
    hipLaunchKernelGGL(column_sums_normalized_kernel, dim3(cs_blocks), dim3(256),
                       sizeof(double) * d, s, r, n, static_cast<int>(d), ws.colsum.get());
  } else {
    hipLaunchKernelGGL(column_sums_kernel, dim3(cs_blocks),
                       dim3(std::min<int64_t>(1024, std::max<int64_t>(64, ((d + 63) / 64) * 64))),
                       0, s, r, n, static_cast<int>(d), 0, ws.colsum.get());
  }
  KERNEL_CHECK();
  const double inv_n = 1.0 / static_cast<double>(n);

  ws.rprep.ensure(static_cast<size_t>(n) * dpad * esz);
  ws.rnorm2.ensure(static_cast<size_t>(knn_norm_rows(n)));
  if (!self) ws.qprep.ensure(static_cast<size_t>(m) * dpad * esz);
  auto prep = [&](const float* x, int64_t rows, void* out, float* nrm) {
    const int64_t norm_rows = nrm ? knn_norm_rows(rows) : rows;
    const unsigned blocks = blocks_for(static_cast<size_t>(norm_rows), 4);
    if (bf16)
      hipLaunchKernelGGL(prep_rows_kernel<uint16_t>, dim3(blocks), dim3(256), 0, s, x, rows,
                         norm_rows, static_cast<int>(d), dpad, metric == 1 ? 1 : 0,
                         ws.colsum.get(), inv_n, static_cast<uint16_t*>(out), nrm);
    else
      hipLaunchKernelGGL(prep_rows_kernel<float>, dim3(blocks), dim3(256), 0, s, x, rows,
                         norm_rows, static_cast<int>(d), dpad, metric == 1 ? 1 : 0,
                         ws.colsum.get(), inv_n, static_cast<float*>(out), nrm);
    KERNEL_CHECK();
  };
  prep(r, n, ws.rprep.get(), ws.rnorm2.get());
  const void* qp = ws.rprep.get();
  if (!self) {
    prep(q, m, ws.qprep.get(), nullptr);
    qp = ws.qprep.get();
  }

  ws.cand.ensure(static_cast<size_t>(m) * kc);
  if (stats) HIP_CHECK(hipEventRecord(stats->cand_start, s));
  if (wide_pad > 0) {
    if (kc == 32)
      dispatch_wide<16>(dpad, static_cast<const uint16_t*>(qp), m,
                        reinterpret_cast<const uint16_t*>(ws.rprep.get()), ws.rnorm2.get(), n,
                        ws.cand.get(), ws.pairs, s);
    else
      dispatch_wide<32>(dpad, static_cast<const uint16_t*>(qp), m,
                        reinterpret_cast<const uint16_t*>(ws.rprep.get()), ws.rnorm2.get(), n,
                        ws.cand.get(), ws.pairs, s);
  } else if (wide) {
    knn_wide_candidates(qp, m, ws.rprep.get(), ws.rnorm2.get(), r, n, d, dpad, bf16, kc,
                        ws.cand.get(), ws, s);
  } else if (kc == 64 && k > 16 && !ws.pairs.borrowed) {
    // 17 .. 32 neighbours through the 16-entry lists (the allocation-free
    // entry point, whose buffers are borrowed, keeps the 32-entry kernel)
    if (bf16)
      dispatch_dpad_k32<uint16_t>(dpad, static_cast<const uint16_t*>(qp), m,
                                  reinterpret_cast<const uint16_t*>(ws.rprep.get()),
                                  ws.rnorm2.get(), n, ws.cand.get(), ws.pairs, s);
    else
      dispatch_dpad_k32<float>(dpad, static_cast<const float*>(qp), m,
                               reinterpret_cast<const float*>(ws.rprep.get()), ws.rnorm2.get(),
                               n, ws.cand.get(), ws.pairs, s);
  } else if (bf16) {
    if (kc == 32)
      dispatch_dpad<uint16_t, 16>(dpad, static_cast<const uint16_t*>(qp), m,
                                  reinterpret_cast<const uint16_t*>(ws.rprep.get()),
                                  ws.rnorm2.get(), n, ws.cand.get(), ws.pairs, s);
    else
      dispatch_dpad<uint16_t, 32>(dpad, static_cast<const uint16_t*>(qp), m,
                                  reinterpret_cast<const uint16_t*>(ws.rprep.get()),
                                  ws.rnorm2.get(), n, ws.cand.get(), ws.pairs, s);
  } else {
    if (kc == 32)
      dispatch_dpad<float, 16>(dpad, static_cast<const float*>(qp), m,
                               reinterpret_cast<const float*>(ws.rprep.get()), ws.rnorm2.get(),
                               n, ws.cand.get(), ws.pairs, s);
    else
      dispatch_dpad<float, 32>(dpad, static_cast<const float*>(qp), m,
                               reinterpret_cast<const float*>(ws.rprep.get()), ws.rnorm2.get(),
                               n, ws.cand.get(), ws.pairs, s);
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
