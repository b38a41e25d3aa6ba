// SGD layout epochs on gfx950.
//
// The reference runs one Gauss-Seidel pass over the head-sorted edges per
// epoch on a single thread (optimize_layout_training_flat,
// rust_umap/src/lib.rs:5137-5318; densMAP loop :5324-5441).  Here each epoch is
// one edge-parallel launch (Hogwild): the edges' f32 schedules are checked and
// the firing edges compacted per wave; a thread per firing edge applies the
// attraction to its head (summed per head across the wave) and to its tail
// (float atomics, or for the fit's symmetric graph by the reverse edge's own
// thread), then draws its negatives from a counter-based hash of (seed,
// epoch, edge, draw) and accumulates the repulsion on a register copy of the
// head.  The per-edge
// schedule (epoch_of_next_sample, epoch_of_next_negative_sample) uses the
// reference's exact f32 operations, so every edge fires and draws negatives
// the same number of times as in the reference.
//
// The reference applies a head's moves one after another, each from the
// position the previous one left.  Within a wave the edge-parallel kernel does
// the same for each run of one head's edges (heads are sorted, so a run is a
// set of adjacent lanes that pass the head position along); runs in different
// waves start from the position the launch read and their net moves add.  A
// hub (a row hundreds of entries long, common in high-dimensional kNN graphs)
// spread over many waves would sum many such runs toward the same
// neighbourhood and overshoot it by that factor, so rows longer than kHubRow
// entries are cut into segments of kHubSeg edges, each walked by one block in
// CSR order with the head in registers (the reference's order on the head,
// lib.rs:5180-5241); a row's segments add their net moves like the waves of an
// ordinary row (edge lists that move tails with atomics walk a hub row whole).
// These walks run in the same launch as the edge-parallel work (their blocks
// first), so they overlap it.
#include "sgd.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <tuple>

namespace umapx {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Uniform index in [0, n) for draw s of stream key (Lemire's multiply-shift
// on the top 32 bits; n < 2^31).
__device__ __forceinline__ uint32_t draw_index(uint64_t key, uint32_t s, uint32_t n) {
  const uint64_t h = mix64(key + (static_cast<uint64_t>(s) + 1ULL) * 0x9e3779b97f4a7c15ULL);
  return static_cast<uint32_t>(((h >> 32) * static_cast<uint64_t>(n)) >> 32);
}

__device__ __forceinline__ float clip4(float v) { return fminf(fmaxf(v, -4.f), 4.f); }

// x^p for x > 0 from the hardware log2 / exp2 (about 1 ulp each); the SGD
// gradients are statistical, so the fast forms are used throughout.
__device__ __forceinline__ float pow_pos(float x, float p) {
  return __builtin_amdgcn_exp2f(p * __builtin_amdgcn_logf(x));
}
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Schedule step for one edge in one epoch.  Returns false when the edge does
// not fire; otherwise reports the number of negatives and updates state.
struct EdgeSchedule {
  __device__ static inline bool fire(float* __restrict__ eons, const float* __restrict__ eps,
                                     int64_t e, float ef) {
    const float cur = eons[e];
    if (cur > ef) return false;
    eons[e] = fadd(cur, eps[e]);
    return true;
  }
  __device__ static inline int negatives(float* __restrict__ eonns, const float* __restrict__ eps,
                                         int64_t e, float ef, float nsr) {
    const float en = fdiv(eps[e], nsr);
    if (!isfinite(en) || en <= 0.f) return -1;
    const float cur = eonns[e];
    const float fl = floorf(fdiv(fsub(ef, cur), en));
    const int nneg = fl > 0.f ? static_cast<int>(fl) : 0;
    eonns[e] = fadd(cur, fmul(static_cast<float>(nneg), en));
    return nneg;
  }
};

template <int DIM>
__global__ __launch_bounds__(256) void sgd_epoch_kernel(
    float* __restrict__ y, const int32_t* __restrict__ head, const int32_t* __restrict__ tail,
    const float* __restrict__ eps, float* __restrict__ eons, float* __restrict__ eonns,
    int64_t ne, uint32_t n, int dim_rt, float ef, float alpha, float a, float b, float rep,
    float nsr, uint64_t key_epoch, const float* __restrict__ radii, int dens_on, float dens_lambda,
    float var_shift, int32_t* __restrict__ pos_counts, int32_t* __restrict__ neg_counts) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  if (!EdgeSchedule::fire(eons, eps, e, ef)) return;
  const int dim = DIM > 0 ? DIM : dim_rt;
  constexpr int MAXD = DIM > 0 ? DIM : 8;
  const int32_t h = head[e];
  const int32_t t = tail[e];
  float yh[MAXD], yt[MAXD];
  for (int c = 0; c < dim; ++c) {
    yh[c] = y[static_cast<int64_t>(h) * dim + c];
    yt[c] = y[static_cast<int64_t>(t) * dim + c];
  }
  float dh[MAXD];
  for (int c = 0; c < dim; ++c) dh[c] = 0.f;
  if (h != t) {
    float d2 = 0.f;
    for (int c = 0; c < dim; ++c) {
      const float df = yh[c] - yt[c];
      d2 += df * df;
    }
    if (d2 > 0.f) {
      const float pb = pow_pos(d2, b);
      const float gc = (-2.f * a * b) * (pb / d2) / (a * pb + 1.f);
      for (int c = 0; c < dim; ++c) {
        const float g = clip4(gc * (yh[c] - yt[c])) * alpha;
        yh[c] += g;
        yt[c] -= g;
        dh[c] += g;
        atomicAdd(&y[static_cast<int64_t>(t) * dim + c], -g);
      }
    }
    if (dens_on) {
      // apply_density_gradient (lib.rs:5110-5134) on the updated pair
      float s2 = 0.f;
      for (int c = 0; c < dim; ++c) {
        const float df = yh[c] - yt[c];
        s2 += df * df;
      }
      const float cur = sqrtf(s2);
      const float target = fmaxf(expf((radii[h] + radii[t]) * 0.5f), 1e-6f);
      const float coeff = 0.1f * dens_lambda * (cur - target) / (cur + var_shift);
      for (int c = 0; c < dim; ++c) {
        const float dir = (yh[c] - yt[c]) / (cur + 1e-6f);
        const float g = clip4(coeff * dir) * alpha;
        yh[c] -= g;
        dh[c] -= g;
        atomicAdd(&y[static_cast<int64_t>(t) * dim + c], g);
      }
    }
  }
  if (pos_counts) pos_counts[e] += 1;
  const int nneg = EdgeSchedule::negatives(eonns, eps, e, ef, nsr);
  if (nneg > 0) {
    if (neg_counts) neg_counts[e] += nneg;
    const uint64_t key = mix64(key_epoch ^ static_cast<uint64_t>(e));
    for (int s = 0; s < nneg; ++s) {
      const uint32_t k = draw_index(key, static_cast<uint32_t>(s), n);
      if (static_cast<int32_t>(k) == h) continue;
      float d2 = 0.f;
      float yk[MAXD];
      for (int c = 0; c < dim; ++c) {
        yk[c] = y[static_cast<int64_t>(k) * dim + c];
        const float df = yh[c] - yk[c];
        d2 += df * df;
      }
      if (d2 > 0.f) {
        const float pb = pow_pos(d2, b);
        const float gc = rep / ((0.001f + d2) * (a * pb + 1.f));
        if (gc > 0.f)
          for (int c = 0; c < dim; ++c) {
            const float g = clip4(gc * (yh[c] - yk[c])) * alpha;
            yh[c] += g;
            dh[c] += g;
          }
      }
    }
  }
  for (int c = 0; c < dim; ++c)
    if (dh[c] != 0.f) atomicAdd(&y[static_cast<int64_t>(h) * dim + c], dh[c]);
}

// Attraction (lib.rs:5196-5212) and density term (lib.rs:5110-5134) of edge
// (ia -> ib) on register copies, accumulating both endpoint moves.
struct PullArgs {
  float a, b, alpha;
  const float* radii;
  int dens_on;
  float dens_lambda, var_shift;
};

__device__ __forceinline__ void pull2(const PullArgs& P, float2& ya, float2& yb, int32_t ia,
                                      int32_t ib, float2& da, float2& db) {
  const float dx = ya.x - yb.x, dy = ya.y - yb.y;
  const float d2 = dx * dx + dy * dy;
  if (d2 > 0.f) {
    const float pb = pow_pos(d2, P.b);
    const float gc = (-2.f * P.a * P.b) * pb * rcp(d2 * (P.a * pb + 1.f));
    const float gx = clip4(gc * dx) * P.alpha, gy = clip4(gc * dy) * P.alpha;
    ya.x += gx;
    ya.y += gy;
    yb.x -= gx;
    yb.y -= gy;
    da.x += gx;
    da.y += gy;
    db.x -= gx;
    db.y -= gy;
  }
  if (P.dens_on) {
    const float sx = ya.x - yb.x, sy = ya.y - yb.y;
    const float cur = sqrtf(sx * sx + sy * sy);
    const float target = fmaxf(expf((P.radii[ia] + P.radii[ib]) * 0.5f), 1e-6f);
    const float coeff = 0.1f * P.dens_lambda * (cur - target) / (cur + P.var_shift);
    const float inv = 1.f / (cur + 1e-6f);
    const float gx = clip4(coeff * sx * inv) * P.alpha, gy = clip4(coeff * sy * inv) * P.alpha;
    ya.x -= gx;
    ya.y -= gy;
    yb.x += gx;
    yb.y += gy;
    da.x -= gx;
    da.y -= gy;
    db.x += gx;
    db.y += gy;
  }
}

// Repulsion from one negative sample k (lib.rs:5214-5238) on the head.
__device__ __forceinline__ void push2(const PullArgs& P, float rep, float2& yh, float2 yk,
                                      float2& dh) {
  const float dx = yh.x - yk.x, dy = yh.y - yk.y;
  const float d2 = dx * dx + dy * dy;
  if (d2 > 0.f) {
    const float pb = pow_pos(d2, P.b);
    const float gc = rep * rcp((0.001f + d2) * (P.a * pb + 1.f));
    if (gc > 0.f) {
      const float gx = clip4(gc * dx) * P.alpha, gy = clip4(gc * dy) * P.alpha;
      yh.x += gx;
      yh.y += gy;
      dh.x += gx;
      dh.y += gy;
    }
  }
}

// Rows longer than this many entries are walked by blocks of their own (the hub
// walk below) instead of one thread per edge or per short chunk.
constexpr int kHubRow = 256;
constexpr int kHubSeg = 256;  // edges of a hub row walked by one block (the fit's union)
// (a walk lasts as long as its chain of dependent moves: one block per whole
// row of ~1200 entries set the epoch time at 200000 x 128, 26.0 ms of SGD per
// fit against 21.9 ms with 256-edge segments; 128- and 64-edge segments let
// hub rows overshoot again).  Edge lists that move tails with atomics (not
// the fit's union) walk each hub row whole: their hubs also take concurrent
// pulls as tails, and 256-edge segments on top of those threw C5's hubs out
// to radius ~90 in an embedding of radius ~16.
constexpr int32_t kWholeRow = std::numeric_limits<int32_t>::max();
constexpr int kHubThreads = 256;  // threads of a hub block (edges staged per pass)
constexpr int kHubNeg = 8;        // negatives gathered ahead per edge

// The hub walk: one block of kHubThreads threads takes head h's edges
// [r0, r1) in CSR order, kHubThreads edges per pass.  Each thread runs its
// edge's schedule, gathers its tail row and the rows of its negatives, and
// works out the move those negatives make, one after another in draw order,
// from the head's position at the start of the pass.  The firing edges are
// listed in LDS in edge order, and wave 0 then applies them one after another
// with the head in registers, each from the position the previous left:
// the edge's attraction (with MIRROR the pair's two moves, row min(h, t)
// first), then its negatives' move -- the reference's order on the head
// (lib.rs:5180-5241), with only the negatives' repulsion evaluated at the
// pass's starting position.  Without MIRROR a tail move is added to the tail
// with a float atomic by the edge's own thread.  The net head move is added
// atomically at the end (in MIRROR mode no other thread writes row h).
// Every thread of the block reaches every barrier.
template <bool MIRROR>
__device__ __forceinline__ void hub_block2(float2* __restrict__ y, int32_t h, int32_t r0,
                                           int32_t r1, const int32_t* __restrict__ tail,
                                           const float* __restrict__ eps,
                                           float* __restrict__ eons, float* __restrict__ eonns,
                                           uint32_t n, float ef, float nsr, float rep,
                                           uint64_t key_epoch, const PullArgs& P,
                                           int32_t* __restrict__ pos_counts,
                                           int32_t* __restrict__ neg_counts) {
  __shared__ int32_t s_t[kHubThreads];
  __shared__ float2 s_yt[kHubThreads], s_dn[kHubThreads], s_dt[kHubThreads];
  __shared__ int32_t s_wcnt[kHubThreads / 64];
  __shared__ float2 s_yh;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const float2 start = y[h];
  float2 yh = start;
  for (int32_t c = r0; c < r1; c += kHubThreads) {
    const int32_t e = c + tid;
    const bool in = e < r1;
    // every load that does not depend on another goes out at once: the edge's
    // schedule words and tail, then the tail row and the negatives' rows
    float eo = INFINITY, ep = 0.f, eon = 0.f;
    int32_t t = h;
    if (in) {
      eo = eons[e];
      ep = eps[e];
      eon = eonns[e];
      t = tail[e];
    }
    const bool fire = in && !(eo > ef);  // EdgeSchedule::fire
    const float2 yt = y[t];
    float2 dn = make_float2(0.f, 0.f);
    int nneg = 0;
    uint64_t key = 0;
    if (fire) {
      eons[e] = fadd(eo, ep);
      if (pos_counts) pos_counts[e] += 1;
      // EdgeSchedule::negatives on the words read above
      const float en = fdiv(ep, nsr);
      if (isfinite(en) && en > 0.f) {
        const float fl = floorf(fdiv(fsub(ef, eon), en));
        nneg = fl > 0.f ? static_cast<int>(fl) : 0;
        eonns[e] = fadd(eon, fmul(static_cast<float>(nneg), en));
      }
      if (nneg > 0 && neg_counts) neg_counts[e] += nneg;
      key = mix64(key_epoch ^ static_cast<uint64_t>(e));
    }
    if (nneg > 0) {
      float2 yl = yh;
      for (int s0 = 0; s0 < nneg; s0 += kHubNeg) {
        int32_t kk[kHubNeg];
        float2 yk[kHubNeg];
#pragma unroll
        for (int q = 0; q < kHubNeg; ++q) {
          kk[q] = s0 + q < nneg
                      ? static_cast<int32_t>(draw_index(key, static_cast<uint32_t>(s0 + q), n))
                      : h;
          yk[q] = y[kk[q]];
        }
#pragma unroll
        for (int q = 0; q < kHubNeg; ++q)
          if (kk[q] != h) push2(P, rep, yl, yk[q], dn);  // also skips the padding
      }
    }
    // list the firing edges in edge order: slot = firing edges before this one
    const uint64_t mask = __ballot(fire);
    if (lane == 0) s_wcnt[wv] = __popcll(mask);
    __syncthreads();
    int slot = __popcll(mask & ((uint64_t{1} << lane) - 1));
    int total = 0;
#pragma unroll
    for (int w = 0; w < kHubThreads / 64; ++w) {
      if (w < wv) slot += s_wcnt[w];
      total += s_wcnt[w];
    }
    if (fire) {
      s_t[slot] = t;
      s_yt[slot] = yt;
      s_dn[slot] = dn;
    }
    __syncthreads();
    if (wv == 0) {  // the edges' moves, one after another
      int32_t tn = total > 0 ? s_t[0] : h;
      float2 ytn = total > 0 ? s_yt[0] : yh;
      float2 dnn = total > 0 ? s_dn[0] : make_float2(0.f, 0.f);
      for (int j = 0; j < total; ++j) {
        const int32_t tl = tn;
        float2 ytl = ytn;
        const float2 dnl = dnn;
        if (j + 1 < total) {  // next entry read ahead of this one's arithmetic
          tn = s_t[j + 1];
          ytn = s_yt[j + 1];
          dnn = s_dn[j + 1];
        }
        float2 dh = make_float2(0.f, 0.f), dt = make_float2(0.f, 0.f);
        if (tl != h) {
          if (MIRROR) {
            if (h < tl) {
              pull2(P, yh, ytl, h, tl, dh, dt);
              pull2(P, ytl, yh, tl, h, dt, dh);
            } else {
              pull2(P, ytl, yh, tl, h, dt, dh);
              pull2(P, yh, ytl, h, tl, dh, dt);
            }
          } else {
            pull2(P, yh, ytl, h, tl, dh, dt);
          }
        }
        yh.x += dnl.x;
        yh.y += dnl.y;
        if (!MIRROR && lane == 0) s_dt[j] = dt;
      }
      if (lane == 0) s_yh = yh;
    }
    __syncthreads();
    yh = s_yh;
    if (!MIRROR && fire) {
      const float2 dt = s_dt[slot];
      if (dt.x != 0.f || dt.y != 0.f) {
        atomicAdd(&y[t].x, dt.x);
        atomicAdd(&y[t].y, dt.y);
      }
    }
    __syncthreads();  // the lists are rewritten by the next pass
  }
  if (tid == 0 && (yh.x != start.x || yh.y != start.y)) {
    atomicAdd(&y[h].x, yh.x - start.x);
    atomicAdd(&y[h].y, yh.y - start.y);
  }
}

// 2-D specialisation (lib.rs:5172-5245) with the firing edges compacted.  A
// wave owns kWaveEdges consecutive edges: it first runs their schedules (one
// edge per lane per pass) and lists the firing ones in LDS in edge order,
// then works through the list 64 at a time, so that the heavy part (the
// attraction and the negatives) runs on full waves although only a fraction
// of the edges fire in an epoch.  Heads are sorted, so equal heads sit in
// adjacent lanes in every round: head deltas are summed over equal-head lane
// runs and added with one atomic per run.
//
// Tail moves: with MIRROR (the fit's symmetric union, where the reverse edge
// (t, h) exists and fires in the same epochs, having the same weight and
// hence the same schedule) every edge replays the pair's two moves, row
// min(h, t) first and row max(h, t) second as in the reference's row order,
// from its own reads of both rows, and keeps only its head's share: the
// reverse edge's thread keeps the other.  No thread then writes a row it does
// not own.  Otherwise the tail move is added with float atomics.  A launch
// covers edges [e_begin, e_end).
//
// The first hub_blocks blocks walk hub-row segments instead (block b walks
// row hubs[2b] from edge hubs[2b + 1], longest first); hub_skip has a bit set
// for every edge of a hub row, which the edge-parallel waves leave to the
// walks.
constexpr int64_t kMirrorMinEdges = int64_t{1} << 20;
constexpr int kRangeLaunches = 8;
constexpr int kChunkLen = 4;  // longest run of one head's edges walked by one thread
constexpr int kPerLane = 2;
constexpr int kWaveEdges = 64 * kPerLane;
constexpr int kEpochWaves = 4;  // waves per block of the epoch kernel
static_assert(64 * kEpochWaves == kHubThreads, "hub walks run in epoch-kernel blocks");

template <bool MIRROR>
__global__ __launch_bounds__(64 * kEpochWaves) __attribute__((amdgpu_waves_per_eu(8)))
void sgd_epoch2_kernel(
    float2* __restrict__ y, const int32_t* __restrict__ head, const int32_t* __restrict__ tail,
    const float* __restrict__ eps, float* __restrict__ eons, float* __restrict__ eonns,
    uint32_t n, float ef, float alpha, float a, float b, float rep, float nsr, uint64_t key_epoch,
    const float* __restrict__ radii, int dens_on, float dens_lambda, float var_shift,
    int32_t* __restrict__ pos_counts, int32_t* __restrict__ neg_counts, int64_t e_begin,
    int64_t e_end, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ hubs,
    int n_hubs, int hub_blocks, const uint64_t* __restrict__ hub_skip) {
  __shared__ uint16_t list[kEpochWaves][kWaveEdges];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const PullArgs P{a, b, alpha, radii, dens_on, dens_lambda, var_shift};
  if (static_cast<int>(blockIdx.x) < hub_blocks) {  // block-uniform: a hub walk
    const int32_t h = hubs[2 * blockIdx.x], r0 = hubs[2 * blockIdx.x + 1];
    const int32_t r1 = MIRROR ? min(r0 + kHubSeg, row_ptr[h + 1]) : row_ptr[h + 1];
    hub_block2<MIRROR>(y, h, r0, r1, tail, eps, eons, eonns, n, ef, nsr, rep, key_epoch, P,
                       pos_counts, neg_counts);
    return;  // the edge-parallel blocks below have no block barriers
  }
  const int64_t wbase =
      e_begin +
      ((static_cast<int64_t>(blockIdx.x) - hub_blocks) * kEpochWaves + wv) * kWaveEdges;
  const uint64_t below = (uint64_t{1} << lane) - 1;
  int cnt = 0;
#pragma unroll
  for (int r = 0; r < kPerLane; ++r) {
    const int64_t e = wbase + r * 64 + lane;
    // e_begin and wbase are multiples of 64: one skip word per pass
    const uint64_t skip = hub_skip && wbase + r * 64 < e_end ? hub_skip[(wbase >> 6) + r] : 0;
    const bool fire =
        e < e_end && !((skip >> lane) & 1ULL) && EdgeSchedule::fire(eons, eps, e, ef);
    const uint64_t mask = __ballot(fire);
    if (fire) list[wv][cnt + __popcll(mask & below)] = static_cast<uint16_t>(r * 64 + lane);
    cnt += __popcll(mask);
  }
  // the list is written and read by lanes of one wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  for (int base = 0; base < cnt; base += 64) {
    const int j = base + lane;
    const bool valid = j < cnt;
    int32_t h = -1, t = -1;
    int64_t e = 0;
    float2 y0 = make_float2(0.f, 0.f), yt = y0, dn = y0;
    if (valid) {
      e = wbase + list[wv][j];
      h = head[e];
      t = tail[e];
      y0 = y[h];
      yt = y[t];
      if (pos_counts) pos_counts[e] += 1;
      // this edge's negatives' move, one after another in draw order, from
      // the head's position at the start of the round
      const int nneg = EdgeSchedule::negatives(eonns, eps, e, ef, nsr);
      if (nneg > 0) {
        if (neg_counts) neg_counts[e] += nneg;
        const uint64_t key = mix64(key_epoch ^ static_cast<uint64_t>(e));
        constexpr int kChunk = 8;
        float2 yl = y0;
        for (int s0 = 0; s0 < nneg; s0 += kChunk) {
          float2 yk[kChunk];
          int32_t kk[kChunk];
#pragma unroll
          for (int q = 0; q < kChunk; ++q) {
            kk[q] = (s0 + q < nneg) ? static_cast<int32_t>(draw_index(key, static_cast<uint32_t>(s0 + q), n)) : h;
            yk[q] = y[kk[q]];
          }
#pragma unroll
          for (int q = 0; q < kChunk; ++q) {
            if (kk[q] == h) continue;  // also the padding past nneg
            push2(P, rep, yl, yk[q], dn);
          }
        }
      }
    }
    // Heads are sorted, so one head's edges of this round sit in adjacent
    // lanes: pos is the edge's place in its head's run.  The run's moves are
    // applied one after another, each from the position the previous left
    // (the reference's order on the head): at step s the lanes at place s
    // move, then hand their head position to the next lane.
    const int32_t h_prev = __shfl_up(h, 1), h_next = __shfl_down(h, 1);  // all lanes take part
    const uint64_t starts = __ballot(valid && (lane == 0 || h_prev != h));
    const uint64_t upto = lane == 63 ? ~uint64_t{0} : ((uint64_t{1} << (lane + 1)) - 1);
    const int pos = lane - (63 - __clzll(starts & upto));
    int maxpos = valid ? pos : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) maxpos = max(maxpos, __shfl_xor(maxpos, o));
    float2 cur = y0, dt_sum = make_float2(0.f, 0.f);
    for (int step = 0; step <= maxpos; ++step) {
      if (valid && pos == step) {
        if (h != t) {
          float2 ytl = yt;
          float2 dh = make_float2(0.f, 0.f), dt = make_float2(0.f, 0.f);
          if (MIRROR) {
            // row min(h, t)'s pull first; the roles are swapped by selects,
            // not branches, so lanes of both orders share one instruction
            // stream
            const bool hf = h < t;
            float2 ya = hf ? cur : ytl, yb = hf ? ytl : cur;
            float2 da = dh, db = dt;
            pull2(P, ya, yb, hf ? h : t, hf ? t : h, da, db);
            pull2(P, yb, ya, hf ? t : h, hf ? h : t, db, da);
            cur = hf ? ya : yb;
          } else {
            pull2(P, cur, ytl, h, t, dh, dt);
            dt_sum = dt;
          }
        }
        cur.x += dn.x;
        cur.y += dn.y;
      }
      const float nx = __shfl_up(cur.x, 1), ny = __shfl_up(cur.y, 1);
      if (pos == step + 1) cur = make_float2(nx, ny);
    }
    if (!MIRROR && valid && (dt_sum.x != 0.f || dt_sum.y != 0.f)) {
      atomicAdd(&y[t].x, dt_sum.x);
      atomicAdd(&y[t].y, dt_sum.y);
    }
    // the last lane of each run adds the run's net head move
    const bool run_end = valid && (lane == 63 || h_next != h);
    if (run_end && (cur.x != y0.x || cur.y != y0.y)) {
      atomicAdd(&y[h].x, cur.x - y0.x);
      atomicAdd(&y[h].y, cur.y - y0.y);
    }
  }
}

// Vertex-parallel 2-D epoch: one thread per chunk of at most chunk_len
// consecutive edges of one head walks them in CSR order as the reference's
// loop does (lib.rs:5180-5241), keeping the head in registers so every later
// edge and negative of the chunk sees its earlier moves (Gauss-Seidel on the
// head within a chunk).  Rows are cut into chunks of chunk_len edges: a
// launch lasts as long as its longest chain of dependent gathers, so short
// chunks keep rows of ordinary length from serialising an epoch.  Tail moves
// are float atomics; the chunk's net head move is added atomically at the end
// so concurrent moves from other threads are kept.  Rows longer than kHubRow
// have no chunks: the first hub_blocks blocks walk their segments
// (hub_block2, block b walks row hubs[2b] from edge hubs[2b + 1]).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void sgd_vertex2_kernel(
    float2* __restrict__ y, const int32_t* __restrict__ chunk_row,
    const int32_t* __restrict__ chunk_start, const int32_t* __restrict__ row_ptr,
    const int32_t* __restrict__ tail, const float* __restrict__ eps, float* __restrict__ eons,
    float* __restrict__ eonns, int64_t c_begin, int64_t c_end, int chunk_len, uint32_t n,
    float ef, float alpha, float a, float b, float rep, float nsr, uint64_t key_epoch,
    const float* __restrict__ radii, int dens_on, float dens_lambda, float var_shift,
    int32_t* __restrict__ pos_counts, int32_t* __restrict__ neg_counts,
    const int32_t* __restrict__ hubs, int n_hubs, int hub_blocks) {
  const PullArgs P{a, b, alpha, radii, dens_on, dens_lambda, var_shift};
  if (static_cast<int>(blockIdx.x) < hub_blocks) {  // block-uniform: a hub walk
    const int32_t h = hubs[2 * blockIdx.x], r0 = hubs[2 * blockIdx.x + 1];
    hub_block2<false>(y, h, r0, row_ptr[h + 1], tail, eps, eons, eonns, n, ef, nsr, rep,
                      key_epoch, P, pos_counts, neg_counts);
    return;
  }
  const int64_t c = c_begin +
                    (static_cast<int64_t>(blockIdx.x) - hub_blocks) * blockDim.x + threadIdx.x;
  if (c >= c_end) return;
  const int32_t h = chunk_row[c];
  const int32_t e0 = chunk_start[c];
  const int32_t e1 = min(e0 + chunk_len, row_ptr[h + 1]);
  float2 yh = make_float2(0.f, 0.f), start = yh;
  bool loaded = false;
  for (int32_t e = e0; e < e1; ++e) {
    if (!EdgeSchedule::fire(eons, eps, e, ef)) continue;
    if (!loaded) {
      yh = y[h];
      start = yh;
      loaded = true;
    }
    const int32_t t = tail[e];
    if (t != h) {
      float2 yt = y[t];
      float2 dh = make_float2(0.f, 0.f), dt = make_float2(0.f, 0.f);
      pull2(P, yh, yt, h, t, dh, dt);
      if (dt.x != 0.f || dt.y != 0.f) {
        atomicAdd(&y[t].x, dt.x);
        atomicAdd(&y[t].y, dt.y);
      }
    }
    if (pos_counts) pos_counts[e] += 1;
    const int nneg = EdgeSchedule::negatives(eonns, eps, e, ef, nsr);
    if (nneg <= 0) continue;
    if (neg_counts) neg_counts[e] += nneg;
    const uint64_t key = mix64(key_epoch ^ static_cast<uint64_t>(e));
    constexpr int kChunk = 8;
    float2 dn = make_float2(0.f, 0.f);
    for (int s0 = 0; s0 < nneg; s0 += kChunk) {
      float2 yk[kChunk];
      int32_t kk[kChunk];
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        kk[j] = (s0 + j < nneg) ? static_cast<int32_t>(draw_index(key, static_cast<uint32_t>(s0 + j), n))
                                : static_cast<int32_t>(h);
        yk[j] = y[kk[j]];
      }
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        if (kk[j] == h) continue;  // also the padding past nneg
        push2(P, rep, yh, yk[j], dn);
      }
    }
  }
  if (loaded) {
    atomicAdd(&y[h].x, yh.x - start.x);
    atomicAdd(&y[h].y, yh.y - start.y);
  }
}

// One thread per query head: the transform SGD moves only the query point
// (lib.rs:5444-5544).  Edge schedule state lives in the workspace arrays.
template <int DIM>
__global__ __launch_bounds__(128) void transform_sgd_kernel(
    float* __restrict__ yq, int64_t m, int dim_rt, const float* __restrict__ yb, uint32_t n_base,
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ tail,
    const float* __restrict__ eps, float* __restrict__ eons, float* __restrict__ eonns,
    int n_epochs, float alpha0, float a, float b, float rep, float nsr, uint64_t seed) {
  const int64_t hq = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (hq >= m) return;
  const int dim = DIM > 0 ? DIM : dim_rt;
  constexpr int MAXD = DIM > 0 ? DIM : 8;
  float cur[MAXD];
  for (int c = 0; c < dim; ++c) cur[c] = yq[hq * dim + c];
  const int32_t s0 = row_ptr[hq], s1 = row_ptr[hq + 1];
  const uint64_t key_head = mix64(seed ^ (static_cast<uint64_t>(hq) * 0x9E3779B97F4A7C15ULL));
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const float alpha = fmul(alpha0, fsub(1.f, fdiv(static_cast<float>(epoch), static_cast<float>(n_epochs))));
    const float ef = static_cast<float>(epoch);
    for (int32_t e = s0; e < s1; ++e) {
      bool finite = true;
      for (int c = 0; c < dim; ++c) finite &= isfinite(cur[c]);
      if (eons[e] > ef || !finite) continue;
      const float* other = yb + static_cast<int64_t>(tail[e]) * dim;
      float d2 = 0.f;
      for (int c = 0; c < dim; ++c) {
        const float df = cur[c] - other[c];
        d2 += df * df;
      }
      float gc = 0.f;
      if (d2 > 0.f) {
        const float pb = pow_pos(d2, b);
        gc = (-2.f * a * b) * (pb / d2) / (a * pb + 1.f);
      }
      for (int c = 0; c < dim; ++c) cur[c] += clip4(gc * (cur[c] - other[c])) * alpha;
      eons[e] = fadd(eons[e], eps[e]);
      const int nneg = EdgeSchedule::negatives(eonns, eps, e, ef, nsr);
      if (nneg <= 0) continue;
      const uint64_t key = mix64(key_head + static_cast<uint64_t>(epoch) * 0xD1B54A32D192ED03ULL +
                                 static_cast<uint64_t>(e - s0));
      for (int s = 0; s < nneg; ++s) {
        const uint32_t k = draw_index(key, static_cast<uint32_t>(s), n_base);
        const float* o2 = yb + static_cast<int64_t>(k) * dim;
        float dd = 0.f;
        for (int c = 0; c < dim; ++c) {
          const float df = cur[c] - o2[c];
          dd += df * df;
        }
        float g2 = 0.f;
        if (dd > 0.f) {
          const float pb = pow_pos(dd, b);
          g2 = rep / ((0.001f + dd) * (a * pb + 1.f));
        }
        if (g2 > 0.f)
          for (int c = 0; c < dim; ++c) cur[c] += clip4(g2 * (cur[c] - o2[c])) * alpha;
      }
    }
  }
  for (int c = 0; c < dim; ++c) yq[hq * dim + c] = cur[c];
}

// The same walk for 2-D embeddings with each query's edges (at most 32, the
// kNN width) held in LDS: tail rows, eps and both schedule words are read
// once, and the rows of an edge's negatives are gathered together before the
// moves are applied one after another.  The arithmetic, the draws and their
// order are those of transform_sgd_kernel<2>, so the results are identical;
// only the memory round trips per edge fall from 2 + nneg to about 1.
constexpr int kXfEdges = 32;
constexpr int kXfNegBatch = 8;

__global__ __launch_bounds__(64) void transform_sgd2_kernel(
    float2* __restrict__ yq, int64_t m, const float2* __restrict__ yb, uint32_t n_base,
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ tail,
    const float* __restrict__ eps, int n_epochs, float alpha0, float a, float b, float rep,
    float nsr, uint64_t seed) {
  __shared__ float2 trow[kXfEdges][64];
  __shared__ float s_eps[kXfEdges][64], s_eons[kXfEdges][64], s_eonns[kXfEdges][64];
  const int lane = threadIdx.x;
  const int64_t hq = static_cast<int64_t>(blockIdx.x) * 64 + lane;
  if (hq >= m) return;  // no block barriers below
  const int32_t s0 = row_ptr[hq];
  const int cnt = row_ptr[hq + 1] - s0;  // <= kXfEdges, checked by the host
  for (int j = 0; j < cnt; ++j) {
    trow[j][lane] = yb[tail[s0 + j]];
    const float e = eps[s0 + j];
    s_eps[j][lane] = e;
    s_eons[j][lane] = e;
    s_eonns[j][lane] = fdiv(e, nsr);
  }
  float2 cur = yq[hq];
  const uint64_t key_head = mix64(seed ^ (static_cast<uint64_t>(hq) * 0x9E3779B97F4A7C15ULL));
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const float alpha = fmul(alpha0, fsub(1.f, fdiv(static_cast<float>(epoch), static_cast<float>(n_epochs))));
    const float ef = static_cast<float>(epoch);
    for (int j = 0; j < cnt; ++j) {
      const bool finite = isfinite(cur.x) && isfinite(cur.y);
      const float eo = s_eons[j][lane];
      if (eo > ef || !finite) continue;
      const float2 other = trow[j][lane];
      const float dx = cur.x - other.x, dy = cur.y - other.y;
      float d2 = 0.f;  // accumulated as the generic kernel's loop does
      d2 += dx * dx;
      d2 += dy * dy;
      float gc = 0.f;
      if (d2 > 0.f) {
        const float pb = pow_pos(d2, b);
        gc = (-2.f * a * b) * (pb / d2) / (a * pb + 1.f);
      }
      cur.x += clip4(gc * (cur.x - other.x)) * alpha;
      cur.y += clip4(gc * (cur.y - other.y)) * alpha;
      const float ej = s_eps[j][lane];
      s_eons[j][lane] = fadd(eo, ej);
      const float en = fdiv(ej, nsr);
      if (!isfinite(en) || en <= 0.f) continue;
      const float eonn = s_eonns[j][lane];
      const float fl = floorf(fdiv(fsub(ef, eonn), en));
      const int nneg = fl > 0.f ? static_cast<int>(fl) : 0;
      s_eonns[j][lane] = fadd(eonn, fmul(static_cast<float>(nneg), en));
      if (nneg <= 0) continue;
      const uint64_t key = mix64(key_head + static_cast<uint64_t>(epoch) * 0xD1B54A32D192ED03ULL +
                                 static_cast<uint64_t>(j));
      for (int sb = 0; sb < nneg; sb += kXfNegBatch) {
        float2 nv[kXfNegBatch];
#pragma unroll
        for (int u = 0; u < kXfNegBatch; ++u)
          nv[u] = sb + u < nneg ? yb[draw_index(key, static_cast<uint32_t>(sb + u), n_base)]
                                : make_float2(0.f, 0.f);
#pragma unroll
        for (int u = 0; u < kXfNegBatch; ++u) {
          if (sb + u >= nneg) break;
          const float ex = cur.x - nv[u].x, ey = cur.y - nv[u].y;
          float dd = 0.f;
          dd += ex * ex;
          dd += ey * ey;
          float g2 = 0.f;
          if (dd > 0.f) {
            const float pb = pow_pos(dd, b);
            g2 = rep / ((0.001f + dd) * (a * pb + 1.f));
          }
          if (g2 > 0.f) {
            cur.x += clip4(g2 * (cur.x - nv[u].x)) * alpha;
            cur.y += clip4(g2 * (cur.y - nv[u].y)) * alpha;
          }
        }
      }
    }
  }
  yq[hq] = cur;
}

// One bit per edge: set when the edge's head row is longer than kHubRow
// entries (left to the hub walk).  One thread per edge, 64-edge words.
__global__ void hub_skip_kernel(const int32_t* __restrict__ head,
                                const int32_t* __restrict__ row_ptr, int64_t ne,
                                uint64_t* __restrict__ skip) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool hub = false;
  if (e < ne) {
    const int32_t h = head[e];
    hub = row_ptr[h + 1] - row_ptr[h] > kHubRow;
  }
  const uint64_t m = __ballot(hub);
  if ((threadIdx.x & 63) == 0 && e < ne) skip[e >> 6] = m;
}

__global__ void copy_kernel(const float* __restrict__ src, float* __restrict__ dst, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void div_kernel(const float* __restrict__ eps, float* __restrict__ out, int64_t n,
                           float nsr) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = fdiv(eps[i], nsr);
}

// Replays every edge's firing schedule with the epoch kernels' own f32
// arithmetic (EdgeSchedule on per-thread copies of eons / eonns, started as
// optimize_layout_device starts them) without touching the embedding, and
// sums the positive updates and the negative samples over all epochs.
__global__ __launch_bounds__(256) void schedule_totals_kernel(const float* __restrict__ eps,
                                                              int64_t ne, int n_epochs, float nsr,
                                                              unsigned long long* __restrict__ totals) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long pos = 0, neg = 0;
  if (e < ne) {
    const float ep = eps[e];
    float eons = ep;
    float eonns = fdiv(ep, nsr);
    for (int t = 0; t < n_epochs; ++t) {
      const float ef = static_cast<float>(t);
      if (!EdgeSchedule::fire(&eons, &ep, 0, ef)) continue;
      ++pos;
      const int k = EdgeSchedule::negatives(&eonns, &ep, 0, ef, nsr);
      if (k > 0) neg += static_cast<unsigned long long>(k);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    pos += __shfl_xor(pos, o);
    neg += __shfl_xor(neg, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&totals[0], pos);
    atomicAdd(&totals[1], neg);
  }
}

// init_graph_transform (lib.rs:4793-4836), one thread per query row.
__global__ void init_transform_kernel(int64_t m, int dim, const int32_t* __restrict__ row_ptr,
                                      const int32_t* __restrict__ tail,
                                      const float* __restrict__ w, const float* __restrict__ yb,
                                      float* __restrict__ yq) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int32_t s0 = row_ptr[i], s1 = row_ptr[i + 1];
  float* o = yq + i * dim;
  if (s0 == s1) {
    for (int c = 0; c < dim; ++c) o[c] = NAN;
    return;
  }
  for (int32_t e = s0; e < s1; ++e)
    if (fabsf(fsub(w[e], 1.f)) < 1.1920929e-7f) {
      for (int c = 0; c < dim; ++c) o[c] = yb[static_cast<int64_t>(tail[e]) * dim + c];
      return;
    }
  float rs = 0.f;
  for (int32_t e = s0; e < s1; ++e) rs = fadd(rs, w[e]);
  if (rs <= 0.f) {
    for (int c = 0; c < dim; ++c) o[c] = NAN;
    return;
  }
  for (int c = 0; c < dim; ++c) o[c] = 0.f;
  for (int32_t e = s0; e < s1; ++e) {
    const float ww = fdiv(w[e], rs);
    for (int c = 0; c < dim; ++c)
      o[c] = fadd(o[c], fmul(ww, yb[static_cast<int64_t>(tail[e]) * dim + c]));
  }
}

// Chunks per row; rows longer than kHubRow have none (the hub walk takes them).
__global__ void chunk_count_kernel(const int32_t* __restrict__ row_ptr, int64_t n, int len,
                                   int32_t* __restrict__ cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t r = row_ptr[i + 1] - row_ptr[i];
  cnt[i] = r > kHubRow ? 0 : (r + len - 1) / len;
}

__global__ void chunk_fill_kernel(const int32_t* __restrict__ row_ptr, int64_t n, int len,
                                  const int32_t* __restrict__ off, int32_t* __restrict__ chunk_row,
                                  int32_t* __restrict__ chunk_start) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (row_ptr[i + 1] - row_ptr[i] > kHubRow) return;
  int32_t o = off[i];
  for (int32_t e = row_ptr[i]; e < row_ptr[i + 1]; e += len, ++o) {
    chunk_row[o] = static_cast<int32_t>(i);
    chunk_start[o] = e;
  }
}

// Work list of (head, first edge) chunks of at most len edges each.
int64_t build_chunks(const Graph& g, int len, SgdWorkspace& ws, hipStream_t s) {
  const int64_t n = g.n;
  DevBuf<int32_t> cnt(n), off(n);
  hipLaunchKernelGGL(chunk_count_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, g.row_ptr.get(), n,
                     len, cnt.get());
  KERNEL_CHECK();
  ScanWorkspace sws;
  const int64_t total = exclusive_scan(cnt.get(), off.get(), n, sws, s);
  ws.chunk_row.ensure(std::max<int64_t>(total, 1));
  ws.chunk_start.ensure(std::max<int64_t>(total, 1));
  hipLaunchKernelGGL(chunk_fill_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, g.row_ptr.get(), n,
                     len, off.get(), ws.chunk_row.get(), ws.chunk_start.get());
  KERNEL_CHECK();
  return total;
}

// Segments of the rows longer than kHubRow as (row, first edge) pairs,
// grouped by the launch of an epoch that walks them and longest first within
// a group.  group_of maps (row, first edge, chunks before the row) to a launch
// index; pairs offsets[g] .. offsets[g + 1] of the returned list are group
// g's.
template <typename GroupOf>
std::vector<int32_t> hub_rows_by_group(const std::vector<int32_t>& rp, int groups, int32_t seg,
                                       GroupOf group_of, std::vector<int>& offsets) {
  const int64_t n = static_cast<int64_t>(rp.size()) - 1;
  std::vector<std::vector<std::tuple<int32_t, int32_t, int32_t>>> by(groups);
  int64_t chunks_before = 0;
  for (int64_t v = 0; v < n; ++v) {
    const int32_t len = rp[v + 1] - rp[v];
    if (len > kHubRow) {
      const int gi = group_of(v, rp[v], chunks_before);
      for (int32_t s0 = 0; s0 < len; s0 += seg)
        by[gi].push_back({std::min(seg, len - s0), static_cast<int32_t>(v), rp[v] + s0});
    } else {
      chunks_before += (len + kChunkLen - 1) / kChunkLen;
    }
  }
  std::vector<int32_t> out;
  offsets.assign(groups + 1, 0);
  for (int gi = 0; gi < groups; ++gi) {
    auto& b = by[gi];
    std::stable_sort(b.begin(), b.end(),
                     [](const auto& l, const auto& r) { return std::get<0>(l) > std::get<0>(r); });
    offsets[gi] = static_cast<int>(out.size() / 2);
    for (const auto& seg : b) {
      out.push_back(std::get<1>(seg));
      out.push_back(std::get<2>(seg));
    }
  }
  offsets[groups] = static_cast<int>(out.size() / 2);
  return out;
}

uint64_t host_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

}  // namespace

void optimize_layout_device(float* y, int64_t n, int dim, const Graph& g, const SgdArgs& args,
                            SgdWorkspace& ws, hipStream_t s, SgdStats* stats) {
  maybe_inject_fault("optimize_layout");
  const int64_t ne = g.n_edges;
  if (ne == 0 || n == 0 || args.n_epochs == 0) return;
  if (dim > 8) throw_invalid("n_components must be <= 8 on the GPU path");
  ws.eps.ensure(ne);
  ws.eons.ensure(ne);
  ws.eonns.ensure(ne);
  epochs_per_sample_device(g.weight.get(), ne, g.max_weight, args.n_epochs, ws.eps.get(), s);
  const float nsr = static_cast<float>(args.negative_sample_rate);
  hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(ne, 256)), dim3(256), 0, s, ws.eps.get(),
                     ws.eons.get(), ne);
  hipLaunchKernelGGL(div_kernel, dim3(blocks_for(ne, 256)), dim3(256), 0, s, ws.eps.get(),
                     ws.eonns.get(), ne, nsr);
  KERNEL_CHECK();
  const float rep = 2.f * args.repulsion_strength * args.b;
  const unsigned blocks = blocks_for(ne, 256);
  // Large symmetric graphs (the fit's union): every edge replays its pair's
  // moves and keeps its head's share, one launch per epoch.  Smaller graphs,
  // where every vertex moves many times per epoch relative to its
  // neighbourhood: one thread per head walks its edges in order with the head
  // in registers (Gauss-Seidel on the head, as the reference's loop), tail
  // moves atomic.  Large edge lists that are not the fit's union: edge
  // parallel with atomic tail moves.  The last two split each epoch into
  // kRangeLaunches launches over consecutive heads, so later rows see the
  // moves of earlier rows as in the reference's row order.
  const bool large = ne >= kMirrorMinEdges;
  const bool mirror = dim == 2 && args.symmetric && large;
  const bool by_head = dim == 2 && !large;
  const unsigned epoch_blocks = blocks_for(ne, kEpochWaves * kWaveEdges);
  const int64_t n_chunks = by_head ? build_chunks(g, kChunkLen, ws, s) : 0;
  // hub rows (longer than kHubRow entries) for the 2-D walks, grouped by the
  // launch that walks them, plus the edge-parallel waves' skip bits
  std::vector<int> hub_off(kRangeLaunches + 1, 0);
  DevBuf<int32_t> hubs;
  DevBuf<uint64_t> skip;
  int64_t range_e[kRangeLaunches + 1];
  for (int r = 0; r <= kRangeLaunches; ++r)  // edge ranges of the atomic-tail launches
    range_e[r] = r == kRangeLaunches ? ne : (ne * r / kRangeLaunches) / kWaveEdges * kWaveEdges;
  int64_t range_c[kRangeLaunches + 1];
  for (int r = 0; r <= kRangeLaunches; ++r)  // chunk ranges of the vertex launches
    range_c[r] = n_chunks * r / kRangeLaunches;
  if (dim == 2) {
    std::vector<int32_t> rp(static_cast<size_t>(n) + 1);
    HIP_CHECK(hipMemcpyAsync(rp.data(), g.row_ptr.get(), sizeof(int32_t) * (n + 1),
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<int32_t> list;
    if (mirror) {
      list = hub_rows_by_group(rp, 1, kHubSeg, [](int64_t, int32_t, int64_t) { return 0; }, hub_off);
      hub_off.resize(kRangeLaunches + 1, hub_off.back());
    } else if (by_head) {
      list = hub_rows_by_group(
          rp, kRangeLaunches, kWholeRow,
          [&](int64_t, int32_t, int64_t cb) {
            int r = 0;
            while (r + 1 < kRangeLaunches && cb >= range_c[r + 1]) ++r;
            return r;
          },
          hub_off);
    } else {
      list = hub_rows_by_group(
          rp, kRangeLaunches, kWholeRow,
          [&](int64_t, int32_t e_first, int64_t) {
            int r = 0;
            while (r + 1 < kRangeLaunches && e_first >= range_e[r + 1]) ++r;
            return r;
          },
          hub_off);
    }
    if (!list.empty()) {
      hubs.alloc(list.size());
      HIP_CHECK(hipMemcpyAsync(hubs.get(), list.data(), sizeof(int32_t) * list.size(),
                               hipMemcpyHostToDevice, s));
      if (!by_head) {
        const int64_t words = (ne + 63) / 64;
        skip.alloc(words);
        hipLaunchKernelGGL(hub_skip_kernel, dim3(blocks_for(words * 64, 256)), dim3(256), 0, s,
                           g.head.get(), g.row_ptr.get(), ne, skip.get());
        KERNEL_CHECK();
      }
      HIP_CHECK(hipStreamSynchronize(s));  // the host list is freed on return
    }
  }
  auto hub_launch = [&](int r, const int32_t*& list_ptr, int& count, int& blocks) {
    count = hub_off[r + 1] - hub_off[r];
    list_ptr = count > 0 ? hubs.get() + 2 * hub_off[r] : nullptr;
    blocks = count;  // one block per hub-row segment
  };
  if (stats) HIP_CHECK(hipEventRecord(stats->start, s));
  for (int epoch = 0; epoch < args.n_epochs; ++epoch) {
    const float alpha = args.learning_rate *
                        (1.f - static_cast<float>(epoch) / static_cast<float>(args.n_epochs));
    const uint64_t key_epoch = host_mix64(args.seed + static_cast<uint64_t>(epoch) * 0xD1B54A32D192ED03ULL);
    const int dens_on = (args.radii != nullptr && epoch >= args.dens_start_epoch) ? 1 : 0;
    if (mirror) {
      const int32_t* hl;
      int hc, hb;
      hub_launch(0, hl, hc, hb);
      hipLaunchKernelGGL(sgd_epoch2_kernel<true>, dim3(hb + epoch_blocks), dim3(64 * kEpochWaves), 0,
                         s, reinterpret_cast<float2*>(y), g.head.get(), g.tail.get(), ws.eps.get(),
                         ws.eons.get(), ws.eonns.get(), static_cast<uint32_t>(n),
                         static_cast<float>(epoch), alpha, args.a, args.b, rep, nsr, key_epoch,
                         args.radii, dens_on, args.dens_lambda, args.dens_var_shift,
                         args.pos_counts, args.neg_counts, int64_t{0}, ne, g.row_ptr.get(), hl, hc,
                         hb, skip.get());
    } else if (by_head) {
      for (int r = 0; r < kRangeLaunches; ++r) {
        const int64_t c0 = range_c[r], c1 = range_c[r + 1];
        const int32_t* hl;
        int hc, hb;
        hub_launch(r, hl, hc, hb);
        const unsigned chunk_blocks = c1 > c0 ? blocks_for(c1 - c0, 256) : 0;
        if (chunk_blocks + hb == 0) continue;
        hipLaunchKernelGGL(sgd_vertex2_kernel, dim3(hb + chunk_blocks), dim3(256), 0, s,
                           reinterpret_cast<float2*>(y), ws.chunk_row.get(), ws.chunk_start.get(),
                           g.row_ptr.get(), g.tail.get(), ws.eps.get(), ws.eons.get(),
                           ws.eonns.get(), c0, c1, kChunkLen, static_cast<uint32_t>(n),
                           static_cast<float>(epoch), alpha, args.a, args.b, rep, nsr, key_epoch,
                           args.radii, dens_on, args.dens_lambda, args.dens_var_shift,
                           args.pos_counts, args.neg_counts, hl, hc, hb);
      }
    } else if (dim == 2) {
      for (int r = 0; r < kRangeLaunches; ++r) {
        // range bounds on wave boundaries
        const int64_t e0 = range_e[r], e1 = range_e[r + 1];
        const int32_t* hl;
        int hc, hb;
        hub_launch(r, hl, hc, hb);
        const unsigned edge_blocks = e1 > e0 ? blocks_for(e1 - e0, kEpochWaves * kWaveEdges) : 0;
        if (edge_blocks + hb == 0) continue;
        hipLaunchKernelGGL(sgd_epoch2_kernel<false>, dim3(hb + edge_blocks),
                           dim3(64 * kEpochWaves), 0, s, reinterpret_cast<float2*>(y),
                           g.head.get(), g.tail.get(), ws.eps.get(), ws.eons.get(),
                           ws.eonns.get(), static_cast<uint32_t>(n), static_cast<float>(epoch),
                           alpha, args.a, args.b, rep, nsr, key_epoch, args.radii, dens_on,
                           args.dens_lambda, args.dens_var_shift, args.pos_counts,
                           args.neg_counts, e0, e1, g.row_ptr.get(), hl, hc, hb, skip.get());
      }
    } else {
      hipLaunchKernelGGL(sgd_epoch_kernel<0>, dim3(blocks), dim3(256), 0, s, y, g.head.get(),
                         g.tail.get(), ws.eps.get(), ws.eons.get(), ws.eonns.get(), ne,
                         static_cast<uint32_t>(n), dim, static_cast<float>(epoch), alpha, args.a,
                         args.b, rep, nsr, key_epoch, args.radii, dens_on, args.dens_lambda,
                         args.dens_var_shift, args.pos_counts, args.neg_counts);
    }
    KERNEL_CHECK();
  }
  if (stats) {
    HIP_CHECK(hipEventRecord(stats->stop, s));
    stats->recorded = true;
  }
}

void optimize_transform_device(float* yq, int64_t m, int dim, const float* y_base, int64_t n_base,
                               const Graph& g, const SgdArgs& args, SgdWorkspace& ws,
                               hipStream_t s) {
  maybe_inject_fault("optimize_layout");
  const int64_t ne = g.n_edges;
  if (ne == 0 || m == 0 || args.n_epochs == 0) return;
  if (dim > 8) throw_invalid("n_components must be <= 8 on the GPU path");
  ws.eps.ensure(ne);
  ws.eons.ensure(ne);
  ws.eonns.ensure(ne);
  epochs_per_sample_device(g.weight.get(), ne, g.max_weight, args.n_epochs, ws.eps.get(), s);
  const float nsr = static_cast<float>(args.negative_sample_rate);
  hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(ne, 256)), dim3(256), 0, s, ws.eps.get(),
                     ws.eons.get(), ne);
  hipLaunchKernelGGL(div_kernel, dim3(blocks_for(ne, 256)), dim3(256), 0, s, ws.eps.get(),
                     ws.eonns.get(), ne, nsr);
  KERNEL_CHECK();
  const float rep = 2.f * args.repulsion_strength * args.b;
  if (dim == 2 && args.max_row_edges > 0 && args.max_row_edges <= kXfEdges)
    hipLaunchKernelGGL(transform_sgd2_kernel, dim3(blocks_for(m, 64)), dim3(64), 0, s,
                       reinterpret_cast<float2*>(yq), m, reinterpret_cast<const float2*>(y_base),
                       static_cast<uint32_t>(n_base), g.row_ptr.get(), g.tail.get(), ws.eps.get(),
                       args.n_epochs, args.learning_rate, args.a, args.b, rep, nsr, args.seed);
  else if (dim == 2)
    hipLaunchKernelGGL(transform_sgd_kernel<2>, dim3(blocks_for(m, 128)), dim3(128), 0, s, yq, m,
                       dim, y_base, static_cast<uint32_t>(n_base), g.row_ptr.get(), g.tail.get(),
                       ws.eps.get(), ws.eons.get(), ws.eonns.get(), args.n_epochs,
                       args.learning_rate, args.a, args.b, rep, nsr, args.seed);
  else
    hipLaunchKernelGGL(transform_sgd_kernel<0>, dim3(blocks_for(m, 128)), dim3(128), 0, s, yq, m,
                       dim, y_base, static_cast<uint32_t>(n_base), g.row_ptr.get(), g.tail.get(),
                       ws.eps.get(), ws.eons.get(), ws.eonns.get(), args.n_epochs,
                       args.learning_rate, args.a, args.b, rep, nsr, args.seed);
  KERNEL_CHECK();
}

void init_graph_transform_device(int64_t m, int dim, const Graph& g, const float* y_base,
                                 float* yq, hipStream_t s) {
  if (m == 0) return;
  hipLaunchKernelGGL(init_transform_kernel, dim3(blocks_for(m, 128)), dim3(128), 0, s, m, dim,
                     g.row_ptr.get(), g.tail.get(), g.weight.get(), y_base, yq);
  KERNEL_CHECK();
}

void schedule_totals_device(const SgdWorkspace& ws, int64_t ne, int n_epochs,
                            int negative_sample_rate, hipStream_t s, int64_t* positive_updates,
                            int64_t* negative_samples) {
  *positive_updates = 0;
  *negative_samples = 0;
  if (ne == 0 || n_epochs == 0) return;
  DevBuf<unsigned long long> tot(2);
  HIP_CHECK(hipMemsetAsync(tot.get(), 0, 2 * sizeof(unsigned long long), s));
  hipLaunchKernelGGL(schedule_totals_kernel, dim3(blocks_for(ne, 256)), dim3(256), 0, s,
                     ws.eps.get(), ne, n_epochs, static_cast<float>(negative_sample_rate), tot.get());
  KERNEL_CHECK();
  unsigned long long h[2];
  HIP_CHECK(hipMemcpyAsync(h, tot.get(), sizeof(h), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  *positive_updates = static_cast<int64_t>(h[0]);
  *negative_samples = static_cast<int64_t>(h[1]);
}

}  // namespace umapx
