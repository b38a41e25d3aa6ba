/*
 * TEST INFRASTRUCTURE ONLY — the oracle.  Never linked into the product.
 *
 * CPU restatement of umapers 1.1.4's dense UMAP fit and plain transform
 * (rust_umap/src/lib.rs).  The reference is Rust and cannot be compiled in
 * this image (no cargo/rustc, crates not vendored), so this restatement keeps
 * the reference's arithmetic instead:
 *   - f32 direct-difference distances, 8-lane partial sums for d >= 64 summed
 *     in lane order then the scalar tail (lib.rs:2445-2492), no FMA contraction
 *     (built with -ffp-contract=off);
 *   - top-k by (distance total order, then index) (lib.rs:2768-2808);
 *   - f32 smooth_knn_dist bisection (lib.rs:3424-3523);
 *   - sort-based fuzzy union (lib.rs:3597-3649) and prune (lib.rs:3804-3822);
 *   - f64 spectral subspace iteration with twice-iterated modified
 *     Gram-Schmidt (lib.rs:4161-4326) and the multi-component layout
 *     (lib.rs:4463-4764);
 *   - the sequential SGD with its exact f32 per-edge schedule
 *     (lib.rs:4937-4960, 5137-5441) and the transform SGD (lib.rs:5444-5544);
 *   - the inverse transform: kNN in the embedding, the local linear
 *     initialisation and the feature-space SGD (lib.rs:1800-1915, 4838-4935,
 *     5547-5645).
 * The stages the reference runs under rayon (exact kNN rows) use OpenMP; the
 * sequential ones stay single-threaded.
 *
 * Parity status (what pins each part):
 *   find_ab_params      bit-exact KAT (lib.rs:5781-5795)          pinned
 *   distances / top-k   FP64 brute force + tie-break KAT           pinned
 *   smooth_knn_dist     psum invariant, self/no-self KAT           pinned
 *   (f64 psum variant)  psum invariant in f64, closed-form rows    pinned
 *   membership / union  closed forms, symmetry, Laplacian KAT      pinned
 *   schedule counts     closed-form per-edge counts                pinned
 *   spectral init       eigh of I + D^-1/2 W D^-1/2 (subspace)     pinned up to sign/rotation
 *   SmallRng core       xoshiro256++ and SplitMix64 equal to the published
 *                       reference vectors; seed_from_u64 = 4 SplitMix64 outputs;
 *                       gen_range(int) = Lemire's widening multiply, checked
 *                       on the raw stream                      pinned
 *   gen_range(float)    the [1, 2) mantissa method restated from rand 0.8.5;
 *                       no reference output to check it against  parity unpinned
 *   Normal noise        Box-Muller, not rand_distr's ziggurat   parity unpinned
 *   symmetric eigen     cyclic Jacobi, not nalgebra 0.33.3      parity unpinned (signs)
 *   inverse init        small LU solve vs numpy.linalg.solve, fixed rows exact,
 *                       affine data reconstructed exactly      pinned (to f32 rounding)
 *   SGD embedding       statistical (trustworthiness, kNN preservation)
 */
#include "umap_oracle.h"

#include <immintrin.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

using std::size_t;
using std::vector;

constexpr float SMOOTH_K_TOLERANCE = 1e-5f;
constexpr float MIN_K_DIST_SCALE = 1e-3f;
constexpr float DEFAULT_BANDWIDTH = 1.0f;
constexpr float INIT_MAX_COORD = 10.0f;
constexpr float INIT_NOISE = 1e-4f;
constexpr size_t SPECTRAL_ITERATIVE_CONNECTED_THRESHOLD = 0;
constexpr size_t SPECTRAL_EXACT_LOW_DIMENSIONAL_MAX_FEATURES = 8;
constexpr size_t SPECTRAL_EXACT_LOW_DIMENSIONAL_MAX_SAMPLES = 2048;
constexpr size_t SPECTRAL_ITERATIVE_COMPONENT_THRESHOLD = 128;
constexpr size_t SPECTRAL_ITERATIVE_MAX_ITERS = 256;
constexpr size_t SPECTRAL_ITERATIVE_OVERSAMPLING = 6;
constexpr size_t SPECTRAL_DENSE_FALLBACK_MAX_SAMPLES = 4096;
constexpr double SPECTRAL_ORTHO_EPS = 1e-12;
constexpr float DENSITY_GRADIENT_SCALE = 0.1f;

double now_sec() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

/* ---------------------------------------------------------------- RNG ----
 * rand 0.8.5 SmallRng on 64-bit targets is xoshiro256++; seed_from_u64 fills
 * the state with SplitMix64 (published algorithms, Blackman & Vigna).
 * gen_range over integers is Lemire's widening multiply with the
 * "leading zeros" rejection zone; over floats it is the [1,2) mantissa trick
 * (rand 0.8.5 UniformInt / UniformFloat::sample_single). */
/* One step of SplitMix64 (Steele, Lea & Flood; Vigna's splitmix64.c). */
static inline uint64_t splitmix64_next(uint64_t& state) {
  state += 0x9e3779b97f4a7c15ULL;
  uint64_t z = state;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct SmallRng {
  uint64_t s[4];
  SmallRng() : s{0, 0, 0, 0} {}
  explicit SmallRng(uint64_t state) {
    for (int i = 0; i < 4; ++i) s[i] = splitmix64_next(state);
    if ((s[0] | s[1] | s[2] | s[3]) == 0) *this = SmallRng(0);
  }
  static SmallRng from_state(const uint64_t st[4]) {
    SmallRng r;
    for (int i = 0; i < 4; ++i) r.s[i] = st[i];
    return r;
  }
  static inline uint64_t rotl(uint64_t x, int k) {
    return (x << k) | (x >> (64 - k));
  }
  uint64_t next_u64() {
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
  uint32_t next_u32() { return static_cast<uint32_t>(next_u64() >> 32); }
  /* gen_range(0..n) for usize. */
  uint64_t gen_index(uint64_t n) {
    const uint64_t range = n;
    const uint64_t zone = (range << __builtin_clzll(range)) - 1;
    for (;;) {
      const uint64_t v = next_u64();
      const unsigned __int128 m = static_cast<unsigned __int128>(v) * range;
      const uint64_t hi = static_cast<uint64_t>(m >> 64);
      const uint64_t lo = static_cast<uint64_t>(m);
      if (lo <= zone) return hi;
    }
  }
  float gen_f32(float low, float high) {
    float scale = high - low;
    for (;;) {
      const uint32_t bits = (next_u32() >> 9) | 0x3F800000u;
      float v12;
      std::memcpy(&v12, &bits, 4);
      const float v01 = v12 - 1.0f;
      const float res = v01 * scale + low;
      if (res < high) return res;
      uint32_t sb;
      std::memcpy(&sb, &scale, 4);
      sb -= 1;
      std::memcpy(&scale, &sb, 4);
    }
  }
  double gen_f64(double low, double high) {
    double scale = high - low;
    for (;;) {
      const uint64_t bits = (next_u64() >> 12) | 0x3FF0000000000000ULL;
      double v12;
      std::memcpy(&v12, &bits, 8);
      const double v01 = v12 - 1.0;
      const double res = v01 * scale + low;
      if (res < high) return res;
      uint64_t sb;
      std::memcpy(&sb, &scale, 8);
      sb -= 1;
      std::memcpy(&scale, &sb, 8);
    }
  }
  /* N(0, sd) by Box-Muller on this stream (rand_distr uses a ziggurat; the
   * noise is 1e-4 and its exact stream is parity unpinned). */
  double normal(double sd) {
    double u1 = gen_f64(0.0, 1.0);
    const double u2 = gen_f64(0.0, 1.0);
    if (u1 <= 0.0) u1 = 1e-300;
    return sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

/* ---------------------------------------------------------- distances ----
 * lib.rs:2445-2492 squared_distance (scalar for d < 64, 8 AVX2 lanes summed
 * in lane order then scalar tail for d >= 64). */
inline float sq_dist(const float* x, const float* y, size_t d) {
  if (d >= 64) {
    __m256 acc = _mm256_setzero_ps();
    size_t i = 0;
    for (; i + 8 <= d; i += 8) {
      const __m256 diff = _mm256_sub_ps(_mm256_loadu_ps(x + i), _mm256_loadu_ps(y + i));
      acc = _mm256_add_ps(acc, _mm256_mul_ps(diff, diff));
    }
    alignas(32) float lanes[8];
    _mm256_store_ps(lanes, acc);
    float sum = 0.0f;
    for (int l = 0; l < 8; ++l) sum += lanes[l];
    for (; i < d; ++i) {
      const float df = x[i] - y[i];
      sum += df * df;
    }
    return sum;
  }
  float acc = 0.0f;
  for (size_t i = 0; i < d; ++i) {
    const float df = x[i] - y[i];
    acc += df * df;
  }
  return acc;
}

/* lib.rs:2603-2646 l2_norm */
inline float l2_norm(const float* x, size_t d) {
  if (d >= 64) {
    __m256 acc = _mm256_setzero_ps();
    size_t i = 0;
    for (; i + 8 <= d; i += 8) {
      const __m256 v = _mm256_loadu_ps(x + i);
      acc = _mm256_add_ps(acc, _mm256_mul_ps(v, v));
    }
    alignas(32) float lanes[8];
    _mm256_store_ps(lanes, acc);
    float sum = 0.0f;
    for (int l = 0; l < 8; ++l) sum += lanes[l];
    for (; i < d; ++i) sum += x[i] * x[i];
    return std::sqrt(sum);
  }
  float acc = 0.0f;
  for (size_t i = 0; i < d; ++i) acc += x[i] * x[i];
  return std::sqrt(acc);
}

/* lib.rs:2700-2765 dot product + cosine_distance_with_norms */
inline float dot(const float* x, const float* y, size_t d) {
  if (d >= 64) {
    __m256 acc = _mm256_setzero_ps();
    size_t i = 0;
    for (; i + 8 <= d; i += 8)
      acc = _mm256_add_ps(acc, _mm256_mul_ps(_mm256_loadu_ps(x + i), _mm256_loadu_ps(y + i)));
    alignas(32) float lanes[8];
    _mm256_store_ps(lanes, acc);
    float sum = 0.0f;
    for (int l = 0; l < 8; ++l) sum += lanes[l];
    for (; i < d; ++i) sum += x[i] * y[i];
    return sum;
  }
  float acc = 0.0f;
  for (size_t i = 0; i < d; ++i) acc += x[i] * y[i];
  return acc;
}

inline float cosine_with_norms(const float* x, const float* y, size_t d,
                               float nx, float ny) {
  const float dt = dot(x, y, d);
  if (nx == 0.0f && ny == 0.0f) return 0.0f;
  if (nx == 0.0f || ny == 0.0f) return 1.0f;
  float c = dt / (nx * ny);
  c = std::min(std::max(c, -1.0f), 1.0f);
  return 1.0f - c;
}

/* lib.rs:2494-2541 manhattan_distance: scalar for d < 64; for d >= 64 eight
 * AVX2 lane sums of |x - y|, added in lane order, then the scalar tail. */
inline float manhattan(const float* x, const float* y, size_t d) {
  if (d >= 64) {
    const __m256 sign = _mm256_set1_ps(-0.0f);
    __m256 acc = _mm256_setzero_ps();
    size_t i = 0;
    for (; i + 8 <= d; i += 8) {
      const __m256 diff = _mm256_sub_ps(_mm256_loadu_ps(x + i), _mm256_loadu_ps(y + i));
      acc = _mm256_add_ps(acc, _mm256_andnot_ps(sign, diff));
    }
    alignas(32) float lanes[8];
    _mm256_store_ps(lanes, acc);
    float sum = 0.0f;
    for (int l = 0; l < 8; ++l) sum += lanes[l];
    for (; i < d; ++i) sum += std::fabs(x[i] - y[i]);
    return sum;
  }
  float acc = 0.0f;
  for (size_t i = 0; i < d; ++i) acc += std::fabs(x[i] - y[i]);
  return acc;
}

/* lib.rs:2543-2552 chebyshev_distance (f32::max ignores NaN, as fmax does) */
inline float chebyshev(const float* x, const float* y, size_t d) {
  float m = 0.0f;
  for (size_t i = 0; i < d; ++i) m = std::fmax(m, std::fabs(x[i] - y[i]));
  return m;
}

/* lib.rs:2554-2563 minkowski_distance */
inline float minkowski(const float* x, const float* y, size_t d, float p) {
  float acc = 0.0f;
  for (size_t i = 0; i < d; ++i) acc += std::pow(std::fabs(x[i] - y[i]), p);
  return std::pow(acc, 1.0f / p);
}

/* lib.rs:2565-2577 canberra_distance */
inline float canberra(const float* x, const float* y, size_t d) {
  float acc = 0.0f;
  for (size_t i = 0; i < d; ++i) {
    const float den = std::fabs(x[i]) + std::fabs(y[i]);
    if (den > 0.0f) acc += std::fabs(x[i] - y[i]) / den;
  }
  return acc;
}

/* lib.rs:2579-2594 bray_curtis_distance */
inline float braycurtis(const float* x, const float* y, size_t d) {
  float num = 0.0f, den = 0.0f;
  for (size_t i = 0; i < d; ++i) {
    num += std::fabs(x[i] - y[i]);
    den += std::fabs(x[i] + y[i]);
  }
  return den == 0.0f ? 0.0f : num / den;
}

/* lib.rs:2655-2698 correlation_distance and
 * correlation_distance_from_summaries */
inline float correlation(const float* x, const float* y, size_t d) {
  float sx = 0.0f, sy = 0.0f, qx = 0.0f, qy = 0.0f, dt = 0.0f;
  for (size_t i = 0; i < d; ++i) {
    sx += x[i];
    sy += y[i];
    qx += x[i] * x[i];
    qy += y[i] * y[i];
    dt += x[i] * y[i];
  }
  const float nf = static_cast<float>(d);
  const float cdot = dt - (sx * sy / nf);
  const float xv = std::fmax(qx - (sx * sx / nf), 0.0f);
  const float yv = std::fmax(qy - (sy * sy / nf), 0.0f);
  const float xn = std::sqrt(xv), yn = std::sqrt(yv);
  if (xn == 0.0f && yn == 0.0f) return 0.0f;
  if (xn == 0.0f || yn == 0.0f) return 1.0f;
  const float c = std::min(std::max(cdot / (xn * yn), -1.0f), 1.0f);
  return 1.0f - c;
}

/* metric codes: 0 euclidean, 1 cosine, 2 manhattan, 3 chebyshev,
 * 4 minkowski (p), 5 correlation, 6 canberra, 7 braycurtis */
inline float metric_distance(int32_t metric, const float* x, const float* y, size_t d, float p) {
  switch (metric) {
    case 2: return manhattan(x, y, d);
    case 3: return chebyshev(x, y, d);
    case 4: return minkowski(x, y, d, p);
    case 5: return correlation(x, y, d);
    case 6: return canberra(x, y, d);
    default: return braycurtis(x, y, d);
  }
}

/* f32::total_cmp key */
inline int32_t total_key(float v) {
  int32_t b;
  std::memcpy(&b, &v, 4);
  return b ^ static_cast<int32_t>(static_cast<uint32_t>(b >> 31) >> 1);
}

/* lib.rs:2768-2772 neighbor_cmp: (dist total order, index) */
inline bool nb_less(int64_t li, float ld, int64_t ri, float rd) {
  const int32_t a = total_key(ld), b = total_key(rd);
  if (a != b) return a < b;
  return li < ri;
}

/* lib.rs:2775-2808 push_top_k_neighbor */
struct TopK {
  vector<int64_t> idx;
  vector<float> dist;
  size_t k;
  explicit TopK(size_t k_) : k(k_) {
    idx.reserve(k + 1);
    dist.reserve(k + 1);
  }
  inline void push(int64_t i, float d) {
    if (idx.size() == k && !nb_less(i, d, idx[k - 1], dist[k - 1])) return;
    size_t at = idx.size();
    while (at > 0 && nb_less(i, d, idx[at - 1], dist[at - 1])) --at;
    idx.insert(idx.begin() + at, i);
    dist.insert(dist.begin() + at, d);
    if (idx.size() > k) {
      idx.pop_back();
      dist.pop_back();
    }
  }
};

/* lib.rs:2863-2928, 3283-3399 exact kNN (rows under rayon -> OpenMP) */
void knn_impl(const float* q, int64_t m, const float* r, int64_t n, int64_t d,
              int32_t k, int32_t metric, int32_t n_threads, int64_t* out_idx,
              float* out_dist, float metric_p = 2.0f) {
  vector<float> qn, rn;
  if (metric == 1) {
    qn.resize(m);
    rn.resize(n);
    for (int64_t i = 0; i < m; ++i) qn[i] = l2_norm(q + i * d, d);
    for (int64_t j = 0; j < n; ++j) rn[j] = l2_norm(r + j * d, d);
  }
  const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
  for (int64_t i = 0; i < m; ++i) {
    TopK top(static_cast<size_t>(k));
    const float* x = q + i * d;
    if (metric == 1) {
      for (int64_t j = 0; j < n; ++j)
        top.push(j, cosine_with_norms(x, r + j * d, d, qn[i], rn[j]));
    } else if (metric >= 2) {
      /* lib.rs:2891-2909 exact_nearest_neighbors_by: the metric value itself */
      for (int64_t j = 0; j < n; ++j) top.push(j, metric_distance(metric, x, r + j * d, d, metric_p));
    } else {
      for (int64_t j = 0; j < n; ++j) top.push(j, sq_dist(x, r + j * d, d));
      for (auto& v : top.dist) v = std::sqrt(v); /* lib.rs:2829-2834 */
    }
    for (int32_t c = 0; c < k; ++c) {
      out_idx[i * k + c] = c < static_cast<int32_t>(top.idx.size()) ? top.idx[c] : -1;
      out_dist[i * k + c] = c < static_cast<int32_t>(top.dist.size())
                                ? top.dist[c]
                                : std::numeric_limits<float>::infinity();
    }
  }
}

/* lib.rs:3402-3417 */
bool rows_include_self(const int64_t* idx, const float* dist, int64_t n, int32_t k) {
  if (k <= 0) return false;
  for (int64_t i = 0; i < n; ++i)
    if (idx[i * k] != i || !(std::fabs(dist[i * k]) <= SMOOTH_K_TOLERANCE)) return false;
  return true;
}

/* lib.rs:3424-3523 smooth_knn_dist.  fp64_sums: the same 64-step bisection,
 * stop rule and f32 ratio (d - rho) / mid, with psum accumulated in f64 from
 * f64 exponentials and compared with the f64 target (the GPU path's default
 * form; the reference itself sums in f32). */
void smooth_knn(const float* distances, int64_t n, int32_t kc, float k,
                float local_connectivity, float bandwidth, bool include_self,
                float* sigmas, float* rhos, bool fp64_sums = false) {
  const float target = std::log2(k) * bandwidth;
  float mean_distances;
  {
    float total = 0.0f;
    size_t count = 0;
    for (int64_t i = 0; i < n; ++i) {
      const int start = (include_self && kc > 0) ? 1 : 0;
      for (int c = start; c < kc; ++c) {
        total += distances[i * kc + c];
        ++count;
      }
    }
    mean_distances = count > 0 ? total / static_cast<float>(count) : 0.0f;
  }
  vector<float> nz;
  for (int64_t i = 0; i < n; ++i) {
    const float* row = distances + i * kc;
    const int start = (include_self && kc > 0) ? 1 : 0;
    const float* act = row + start;
    const int na = kc - start;
    nz.clear();
    for (int c = 0; c < na; ++c)
      if (act[c] > 0.0f) nz.push_back(act[c]);
    float mean_ith = 0.0f;
    if (na > 0) {
      float s = 0.0f;
      for (int c = 0; c < na; ++c) s += act[c];
      mean_ith = s / static_cast<float>(na);
    }
    float rho = 0.0f;
    if (static_cast<float>(nz.size()) >= local_connectivity) {
      const size_t index = static_cast<size_t>(std::floor(local_connectivity));
      const float interpolation = local_connectivity - static_cast<float>(index);
      if (index > 0) {
        rho = nz[index - 1];
        if (interpolation > SMOOTH_K_TOLERANCE && index < nz.size())
          rho += interpolation * (nz[index] - nz[index - 1]);
      } else if (!nz.empty()) {
        rho = interpolation * nz[0];
      }
    } else if (!nz.empty()) {
      float mx = nz[0];
      for (size_t t = 1; t < nz.size(); ++t) mx = std::max(mx, nz[t]);
      rho = mx;
    }
    rhos[i] = rho;
    float lo = 0.0f, hi = std::numeric_limits<float>::infinity(), mid = 1.0f;
    for (int it = 0; it < 64; ++it) {
      bool stop, above;
      if (fp64_sums) {
        double psum = 0.0;
        for (int c = 0; c < na; ++c) {
          const float dd = act[c] - rho;
          psum += dd > 0.0f ? std::exp(-static_cast<double>(dd / mid)) : 1.0;
        }
        stop = std::fabs(psum - static_cast<double>(target)) < 1e-5;
        above = psum > static_cast<double>(target);
      } else {
        float psum = 0.0f;
        for (int c = 0; c < na; ++c) {
          const float dd = act[c] - rho;
          if (dd > 0.0f)
            psum += std::exp(-(dd / mid));
          else
            psum += 1.0f;
        }
        stop = std::fabs(psum - target) < SMOOTH_K_TOLERANCE;
        above = psum > target;
      }
      if (stop) break;
      if (above) {
        hi = mid;
        mid = (lo + hi) / 2.0f;
      } else {
        lo = mid;
        if (std::isinf(hi))
          mid *= 2.0f;
        else
          mid = (lo + hi) / 2.0f;
      }
    }
    const float min_scale = rho > 0.0f ? MIN_K_DIST_SCALE * mean_ith
                                       : MIN_K_DIST_SCALE * mean_distances;
    sigmas[i] = std::max(mid, min_scale);
  }
}

struct Edge {
  int64_t head, tail;
  float weight;
};

/* lib.rs:3525-3560 compute_membership_strengths */
vector<Edge> membership(const int64_t* idx, const float* dist, const float* sig,
                        const float* rho, int64_t n, int32_t k) {
  vector<Edge> out;
  out.reserve(static_cast<size_t>(n) * k);
  for (int64_t i = 0; i < n; ++i)
    for (int32_t j = 0; j < k; ++j) {
      const int64_t nb = idx[i * k + j];
      const float d = dist[i * k + j];
      float val;
      if (nb == i)
        val = 0.0f;
      else if (d - rho[i] <= 0.0f || sig[i] == 0.0f)
        val = 1.0f;
      else
        val = std::exp(-(d - rho[i]) / sig[i]);
      if (val > 0.0f) out.push_back({i, nb, val});
    }
  return out;
}

/* lib.rs:3562-3595 compute_membership_strengths_bipartite */
vector<Edge> membership_bipartite(const int64_t* idx, const float* dist,
                                  const float* sig, const float* rho, int64_t m,
                                  int32_t k) {
  vector<Edge> out;
  out.reserve(static_cast<size_t>(m) * k);
  for (int64_t i = 0; i < m; ++i)
    for (int32_t j = 0; j < k; ++j) {
      const float d = dist[i * k + j];
      float val;
      if (d - rho[i] <= 0.0f || sig[i] == 0.0f)
        val = 1.0f;
      else
        val = std::exp(-(d - rho[i]) / sig[i]);
      if (val > 0.0f) out.push_back({i, idx[i * k + j], val});
    }
  return out;
}

/* lib.rs:3597-3649 symmetrize_fuzzy_graph */
vector<Edge> symmetrize(const vector<Edge>& directed, float mix) {
  struct P {
    int64_t i, j;
    float a, b;
  };
  vector<P> by_pair;
  by_pair.reserve(directed.size());
  for (const auto& e : directed) {
    if (e.head == e.tail) continue;
    if (e.head < e.tail)
      by_pair.push_back({e.head, e.tail, e.weight, 0.0f});
    else
      by_pair.push_back({e.tail, e.head, 0.0f, e.weight});
  }
  std::stable_sort(by_pair.begin(), by_pair.end(), [](const P& l, const P& r) {
    return l.i != r.i ? l.i < r.i : l.j < r.j;
  });
  vector<Edge> edges;
  edges.reserve(by_pair.size() * 2);
  size_t p = 0;
  while (p < by_pair.size()) {
    const int64_t i = by_pair[p].i, j = by_pair[p].j;
    float wij = 0.0f, wji = 0.0f;
    while (p < by_pair.size() && by_pair[p].i == i && by_pair[p].j == j) {
      if (by_pair[p].a > 0.0f) wij = by_pair[p].a;
      if (by_pair[p].b > 0.0f) wji = by_pair[p].b;
      ++p;
    }
    const float prod = wij * wji;
    const float sym = mix * (wij + wji - prod) + (1.0f - mix) * prod;
    if (sym > 0.0f) {
      edges.push_back({i, j, sym});
      edges.push_back({j, i, sym});
    }
  }
  std::stable_sort(edges.begin(), edges.end(), [](const Edge& l, const Edge& r) {
    return l.head != r.head ? l.head < r.head : l.tail < r.tail;
  });
  return edges;
}

/* lib.rs:3804-3822 prune_edges */
void prune(vector<Edge>& edges, int32_t n_epochs) {
  if (edges.empty()) return;
  float mx = -std::numeric_limits<float>::infinity();
  for (const auto& e : edges) mx = std::max(mx, e.weight);
  const float denom = n_epochs > 10 ? static_cast<float>(n_epochs) : 500.0f;
  const float thr = mx / denom;
  vector<Edge> kept;
  kept.reserve(edges.size());
  for (const auto& e : edges)
    if (e.weight >= thr && e.weight > 0.0f) kept.push_back(e);
  edges.swap(kept);
}

/* lib.rs:3824-3839 random_init */
vector<float> random_init(int64_t n, int dim, uint64_t seed, float low, float high) {
  SmallRng rng(seed);
  vector<float> out(static_cast<size_t>(n) * dim);
  for (auto& v : out) v = rng.gen_f32(low, high);
  return out;
}

/* ---------------------------------------------------- symmetric eigen ----
 * Cyclic Jacobi (textbook).  nalgebra 0.33.3 SymmetricEigen is not available;
 * eigenvalues agree to rounding, eigenvector signs are parity unpinned. */
void jacobi_eigen(vector<double> a, size_t n, vector<double>& evals,
                  vector<double>& evecs) {
  evecs.assign(n * n, 0.0);
  for (size_t i = 0; i < n; ++i) evecs[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j < n; ++j) {
        tot += a[i * n + j] * a[i * n + j];
        if (i != j) off += a[i * n + j] * a[i * n + j];
      }
    if (off <= 1e-30 * (tot > 0 ? tot : 1.0) || off == 0.0) break;
    for (size_t p = 0; p < n; ++p)
      for (size_t q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (apq == 0.0) continue;
        const double app = a[p * n + p], aqq = a[q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (size_t r = 0; r < n; ++r) {
          const double arp = a[r * n + p], arq = a[r * n + q];
          a[r * n + p] = c * arp - s * arq;
          a[r * n + q] = s * arp + c * arq;
        }
        for (size_t r = 0; r < n; ++r) {
          const double apr = a[p * n + r], aqr = a[q * n + r];
          a[p * n + r] = c * apr - s * aqr;
          a[q * n + r] = s * apr + c * aqr;
        }
        for (size_t r = 0; r < n; ++r) {
          const double vrp = evecs[r * n + p], vrq = evecs[r * n + q];
          evecs[r * n + p] = c * vrp - s * vrq;
          evecs[r * n + q] = s * vrp + c * vrq;
        }
      }
  }
  evals.resize(n);
  for (size_t i = 0; i < n; ++i) evals[i] = a[i * n + i];
}

/* lib.rs:3885-3926 spectral_embedding_from_laplacian (eigenvalues ascending,
 * optionally dropping the first). */
bool embedding_from_laplacian(const vector<double>& lap, size_t n, int dim,
                              bool drop_first, vector<float>& coords) {
  if (n == 0) return false;
  vector<double> ev, vec;
  jacobi_eigen(lap, n, ev, vec);
  vector<size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t i, size_t j) { return ev[i] < ev[j]; });
  const size_t off = drop_first ? 1 : 0;
  if (order.size() <= static_cast<size_t>(dim) + off) return false;
  coords.assign(n * dim, 0.0f);
  for (int c = 0; c < dim; ++c) {
    const size_t col = order[c + off];
    for (size_t r = 0; r < n; ++r) coords[r * dim + c] = static_cast<float>(vec[r * n + col]);
  }
  for (float v : coords)
    if (!std::isfinite(v)) return false;
  return true;
}

/* lib.rs:3928-3964 normalized_laplacian_from_affinity */
bool laplacian_from_affinity(const vector<double>& aff, size_t n, vector<double>& lap) {
  if (n == 0) return false;
  vector<double> deg(n, 0.0);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) deg[i] += aff[i * n + j];
  bool any = false;
  for (double g : deg) any |= g > 0.0;
  if (!any) return false;
  vector<double> inv(n);
  for (size_t i = 0; i < n; ++i) inv[i] = deg[i] > 0.0 ? 1.0 / std::sqrt(deg[i]) : 0.0;
  lap.assign(n * n, 0.0);
  for (size_t i = 0; i < n; ++i) lap[i * n + i] = 1.0;
  for (size_t i = 0; i < n; ++i) {
    if (deg[i] <= 0.0) continue;
    for (size_t j = 0; j < n; ++j) {
      if (deg[j] <= 0.0) continue;
      const double w = aff[i * n + j];
      if (w > 0.0) lap[i * n + j] -= w * inv[i] * inv[j];
    }
  }
  return true;
}

/* lib.rs:3966-4023 normalized_laplacian_from_edges */
bool laplacian_from_edges(size_t n, const vector<Edge>& edges, vector<double>& lap) {
  if (n == 0 || edges.empty()) return false;
  lap.assign(n * n, 0.0);
  for (size_t i = 0; i < n; ++i) lap[i * n + i] = 1.0;
  vector<double> deg(n, 0.0);
  vector<std::pair<size_t, size_t>> pairs;
  for (const auto& e : edges) {
    if (e.head == e.tail) continue;
    const double w = static_cast<double>(std::max(e.weight, 0.0f));
    if (w <= 0.0) continue;
    const size_t i = static_cast<size_t>(std::min(e.head, e.tail));
    const size_t j = static_cast<size_t>(std::max(e.head, e.tail));
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
  vector<double> inv(n);
  for (size_t i = 0; i < n; ++i) inv[i] = deg[i] > 0.0 ? 1.0 / std::sqrt(deg[i]) : 0.0;
  for (const auto& pr : pairs) {
    const double v = lap[pr.first * n + pr.second] * inv[pr.first] * inv[pr.second];
    lap[pr.first * n + pr.second] = -v;
    lap[pr.second * n + pr.first] = -v;
  }
  for (size_t i = 0; i < n; ++i) lap[i * n + i] = 1.0;
  return true;
}

/* lib.rs:4105-4159 symmetrized_undirected_edges (max per pair, sorted) */
bool undirected_edges(size_t n, const vector<Edge>& edges,
                      vector<std::tuple<size_t, size_t, double>>& und,
                      vector<double>& deg) {
  if (n == 0 || edges.empty()) return false;
  std::map<std::pair<size_t, size_t>, double> by_pair;
  for (const auto& e : edges) {
    if (e.head == e.tail) continue;
    const double w = static_cast<double>(std::max(e.weight, 0.0f));
    if (w <= 0.0) continue;
    const size_t i = static_cast<size_t>(std::min(e.head, e.tail));
    const size_t j = static_cast<size_t>(std::max(e.head, e.tail));
    auto it = by_pair.find({i, j});
    if (it == by_pair.end())
      by_pair.emplace(std::make_pair(i, j), w);
    else if (w > it->second)
      it->second = w;
  }
  if (by_pair.empty()) return false;
  deg.assign(n, 0.0);
  und.clear();
  und.reserve(by_pair.size());
  for (const auto& kv : by_pair) {
    deg[kv.first.first] += kv.second;
    deg[kv.first.second] += kv.second;
    und.emplace_back(kv.first.first, kv.first.second, kv.second);
  }
  bool any = false;
  for (double g : deg) any |= g > 0.0;
  return any;
}

/* lib.rs:4161-4197 orthonormalize_columns (twice-iterated MGS) */
bool orthonormalize(vector<double>& blk, size_t rows, size_t cols) {
  if (rows == 0 || cols == 0) return false;
  for (size_t col = 0; col < cols; ++col) {
    for (int rep = 0; rep < 2; ++rep)
      for (size_t prev = 0; prev < col; ++prev) {
        double dt = 0.0;
        for (size_t r = 0; r < rows; ++r) dt += blk[r * cols + col] * blk[r * cols + prev];
        for (size_t r = 0; r < rows; ++r) blk[r * cols + col] -= dt * blk[r * cols + prev];
      }
    double ns = 0.0;
    for (size_t r = 0; r < rows; ++r) ns += blk[r * cols + col] * blk[r * cols + col];
    if (!std::isfinite(ns) || ns <= SPECTRAL_ORTHO_EPS) return false;
    const double inv = 1.0 / std::sqrt(ns);
    for (size_t r = 0; r < rows; ++r) blk[r * cols + col] *= inv;
  }
  return true;
}

/* lib.rs:4199-4225 shifted_normalized_adjacency_mul */
void shifted_mul(const vector<double>& blk, size_t cols,
                 const vector<std::tuple<size_t, size_t, double>>& und,
                 const vector<double>& inv, vector<double>& out) {
  out = blk;
  for (const auto& t : und) {
    const size_t i = std::get<0>(t), j = std::get<1>(t);
    const double nw = std::get<2>(t) * inv[i] * inv[j];
    if (nw == 0.0) continue;
    const double* xi = &blk[i * cols];
    const double* xj = &blk[j * cols];
    double* oi = &out[i * cols];
    double* oj = &out[j * cols];
    for (size_t c = 0; c < cols; ++c) {
      oi[c] += nw * xj[c];
      oj[c] += nw * xi[c];
    }
  }
}

/* lib.rs:4227-4326 spectral_embedding_from_edges_iterative */
bool spectral_iterative(size_t n, int dim, const vector<Edge>& edges, uint64_t seed,
                        vector<float>& coords) {
  const size_t sub = std::max(static_cast<size_t>(dim + 1) * SPECTRAL_ITERATIVE_OVERSAMPLING,
                              static_cast<size_t>(dim + 1));
  if (n <= sub || edges.empty()) return false;
  vector<std::tuple<size_t, size_t, double>> und;
  vector<double> deg;
  if (!undirected_edges(n, edges, und, deg)) return false;
  vector<double> inv(n);
  for (size_t i = 0; i < n; ++i) inv[i] = deg[i] > 0.0 ? 1.0 / std::sqrt(deg[i]) : 0.0;
  SmallRng rng(seed ^ 0x7A5B3D91CE42A117ULL);
  vector<double> blk(n * sub, 0.0);
  for (size_t r = 0; r < n; ++r) {
    blk[r * sub] = std::sqrt(deg[r]);
    for (size_t c = 1; c < sub; ++c) blk[r * sub + c] = rng.gen_f64(-1.0, 1.0);
  }
  if (!orthonormalize(blk, n, sub)) return false;
  vector<double> scratch;
  for (size_t it = 0; it < SPECTRAL_ITERATIVE_MAX_ITERS; ++it) {
    shifted_mul(blk, sub, und, inv, scratch);
    if (!orthonormalize(scratch, n, sub)) return false;
    std::swap(blk, scratch);
  }
  shifted_mul(blk, sub, und, inv, scratch);
  vector<double> ritz(sub * sub, 0.0);
  for (size_t i = 0; i < sub; ++i)
    for (size_t j = i; j < sub; ++j) {
      double v = 0.0;
      for (size_t r = 0; r < n; ++r) v += blk[r * sub + i] * scratch[r * sub + j];
      ritz[i * sub + j] = v;
      ritz[j * sub + i] = v;
    }
  vector<double> ev, vec;
  jacobi_eigen(ritz, sub, ev, vec);
  vector<size_t> order(sub);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t i, size_t j) { return ev[j] < ev[i]; });
  if (order.size() <= static_cast<size_t>(dim)) return false;
  coords.assign(n * dim, 0.0f);
  for (int c = 0; c < dim; ++c) {
    const size_t col = order[c + 1];
    for (size_t r = 0; r < n; ++r) {
      double v = 0.0;
      for (size_t bcol = 0; bcol < sub; ++bcol) v += blk[r * sub + bcol] * vec[bcol * sub + col];
      coords[r * dim + c] = static_cast<float>(v);
    }
  }
  for (float v : coords)
    if (!std::isfinite(v)) return false;
  return true;
}

/* lib.rs:4328-4353 spectral_embedding_from_edges */
bool spectral_from_edges(size_t n, int dim, const vector<Edge>& edges, uint64_t seed,
                         size_t iterative_threshold, vector<float>& coords) {
  if (n <= static_cast<size_t>(dim) + 1 || edges.empty()) return false;
  if (n >= iterative_threshold) {
    if (spectral_iterative(n, dim, edges, seed, coords)) return true;
    if (n > SPECTRAL_DENSE_FALLBACK_MAX_SAMPLES) return false;
  }
  vector<double> lap;
  if (!laplacian_from_edges(n, edges, lap)) return false;
  return embedding_from_laplacian(lap, n, dim, true, coords);
}

/* lib.rs:4766-4791 noisy_scale_coords */
void noisy_scale(vector<float>& coords, uint64_t seed, float max_coord, float noise) {
  if (coords.empty()) return;
  float mx = 0.0f;
  for (float v : coords) mx = std::max(mx, std::fabs(v));
  const float expansion = mx > 0.0f ? max_coord / mx : 1.0f;
  SmallRng rng(seed);
  for (auto& v : coords) v = v * expansion + static_cast<float>(rng.normal(noise));
}

/* lib.rs:4610-4622 add_noise */
void add_noise(vector<float>& coords, uint64_t seed, float noise) {
  if (coords.empty()) return;
  SmallRng rng(seed);
  for (auto& v : coords) v += static_cast<float>(rng.normal(noise));
}

/* lib.rs:4394-4437 connected_components_from_edges (components ordered by
 * smallest member, members ascending) */
vector<vector<int64_t>> components_of(int64_t n, const int64_t* heads,
                                      const int64_t* tails, int64_t ne) {
  vector<int64_t> parent(n);
  std::iota(parent.begin(), parent.end(), 0);
  vector<uint8_t> rank(n, 0);
  auto find = [&](int64_t x) {
    int64_t root = x;
    while (parent[root] != root) root = parent[root];
    while (parent[x] != root) {
      const int64_t nx = parent[x];
      parent[x] = root;
      x = nx;
    }
    return root;
  };
  for (int64_t e = 0; e < ne; ++e) {
    if (heads[e] == tails[e]) continue;
    const int64_t ra = find(heads[e]), rb = find(tails[e]);
    if (ra == rb) continue;
    if (rank[ra] < rank[rb])
      parent[ra] = rb;
    else if (rank[ra] > rank[rb])
      parent[rb] = ra;
    else {
      parent[rb] = ra;
      rank[ra] += 1;
    }
  }
  std::unordered_map<int64_t, size_t> slot;
  vector<vector<int64_t>> comps;
  for (int64_t v = 0; v < n; ++v) {
    const int64_t r = find(v);
    auto it = slot.find(r);
    if (it == slot.end()) {
      slot.emplace(r, comps.size());
      comps.push_back({v});
    } else {
      comps[it->second].push_back(v);
    }
  }
  /* members are pushed in ascending order, and components in order of their
   * first (smallest) member, matching sort_by_key(component[0]). */
  return comps;
}

/* lib.rs:4518-4550 meta_component_layout_from_centroids */
bool meta_layout(const vector<vector<double>>& cent, int dim, uint64_t seed,
                 vector<float>& layout) {
  const size_t nc = cent.size();
  if (nc == 0) return false;
  if (nc == 1) {
    layout.assign(dim, 0.0f);
    return true;
  }
  if (nc <= static_cast<size_t>(2 * dim)) {
    const size_t k = (nc + 1) / 2;
    layout.assign(nc * dim, 0.0f);
    for (size_t i = 0; i < nc; ++i) layout[i * dim + (i % k)] = i < k ? 1.0f : -1.0f;
    return true;
  }
  vector<double> aff(nc * nc, 0.0);
  for (size_t i = 0; i < nc; ++i) {
    aff[i * nc + i] = 1.0;
    for (size_t j = i + 1; j < nc; ++j) {
      double d2 = 0.0;
      for (size_t f = 0; f < cent[i].size(); ++f) {
        const double df = cent[i][f] - cent[j][f];
        d2 += df * df;
      }
      const double w = std::exp(-d2);
      aff[i * nc + j] = w;
      aff[j * nc + i] = w;
    }
  }
  vector<double> lap;
  if (!laplacian_from_affinity(aff, nc, lap)) return false;
  if (!embedding_from_laplacian(lap, nc, dim, true, layout)) return false;
  noisy_scale(layout, seed ^ 0xC85E7D6ADA318F21ULL, INIT_MAX_COORD, INIT_NOISE);
  for (float v : layout)
    if (!std::isfinite(v)) return false;
  return true;
}

/* lib.rs:4624-4764 multi_component_spectral_init(_with_layout) */
bool multi_component_init(const float* x, int64_t n, int64_t d, int dim,
                          const vector<Edge>& edges,
                          const vector<vector<int64_t>>& comps, uint64_t seed,
                          vector<float>& result) {
  vector<vector<double>> cent;
  for (const auto& c : comps) {
    vector<double> ce(d, 0.0);
    for (int64_t idx : c)
      for (int64_t f = 0; f < d; ++f) ce[f] += static_cast<double>(x[idx * d + f]);
    const double sc = 1.0 / static_cast<double>(c.size());
    for (auto& v : ce) v *= sc;
    cent.push_back(std::move(ce));
  }
  vector<float> layout;
  if (!meta_layout(cent, dim, seed ^ 0xA13F52A92D4CB801ULL, layout)) return false;
  vector<int64_t> label(n, -1), local(n, -1);
  for (size_t c = 0; c < comps.size(); ++c)
    for (size_t li = 0; li < comps[c].size(); ++li) {
      label[comps[c][li]] = static_cast<int64_t>(c);
      local[comps[c][li]] = static_cast<int64_t>(li);
    }
  result.assign(static_cast<size_t>(n) * dim, 0.0f);
  for (size_t cid = 0; cid < comps.size(); ++cid) {
    const auto& comp = comps[cid];
    const float* anchor = &layout[cid * dim];
    float range = std::numeric_limits<float>::infinity();
    for (size_t o = 0; o < comps.size(); ++o) {
      if (o == cid) continue;
      const float dd = std::sqrt(sq_dist(anchor, &layout[o * dim], dim));
      if (dd > 0.0f) range = std::min(range, dd / 2.0f);
    }
    if (!std::isfinite(range) || range <= 0.0f) range = 1.0f;
    vector<float> lc;
    const size_t cs = comp.size();
    if (cs < static_cast<size_t>(2 * dim) || cs <= static_cast<size_t>(dim) + 1) {
      lc = random_init(static_cast<int64_t>(cs), dim,
                       seed ^ ((static_cast<uint64_t>(cid) + 1) * 0x9E3779B9ULL), -range, range);
    } else {
      vector<Edge> ce;
      for (const auto& e : edges)
        if (label[e.head] == static_cast<int64_t>(cid) && label[e.tail] == static_cast<int64_t>(cid))
          ce.push_back({local[e.head], local[e.tail], e.weight});
      if (!spectral_from_edges(cs, dim, ce,
                               seed ^ ((static_cast<uint64_t>(cid) + 1) * 0xBF58476D1CE4E5B9ULL),
                               SPECTRAL_ITERATIVE_COMPONENT_THRESHOLD, lc))
        return false;
      float mx = 0.0f;
      for (float v : lc) mx = std::max(mx, std::fabs(v));
      const float expansion = mx > 0.0f ? range / mx : 1.0f;
      for (auto& v : lc) v *= expansion;
      add_noise(lc, seed ^ ((static_cast<uint64_t>(cid) + 1) * 0x94D049BBULL), INIT_NOISE);
    }
    for (size_t li = 0; li < cs; ++li)
      for (int c = 0; c < dim; ++c)
        result[comp[li] * dim + c] = lc[li * dim + c] + anchor[c];
  }
  for (float v : result)
    if (!std::isfinite(v)) return false;
  return true;
}

/* lib.rs:3855-3870, 4355-4376 spectral_init(_connected) */
bool spectral_init(const float* x, int64_t n, int64_t d, int dim,
                   const vector<Edge>& edges, uint64_t seed, vector<float>& out) {
  if (n <= dim + 1 || edges.empty()) return false;
  vector<int64_t> hs(edges.size()), ts(edges.size());
  for (size_t e = 0; e < edges.size(); ++e) {
    hs[e] = edges[e].head;
    ts[e] = edges[e].tail;
  }
  const auto comps = components_of(n, hs.data(), ts.data(), static_cast<int64_t>(edges.size()));
  if (comps.size() > 1) return multi_component_init(x, n, d, dim, edges, comps, seed, out);
  const size_t thr = static_cast<size_t>(d) <= SPECTRAL_EXACT_LOW_DIMENSIONAL_MAX_FEATURES
                         ? SPECTRAL_EXACT_LOW_DIMENSIONAL_MAX_SAMPLES + 1
                         : SPECTRAL_ITERATIVE_CONNECTED_THRESHOLD;
  if (!spectral_from_edges(static_cast<size_t>(n), dim, edges, seed, thr, out)) return false;
  noisy_scale(out, seed ^ 0x5DEECE66DULL, INIT_MAX_COORD, INIT_NOISE);
  for (float v : out)
    if (!std::isfinite(v)) return false;
  return true;
}

/* lib.rs:4937-4960 make_epochs_per_sample */
vector<float> epochs_per_sample(const float* w, int64_t ne, int32_t n_epochs) {
  vector<float> out(ne);
  if (ne == 0) return out;
  float mx = -std::numeric_limits<float>::infinity();
  for (int64_t e = 0; e < ne; ++e) mx = std::max(mx, w[e]);
  mx = std::max(mx, 1e-12f);
  const float ef = static_cast<float>(n_epochs);
  for (int64_t e = 0; e < ne; ++e) {
    const float ns = ef * (w[e] / mx);
    out[e] = ns > 0.0f ? ef / ns : std::numeric_limits<float>::infinity();
  }
  return out;
}

inline float clip4(float v) { return std::min(std::max(v, -4.0f), 4.0f); }

/* lib.rs:4990-5000 flat_squared_distance: always a sequential sum. */
inline float flat_sq_dist(const float* x, const float* y, int dim) {
  float s = 0.0f;
  for (int c = 0; c < dim; ++c) {
    const float df = x[c] - y[c];
    s += df * df;
  }
  return s;
}

/* lib.rs:5137-5318 optimize_layout_training_flat, and lib.rs:5324-5441 the
 * densMAP variant (density term, lib.rs:5110-5134). */
void optimize_layout(float* y, int64_t n, int dim, const int64_t* heads,
                     const int64_t* tails, const float* w, int64_t ne,
                     int32_t n_epochs, float a, float b, float alpha0, int32_t nsr,
                     float gamma, uint64_t seed, const float* radii,
                     int32_t dens_start, float dens_lambda, float var_shift) {
  if (ne == 0 || n == 0 || n_epochs == 0) return;
  const vector<float> eps = epochs_per_sample(w, ne, n_epochs);
  vector<float> eons = eps;
  vector<float> eps_neg(ne);
  for (int64_t e = 0; e < ne; ++e) eps_neg[e] = eps[e] / static_cast<float>(nsr);
  vector<float> eonns = eps_neg;
  SmallRng rng(seed);
  const float rep = 2.0f * gamma * b;
  const float att = -2.0f * a * b;
  const bool dens = radii != nullptr;
  const uint64_t nv = static_cast<uint64_t>(n);
  for (int32_t epoch = 0; epoch < n_epochs; ++epoch) {
    const float alpha = alpha0 * (1.0f - static_cast<float>(epoch) / static_cast<float>(n_epochs));
    const float ef = static_cast<float>(epoch);
    for (int64_t e = 0; e < ne; ++e) {
      if (eons[e] > ef) continue;
      const int64_t h = heads[e], t = tails[e];
      float* yh = y + h * dim;
      float* yt = y + t * dim;
      if (h != t) {
        const float d2 = dens ? sq_dist(yh, yt, dim) : flat_sq_dist(yh, yt, dim);
        if (!dens) {
          if (d2 > 0.0f) {
            const float pb = std::pow(d2, b);
            const float gc = att * (pb / d2) / (a * pb + 1.0f);
            for (int c = 0; c < dim; ++c) {
              const float hv = yh[c], tv = yt[c];
              const float g = clip4(gc * (hv - tv));
              yh[c] = hv + g * alpha;
              yt[c] = tv - g * alpha;
            }
          }
        } else {
          float gc = 0.0f;
          if (d2 > 0.0f) {
            const float pb = std::pow(d2, b);
            gc = -2.0f * a * b * std::pow(d2, b - 1.0f) / (a * pb + 1.0f);
          }
          for (int c = 0; c < dim; ++c) {
            const float g = clip4(gc * (yh[c] - yt[c]));
            yh[c] += g * alpha;
            yt[c] -= g * alpha;
          }
        }
      }
      if (dens && epoch >= dens_start && h != t) {
        const float cur = std::sqrt(sq_dist(yh, yt, dim));
        const float target = std::max(std::exp((radii[h] + radii[t]) * 0.5f), 1e-6f);
        const float coeff = DENSITY_GRADIENT_SCALE * dens_lambda * (cur - target) / (cur + var_shift);
        for (int c = 0; c < dim; ++c) {
          const float dir = (yh[c] - yt[c]) / (cur + 1e-6f);
          const float g = clip4(coeff * dir);
          yh[c] -= g * alpha;
          yt[c] += g * alpha;
        }
      }
      eons[e] += eps[e];
      const float en = eps_neg[e];
      if (!std::isfinite(en) || en <= 0.0f) continue;
      const float fl = std::floor((ef - eonns[e]) / en);
      const size_t nneg = fl > 0.0f ? static_cast<size_t>(fl) : 0;
      for (size_t s = 0; s < nneg; ++s) {
        const int64_t k = static_cast<int64_t>(rng.gen_index(nv));
        if (k == h) continue;
        float* yk = y + k * dim;
        const float d2 = dens ? sq_dist(yh, yk, dim) : flat_sq_dist(yh, yk, dim);
        if (d2 > 0.0f) {
          const float pb = std::pow(d2, b);
          const float gc = rep / ((0.001f + d2) * (a * pb + 1.0f));
          if (gc > 0.0f)
            for (int c = 0; c < dim; ++c) {
              const float hv = yh[c], kv = yk[c];
              const float g = clip4(gc * (hv - kv));
              yh[c] = hv + g * alpha;
            }
        }
      }
      eonns[e] += static_cast<float>(nneg) * en;
    }
  }
}

/* lib.rs:5057-5083 local_radii_from_knn */
void local_radii(const int64_t* idx, const float* dist, int64_t n, int32_t k,
                 float vs, float* out) {
  for (int64_t i = 0; i < n; ++i) {
    float s = 0.0f;
    size_t c = 0;
    for (int32_t j = 0; j < k; ++j) {
      const int64_t nb = idx[i * k + j];
      const float d = dist[i * k + j];
      if (nb == i && std::fabs(d) <= SMOOTH_K_TOLERANCE) continue;
      s += std::log(std::max(d, 0.0f) + vs);
      ++c;
    }
    out[i] = c == 0 ? std::log(vs) : s / static_cast<float>(c);
  }
}

/* lib.rs:5085-5108 embedding_radii_from_knn */
void embedding_radii(const float* y, int64_t n, int dim, const int64_t* idx,
                     int32_t k, float vs, float* out) {
  for (int64_t i = 0; i < n; ++i) {
    float s = 0.0f;
    size_t c = 0;
    for (int32_t j = 0; j < k; ++j) {
      const int64_t nb = idx[i * k + j];
      if (nb == i) continue;
      const float d = std::sqrt(sq_dist(y + i * dim, y + nb * dim, dim));
      s += std::log(std::max(d, 0.0f) + vs);
      ++c;
    }
    out[i] = c == 0 ? std::log(vs) : s / static_cast<float>(c);
  }
}

/* lib.rs:5028-5033 query_head_seed */
inline uint64_t query_head_seed(uint64_t seed, uint64_t head) {
  uint64_t m = seed ^ (head * 0x9E3779B97F4A7C15ULL);
  m = (m ^ (m >> 30)) * 0xBF58476D1CE4E5B9ULL;
  m = (m ^ (m >> 27)) * 0x94D049BB133111EBULL;
  return m ^ (m >> 31);
}

/* lib.rs:4793-4836 init_graph_transform */
void init_graph_transform(int64_t m, int dim, const vector<Edge>& edges,
                          const float* base, float* out) {
  vector<vector<std::pair<int64_t, float>>> rows(m);
  for (const auto& e : edges) rows[e.head].push_back({e.tail, e.weight});
  const float nanv = std::numeric_limits<float>::quiet_NaN();
  for (int64_t i = 0; i < m; ++i) {
    float* o = out + i * dim;
    std::fill(o, o + dim, 0.0f);
    if (rows[i].empty()) {
      std::fill(o, o + dim, nanv);
      continue;
    }
    bool copied = false;
    for (const auto& pr : rows[i])
      if (std::fabs(pr.second - 1.0f) < std::numeric_limits<float>::epsilon()) {
        std::copy(base + pr.first * dim, base + pr.first * dim + dim, o);
        copied = true;
        break;
      }
    if (copied) continue;
    float rs = 0.0f;
    for (const auto& pr : rows[i]) rs += pr.second;
    if (rs <= 0.0f) {
      std::fill(o, o + dim, nanv);
      continue;
    }
    for (const auto& pr : rows[i]) {
      const float ww = pr.second / rs;
      for (int c = 0; c < dim; ++c) o[c] += ww * base[pr.first * dim + c];
    }
  }
}

/* lib.rs:5444-5544 optimize_layout_transform (rayon over heads -> OpenMP) */
void optimize_layout_transform(float* emb, int64_t m, int dim, const float* base,
                               int64_t n_base, const vector<Edge>& edges,
                               int32_t n_epochs, float a, float b, float alpha0,
                               int32_t nsr, float gamma, uint64_t seed, int nt) {
  if (edges.empty() || m == 0 || n_epochs == 0) return;
  vector<float> w(edges.size());
  for (size_t e = 0; e < edges.size(); ++e) w[e] = edges[e].weight;
  const vector<float> eps = epochs_per_sample(w.data(), static_cast<int64_t>(w.size()), n_epochs);
  vector<float> eps_neg(eps.size());
  for (size_t e = 0; e < eps.size(); ++e) eps_neg[e] = eps[e] / static_cast<float>(nsr);
  vector<size_t> off(m + 1, 0);
  {
    size_t ei = 0;
    for (int64_t h = 0; h < m; ++h) {
      off[h] = ei;
      while (ei < edges.size() && edges[ei].head == h) ++ei;
    }
    off[m] = ei;
    if (ei != edges.size()) return; /* not grouped by head */
  }
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt) if (m >= 32)
  for (int64_t h = 0; h < m; ++h) {
    float* cur = emb + h * dim;
    const size_t s0 = off[h], s1 = off[h + 1];
    vector<float> eons(eps.begin() + s0, eps.begin() + s1);
    vector<float> eonns(eps_neg.begin() + s0, eps_neg.begin() + s1);
    SmallRng rng(query_head_seed(seed, static_cast<uint64_t>(h)));
    for (int32_t epoch = 0; epoch < n_epochs; ++epoch) {
      const float alpha = alpha0 * (1.0f - static_cast<float>(epoch) / static_cast<float>(n_epochs));
      const float ef = static_cast<float>(epoch);
      for (size_t e = s0; e < s1; ++e) {
        const size_t li = e - s0;
        bool finite = true;
        for (int c = 0; c < dim; ++c) finite &= std::isfinite(cur[c]);
        if (eons[li] > ef || !finite) continue;
        const float* other = base + edges[e].tail * dim;
        const float d2 = sq_dist(cur, other, dim);
        float gc = 0.0f;
        if (d2 > 0.0f) {
          const float pb = std::pow(d2, b);
          gc = -2.0f * a * b * std::pow(d2, b - 1.0f) / (a * pb + 1.0f);
        }
        for (int c = 0; c < dim; ++c) {
          const float g = clip4(gc * (cur[c] - other[c]));
          cur[c] += g * alpha;
        }
        eons[li] += eps[e];
        const float en = eps_neg[e];
        if (!std::isfinite(en) || en <= 0.0f) continue;
        const float fl = std::floor((ef - eonns[li]) / en);
        const size_t nneg = fl > 0.0f ? static_cast<size_t>(fl) : 0;
        for (size_t s = 0; s < nneg; ++s) {
          const int64_t k = static_cast<int64_t>(rng.gen_index(static_cast<uint64_t>(n_base)));
          const float* o2 = base + k * dim;
          const float dd = sq_dist(cur, o2, dim);
          float g2 = 0.0f;
          if (dd > 0.0f) {
            const float pb = std::pow(dd, b);
            g2 = 2.0f * gamma * b / ((0.001f + dd) * (a * pb + 1.0f));
          }
          if (g2 > 0.0f)
            for (int c = 0; c < dim; ++c) {
              const float g = clip4(g2 * (cur[c] - o2[c]));
              cur[c] += g * alpha;
            }
        }
        eonns[li] += static_cast<float>(nneg) * en;
      }
    }
  }
}

/* ------------------------------------------------ approximate kNN ----
 * lib.rs:2932-3168 approximate_nearest_neighbors_with_distance: random
 * initial pools, then rounds over neighbours-of-neighbours candidates with a
 * few random explorations, each round built from the previous table, until
 * no row changes, under any of the metrics (lib.rs:3213-3281). */
uint64_t mix_approx_seed(uint64_t seed, uint64_t row, uint64_t iteration) {
  uint64_t v = seed ^ (row * 0x9E3779B97F4A7C15ULL) ^ (iteration * 0xBF58476D1CE4E5B9ULL);
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
  return v ^ (v >> 31);
}

inline bool push_unique(vector<int64_t>& out, int64_t idx) {
  for (int64_t v : out)
    if (v == idx) return false;
  out.push_back(idx);
  return true;
}

/* lib.rs:2932-2938 sorted_unique_neighbors (inputs are unique already) */
void sorted_top(vector<std::pair<int64_t, float>>& pairs, int32_t k) {
  std::sort(pairs.begin(), pairs.end(), [](const std::pair<int64_t, float>& a,
                                           const std::pair<int64_t, float>& b) {
    return nb_less(a.first, a.second, b.first, b.second);
  });
  if (static_cast<int32_t>(pairs.size()) > k) pairs.resize(k);
}

void knn_approx_impl(const float* x, int64_t n, int64_t d, int32_t k, int32_t metric,
                     int32_t candidate_pool, int32_t n_iters, uint64_t seed, int nt,
                     int64_t* out_idx, float* out_dist, float metric_p = 2.0f) {
  vector<float> norms;
  if (metric == 1) {
    norms.resize(n);
    for (int64_t i = 0; i < n; ++i) norms[i] = l2_norm(x + i * d, d);
  }
  auto dist = [&](int64_t i, int64_t j) -> float {
    if (metric == 1) return cosine_with_norms(x + i * d, x + j * d, d, norms[i], norms[j]);
    if (metric >= 2) return metric_distance(metric, x + i * d, x + j * d, d, metric_p);
    return std::sqrt(sq_dist(x + i * d, x + j * d, d));
  };
  const int64_t pool = std::min<int64_t>(std::max<int64_t>(candidate_pool, k), n - 1);
  const int64_t max_cand = static_cast<int64_t>(std::max<int32_t>(candidate_pool, k)) * 4 + 1;
  vector<int64_t> tab_i(static_cast<size_t>(n) * k), nxt_i(tab_i.size());
  vector<float> tab_d(tab_i.size()), nxt_d(tab_i.size());
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
  for (int64_t i = 0; i < n; ++i) {
    SmallRng rng(mix_approx_seed(seed, static_cast<uint64_t>(i), 0));
    vector<int64_t> cand{i};
    while (static_cast<int64_t>(cand.size()) < pool + 1)
      push_unique(cand, static_cast<int64_t>(rng.gen_index(static_cast<uint64_t>(n))));
    vector<std::pair<int64_t, float>> pairs;
    for (int64_t j : cand) pairs.push_back({j, dist(i, j)});
    sorted_top(pairs, k);
    for (int32_t c = 0; c < k; ++c) {
      tab_i[i * k + c] = pairs[c].first;
      tab_d[i * k + c] = pairs[c].second;
    }
  }
  const int64_t exploration = std::min<int64_t>(std::max<int64_t>(k / 2, 4), max_cand);
  for (int32_t it = 0; it < n_iters; ++it) {
    int changed = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt) reduction(| : changed)
    for (int64_t i = 0; i < n; ++i) {
      SmallRng rng(mix_approx_seed(seed, static_cast<uint64_t>(i), static_cast<uint64_t>(it) + 1));
      /* lib.rs:3053-3081 collect_approx_candidate_indices */
      vector<int64_t> cand;
      for (int32_t a = 0; a < k; ++a) {
        const int64_t nb = tab_i[i * k + a];
        push_unique(cand, nb);
        for (int32_t b = 0; b < k; ++b) push_unique(cand, tab_i[nb * k + b]);
      }
      push_unique(cand, i);
      for (int64_t e = 0; e < exploration; ++e)
        push_unique(cand, static_cast<int64_t>(rng.gen_index(static_cast<uint64_t>(n))));
      std::sort(cand.begin(), cand.end());
      if (static_cast<int64_t>(cand.size()) > max_cand) {
        for (int64_t s2 = 0; s2 < max_cand; ++s2) {
          const int64_t r = s2 + static_cast<int64_t>(
                                     rng.gen_index(static_cast<uint64_t>(cand.size() - s2)));
          std::swap(cand[s2], cand[r]);
        }
        cand.resize(max_cand);
      }
      vector<std::pair<int64_t, float>> pairs;
      for (int64_t j : cand) pairs.push_back({j, dist(i, j)});
      sorted_top(pairs, k);
      bool same = true;
      for (int32_t c = 0; c < k; ++c) {
        same = same && tab_i[i * k + c] == pairs[c].first && tab_d[i * k + c] == pairs[c].second;
        nxt_i[i * k + c] = pairs[c].first;
        nxt_d[i * k + c] = pairs[c].second;
      }
      if (!same) changed = 1;
    }
    tab_i.swap(nxt_i);
    tab_d.swap(nxt_d);
    if (!changed) break;
  }
  std::copy(tab_i.begin(), tab_i.end(), out_idx);
  std::copy(tab_d.begin(), tab_d.end(), out_dist);
}

/* lib.rs:3420-3422 inverse_neighbor_count */
int32_t inverse_neighbor_count(int32_t n_neighbors, int64_t n_train) {
  return static_cast<int32_t>(std::max<int64_t>(std::min<int64_t>(n_neighbors, n_train), 1));
}

/* Partial-pivot LU of a small f32 matrix and its solve, in the steps of
 * nalgebra 0.33's LU::new / LU::solve (the crate is not in the reference
 * tree; restated from its published source): pivot = first row of maximal
 * |.| in the column, a zero pivot skips the column, coefficients are scaled
 * by 1/pivot, rows updated as y + (-p) * x; the solve permutes b, runs the
 * unit-lower then the upper substitution and fails on a zero diagonal. */
struct SmallLu {
  int n = 0;
  vector<float> lu;            /* row-major n x n */
  vector<std::pair<int, int>> swaps;
  void factor(const vector<float>& m, int dim) {
    n = dim;
    lu = m;
    swaps.clear();
    for (int i = 0; i < n; ++i) {
      int piv = i;
      float best = std::fabs(lu[i * n + i]);
      for (int r = i + 1; r < n; ++r)
        if (std::fabs(lu[r * n + i]) > best) {
          best = std::fabs(lu[r * n + i]);
          piv = r;
        }
      const float diag = lu[piv * n + i];
      if (diag == 0.0f) continue;
      if (piv != i) {
        swaps.push_back({i, piv});
        for (int c = 0; c < n; ++c) std::swap(lu[i * n + c], lu[piv * n + c]);
      }
      const float inv = 1.0f / diag;
      for (int r = i + 1; r < n; ++r) lu[r * n + i] *= inv;
      for (int c = i + 1; c < n; ++c) {
        const float a = -lu[i * n + c];
        for (int r = i + 1; r < n; ++r) lu[r * n + c] = a * lu[r * n + i] + lu[r * n + c];
      }
    }
  }
  /* b: n x cols row-major, solved in place; false when singular */
  bool solve(vector<float>& b, int64_t cols) const {
    for (const auto& sw : swaps)
      for (int64_t c = 0; c < cols; ++c) std::swap(b[sw.first * cols + c], b[sw.second * cols + c]);
    for (int64_t c = 0; c < cols; ++c) {
      for (int i = 0; i + 1 < n; ++i) {
        const float coeff = b[i * cols + c];
        for (int r = i + 1; r < n; ++r) b[r * cols + c] = -coeff * lu[r * n + i] + b[r * cols + c];
      }
      for (int i = n - 1; i >= 0; --i) {
        const float diag = lu[i * n + i];
        if (diag == 0.0f) return false;
        const float coeff = b[i * cols + c] / diag;
        b[i * cols + c] = coeff;
        for (int r = 0; r < i; ++r) b[r * cols + c] = -coeff * lu[r * n + i] + b[r * cols + c];
      }
    }
    return true;
  }
};

/* lib.rs:1856-1890: the inverse kNN weights 1 / (1 + a d^2b), normalised per
 * row, the fixed rows (nearest embedded point at distance <= 1e-6) and the
 * edge list (query, training row, raw weight). */
void inverse_weights(const float* dist, int64_t m, int32_t k, float a, float b,
                     vector<float>& raw, vector<float>& norm, vector<char>& fixed) {
  raw.assign(static_cast<size_t>(m) * k, 0.0f);
  norm.assign(static_cast<size_t>(m) * k, 0.0f);
  fixed.assign(m, 0);
  const float two_b = 2.0f * b;
  for (int64_t i = 0; i < m; ++i) {
    float row_sum = 0.0f;
    for (int32_t j = 0; j < k; ++j) {
      const float w = 1.0f / (1.0f + a * std::pow(dist[i * k + j], two_b));
      if (w > 0.0f && std::isfinite(w)) {
        raw[i * k + j] = w;
        norm[i * k + j] = w;
        row_sum += w;
      }
    }
    if (row_sum > 0.0f)
      for (int32_t j = 0; j < k; ++j) norm[i * k + j] /= row_sum;
    fixed[i] = dist[i * k] <= 1e-6f;
  }
}

/* lib.rs:4838-4935 init_inverse_points */
void init_inverse_points(const float* q, int64_t m, int low_dim, const int64_t* idx,
                         const vector<float>& norm, const vector<char>& fixed, int32_t k,
                         const float* y_train, const float* x_train, int64_t d, float* out) {
  const float nanv = std::numeric_limits<float>::quiet_NaN();
  for (int64_t i = 0; i < m; ++i) {
    float* res = out + i * d;
    if (fixed[i]) {
      std::copy(x_train + idx[i * k] * d, x_train + idx[i * k] * d + d, res);
      continue;
    }
    vector<float> z_bar(low_dim, 0.0f), x_bar(d, 0.0f);
    float weight_sum = 0.0f;
    for (int32_t j = 0; j < k; ++j) {
      const float w = norm[i * k + j];
      if (w <= 0.0f || !std::isfinite(w)) continue;
      weight_sum += w;
      const int64_t r = idx[i * k + j];
      for (int c = 0; c < low_dim; ++c) z_bar[c] += w * y_train[r * low_dim + c];
      for (int64_t f = 0; f < d; ++f) x_bar[f] += w * x_train[r * d + f];
    }
    if (weight_sum <= 0.0f) {
      std::fill(res, res + d, nanv);
      continue;
    }
    vector<float> gram(static_cast<size_t>(low_dim) * low_dim, 0.0f);
    vector<float> rhs(static_cast<size_t>(low_dim) * d, 0.0f);
    int non_zero = 0;
    for (int32_t j = 0; j < k; ++j) {
      const float w = norm[i * k + j];
      if (w <= 0.0f || !std::isfinite(w)) continue;
      ++non_zero;
      const int64_t r = idx[i * k + j];
      for (int aa = 0; aa < low_dim; ++aa) {
        const float z_a = y_train[r * low_dim + aa] - z_bar[aa];
        for (int bb = 0; bb < low_dim; ++bb) {
          const float z_b = y_train[r * low_dim + bb] - z_bar[bb];
          gram[aa * low_dim + bb] += w * z_a * z_b;
        }
        for (int64_t f = 0; f < d; ++f) {
          const float x_d = x_train[r * d + f] - x_bar[f];
          rhs[aa * d + f] += w * z_a * x_d;
        }
      }
    }
    for (int c = 0; c < low_dim; ++c) gram[c * low_dim + c] += 1e-3f;
    vector<float> pred = x_bar;
    if (non_zero > low_dim) {
      SmallLu lu;
      lu.factor(gram, low_dim);
      vector<float> beta = rhs;
      if (lu.solve(beta, d)) {
        bool valid = true;
        for (int64_t f = 0; f < d && valid; ++f) {
          float value = x_bar[f];
          for (int aa = 0; aa < low_dim; ++aa)
            value += (q[i * low_dim + aa] - z_bar[aa]) * beta[aa * d + f];
          if (!std::isfinite(value)) valid = false;
          else pred[f] = value;
        }
        if (!valid) pred = x_bar;
      }
    }
    std::copy(pred.begin(), pred.end(), res);
  }
}

/* lib.rs:5035-5043 euclidean_distance_with_grad_denom */
inline float dist_sequential(const float* x, const float* y, int64_t d) {
  float sq = 0.0f;
  for (int64_t f = 0; f < d; ++f) {
    const float df = x[f] - y[f];
    sq += df * df;
  }
  return std::sqrt(sq);
}

/* lib.rs:5547-5645 optimize_layout_inverse (one SmallRng stream, edges in
 * query order) */
void optimize_layout_inverse(float* head, int64_t m, const float* tail, int64_t n, int64_t d,
                             const vector<Edge>& edges, const vector<char>& fixed,
                             const float* sigmas, const float* rhos, int32_t n_epochs,
                             float alpha0, int32_t nsr, float gamma, uint64_t seed) {
  if (edges.empty() || m == 0 || n_epochs == 0) return;
  vector<float> w(edges.size());
  for (size_t e = 0; e < edges.size(); ++e) w[e] = edges[e].weight;
  const vector<float> eps = epochs_per_sample(w.data(), static_cast<int64_t>(w.size()), n_epochs);
  vector<float> eons = eps;
  vector<float> eps_neg(eps.size());
  for (size_t e = 0; e < eps.size(); ++e) eps_neg[e] = eps[e] / static_cast<float>(nsr);
  vector<float> eonns = eps_neg;
  SmallRng rng(seed);
  for (int32_t epoch = 0; epoch < n_epochs; ++epoch) {
    const float alpha = alpha0 * (1.0f - static_cast<float>(epoch) / static_cast<float>(n_epochs));
    const float ef = static_cast<float>(epoch);
    for (size_t e = 0; e < edges.size(); ++e) {
      if (eons[e] > ef) continue;
      const int64_t h = edges[e].head, t = edges[e].tail;
      float* cur = head + h * d;
      bool finite = true;
      for (int64_t f = 0; f < d; ++f) finite &= std::isfinite(cur[f]);
      if (fixed[h] || !finite) continue;
      const float* other = tail + t * d;
      const float gd = 1e-6f + dist_sequential(cur, other, d);
      const float gc = -(1.0f / (edges[e].weight * sigmas[t] + 1e-6f));
      for (int64_t f = 0; f < d; ++f) {
        const float gdist = (cur[f] - other[f]) / gd;
        cur[f] += clip4(gc * gdist) * alpha;
      }
      eons[e] += eps[e];
      const float en = eps_neg[e];
      if (!std::isfinite(en) || en <= 0.0f) continue;
      const float fl = std::floor((ef - eonns[e]) / en);
      const size_t nneg = fl > 0.0f ? static_cast<size_t>(fl) : 0;
      for (size_t s = 0; s < nneg; ++s) {
        const int64_t kk = static_cast<int64_t>(rng.gen_index(static_cast<uint64_t>(n)));
        const float* o2 = tail + kk * d;
        const float dn = dist_sequential(cur, o2, d);
        const float gd2 = 1e-6f + dn;
        const float wh = std::exp(-std::fmax(dn - rhos[kk], 1e-6f) / (sigmas[kk] + 1e-6f));
        const float gc2 = -gamma * ((0.0f - wh) / ((1.0f - wh) * sigmas[kk] + 1e-6f));
        for (int64_t f = 0; f < d; ++f) {
          const float gdist = (cur[f] - o2[f]) / gd2;
          cur[f] += clip4(gc2 * gdist) * alpha;
        }
      }
      eonns[e] += static_cast<float>(nneg) * en;
    }
  }
}

/* lib.rs:5054-5055 */
int32_t density_start(int32_t n_epochs, float frac) {
  const int64_t active = static_cast<int64_t>(std::ceil(static_cast<float>(n_epochs) * frac));
  const int64_t s = static_cast<int64_t>(n_epochs) - std::max<int64_t>(active, 1);
  return static_cast<int32_t>(std::max<int64_t>(s, 0));
}

/* lib.rs:5669-5746 find_ab_params */
float curve_loss(float a, float b, const vector<float>& xv, const vector<float>& yv) {
  float loss = 0.0f;
  const float ex = 2.0f * b;
  for (size_t i = 0; i < xv.size(); ++i) {
    const float model = 1.0f / (1.0f + a * std::pow(xv[i], ex));
    const float e = model - yv[i];
    loss += e * e;
  }
  return loss / static_cast<float>(xv.size());
}

}  // namespace

extern "C" {

/* Generator streams for known-answer tests: xoshiro256++ from a raw state,
 * SplitMix64 from a seed, and SmallRng::seed_from_u64 (the two composed). */
void orc_xoshiro256pp(const uint64_t* state, int64_t n, uint64_t* out) {
  SmallRng r = SmallRng::from_state(state);
  for (int64_t i = 0; i < n; ++i) out[i] = r.next_u64();
}
void orc_splitmix64(uint64_t seed, int64_t n, uint64_t* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = splitmix64_next(seed);
}
void orc_smallrng_u64(uint64_t seed, int64_t n, uint64_t* out) {
  SmallRng r(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = r.next_u64();
}
void orc_smallrng_index(uint64_t seed, uint64_t range, int64_t n, uint64_t* out) {
  SmallRng r(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = r.gen_index(range);
}

void orc_find_ab_params(float spread, float min_dist, float* a_out, float* b_out) {
  const int n = 300;
  const float xmax = spread * 3.0f;
  vector<float> xv(n), yv(n);
  for (int i = 0; i < n; ++i) {
    const float x = xmax * static_cast<float>(i) / (static_cast<float>(n) - 1.0f);
    xv[i] = x;
    yv[i] = x < min_dist ? 1.0f : std::exp(-(x - min_dist) / spread);
  }
  float best_a = 1.0f, best_b = 1.0f, best_loss = std::numeric_limits<float>::infinity();
  vector<float> xp(n);
  for (int bi = 0; bi <= 70; ++bi) {
    const float b = 0.3f + 2.7f * static_cast<float>(bi) / 70.0f;
    const float ex = 2.0f * b;
    for (int i = 0; i < n; ++i) xp[i] = std::pow(xv[i], ex);
    for (int ai = 0; ai <= 120; ++ai) {
      const float la = -3.0f + 6.0f * static_cast<float>(ai) / 120.0f;
      const float a = std::pow(10.0f, la);
      float loss = 0.0f;
      for (int i = 0; i < n; ++i) {
        const float model = 1.0f / (1.0f + a * xp[i]);
        const float e = model - yv[i];
        loss += e * e;
      }
      loss = loss / static_cast<float>(n);
      if (loss < best_loss) {
        best_loss = loss;
        best_a = a;
        best_b = b;
      }
    }
  }
  float step_a = std::max(best_a * 0.35f, 0.05f);
  float step_b = 0.2f;
  for (int it = 0; it < 80; ++it) {
    const float cand[8][2] = {
        {best_a + step_a, best_b},
        {std::max(best_a - step_a, 1e-6f), best_b},
        {best_a, best_b + step_b},
        {best_a, std::max(best_b - step_b, 1e-6f)},
        {best_a + step_a, best_b + step_b},
        {best_a + step_a, std::max(best_b - step_b, 1e-6f)},
        {std::max(best_a - step_a, 1e-6f), best_b + step_b},
        {std::max(best_a - step_a, 1e-6f), std::max(best_b - step_b, 1e-6f)},
    };
    bool improved = false;
    for (const auto& c : cand) {
      const float loss = curve_loss(c[0], c[1], xv, yv);
      if (loss < best_loss) {
        best_loss = loss;
        best_a = c[0];
        best_b = c[1];
        improved = true;
      }
    }
    if (!improved) {
      step_a *= 0.7f;
      step_b *= 0.7f;
      if (step_a < 1e-5f && step_b < 1e-5f) break;
    }
  }
  *a_out = best_a;
  *b_out = best_b;
}

void orc_knn(const float* x, int64_t n, int64_t d, int32_t k, int32_t metric,
             int32_t n_threads, int64_t* idx, float* dist) {
  knn_impl(x, n, x, n, d, k, metric, n_threads, idx, dist);
}

void orc_knn_metric(const float* q, int64_t m, const float* r, int64_t n, int64_t d, int32_t k,
                    int32_t metric, float metric_p, int32_t n_threads, int64_t* idx,
                    float* dist) {
  knn_impl(q, m, r, n, d, k, metric, n_threads, idx, dist, metric_p);
}

void orc_knn_approx(const float* x, int64_t n, int64_t d, int32_t k, int32_t metric,
                    float metric_p, int32_t candidate_pool, int32_t n_iters, uint64_t seed,
                    int32_t n_threads, int64_t* idx, float* dist) {
  const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
  knn_approx_impl(x, n, d, k, metric, candidate_pool, n_iters, seed, nt, idx, dist, metric_p);
}

void orc_knn_to_reference(const float* q, int64_t m, const float* r, int64_t n,
                          int64_t d, int32_t k, int32_t metric, int32_t n_threads,
                          int64_t* idx, float* dist) {
  knn_impl(q, m, r, n, d, k, metric, n_threads, idx, dist);
}

float orc_sq_euclidean(const float* x, const float* y, int64_t d) {
  return sq_dist(x, y, static_cast<size_t>(d));
}

float orc_cosine_distance(const float* x, const float* y, int64_t d) {
  return cosine_with_norms(x, y, d, l2_norm(x, d), l2_norm(y, d));
}

int32_t orc_knn_rows_include_self(const int64_t* idx, const float* dist, int64_t n, int32_t k) {
  return rows_include_self(idx, dist, n, k) ? 1 : 0;
}

void orc_smooth_knn_dist(const float* dist, int64_t n, int32_t k, float k_value,
                         float local_connectivity, float bandwidth, int32_t include_self,
                         float* sigmas, float* rhos) {
  smooth_knn(dist, n, k, k_value, local_connectivity, bandwidth, include_self != 0, sigmas, rhos);
}

void orc_smooth_knn_dist_fp64(const float* dist, int64_t n, int32_t k, float k_value,
                              float local_connectivity, float bandwidth, int32_t include_self,
                              float* sigmas, float* rhos) {
  smooth_knn(dist, n, k, k_value, local_connectivity, bandwidth, include_self != 0, sigmas, rhos,
             true);
}

int64_t orc_fuzzy_graph(const int64_t* idx, const float* dist, const float* sigmas,
                        const float* rhos, int64_t n, int32_t k, float mix,
                        int32_t n_epochs, int64_t cap, int64_t* heads, int64_t* tails,
                        float* w) {
  auto directed = membership(idx, dist, sigmas, rhos, n, k);
  auto edges = symmetrize(directed, mix);
  if (n_epochs > 0) prune(edges, n_epochs);
  if (static_cast<int64_t>(edges.size()) > cap) return -1;
  for (size_t e = 0; e < edges.size(); ++e) {
    heads[e] = edges[e].head;
    tails[e] = edges[e].tail;
    w[e] = edges[e].weight;
  }
  return static_cast<int64_t>(edges.size());
}

int64_t orc_connected_components(int64_t n, const int64_t* heads, const int64_t* tails,
                                 int64_t ne, int64_t* labels) {
  const auto comps = components_of(n, heads, tails, ne);
  for (size_t c = 0; c < comps.size(); ++c)
    for (int64_t v : comps[c]) labels[v] = static_cast<int64_t>(c);
  return static_cast<int64_t>(comps.size());
}

int32_t orc_initialize_embedding(const float* x, int64_t n, int64_t d, int32_t dim,
                                 const int64_t* heads, const int64_t* tails,
                                 const float* w, int64_t ne, int32_t init,
                                 uint64_t seed, float* y) {
  vector<Edge> edges(ne);
  for (int64_t e = 0; e < ne; ++e) edges[e] = {heads[e], tails[e], w[e]};
  vector<float> out;
  int32_t ok = 0;
  if (init == 1 && spectral_init(x, n, d, dim, edges, seed, out)) {
    ok = 1;
  } else {
    out = random_init(n, dim, seed, -10.0f, 10.0f);
  }
  std::copy(out.begin(), out.end(), y);
  return ok;
}

void orc_make_epochs_per_sample(const float* w, int64_t ne, int32_t n_epochs, float* eps) {
  const auto v = epochs_per_sample(w, ne, n_epochs);
  std::copy(v.begin(), v.end(), eps);
}

void orc_schedule_counts(const float* w, int64_t ne, int32_t n_epochs, int32_t nsr,
                         int32_t* pos_counts, int32_t* neg_counts,
                         int64_t* active_per_epoch, int64_t* neg_per_epoch) {
  const vector<float> eps = epochs_per_sample(w, ne, n_epochs);
  vector<float> eons = eps, eps_neg(ne), eonns(ne);
  for (int64_t e = 0; e < ne; ++e) eps_neg[e] = eps[e] / static_cast<float>(nsr);
  eonns = eps_neg;
  for (int64_t e = 0; e < ne; ++e) pos_counts[e] = neg_counts[e] = 0;
  for (int32_t ep = 0; ep < n_epochs; ++ep) {
    const float ef = static_cast<float>(ep);
    int64_t act = 0, negs = 0;
    for (int64_t e = 0; e < ne; ++e) {
      if (eons[e] > ef) continue;
      ++act;
      pos_counts[e] += 1;
      eons[e] += eps[e];
      const float en = eps_neg[e];
      if (!std::isfinite(en) || en <= 0.0f) continue;
      const float fl = std::floor((ef - eonns[e]) / en);
      const size_t nneg = fl > 0.0f ? static_cast<size_t>(fl) : 0;
      neg_counts[e] += static_cast<int32_t>(nneg);
      negs += static_cast<int64_t>(nneg);
      eonns[e] += static_cast<float>(nneg) * en;
    }
    if (active_per_epoch) active_per_epoch[ep] = act;
    if (neg_per_epoch) neg_per_epoch[ep] = negs;
  }
}

void orc_optimize_layout(float* y, int64_t n, int32_t dim, const int64_t* heads,
                         const int64_t* tails, const float* w, int64_t ne,
                         int32_t n_epochs, float a, float b, float lr, int32_t nsr,
                         float gamma, uint64_t seed, const float* radii,
                         int32_t dens_start, float dens_lambda, float var_shift) {
  optimize_layout(y, n, dim, heads, tails, w, ne, n_epochs, a, b, lr, nsr, gamma, seed,
                  radii, dens_start, dens_lambda, var_shift);
}

int32_t orc_normalized_laplacian(int64_t n, const int64_t* heads, const int64_t* tails,
                                 const float* w, int64_t ne, int32_t from_affinity,
                                 double* out) {
  vector<double> lap;
  bool ok;
  if (from_affinity) {
    /* the dense affinity path: max-combined, symmetrised (lib.rs:7492-7519) */
    vector<double> aff(static_cast<size_t>(n * n), 0.0);
    for (int64_t e = 0; e < ne; ++e) {
      if (heads[e] == tails[e]) continue;
      const double v = static_cast<double>(std::max(w[e], 0.0f));
      double& a = aff[heads[e] * n + tails[e]];
      if (v > a) a = v;
    }
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = i + 1; j < n; ++j) {
        const double sym = std::max(aff[i * n + j], aff[j * n + i]);
        aff[i * n + j] = aff[j * n + i] = sym;
      }
    ok = laplacian_from_affinity(aff, static_cast<size_t>(n), lap);
  } else {
    vector<Edge> edges(ne);
    for (int64_t e = 0; e < ne; ++e) edges[e] = {heads[e], tails[e], w[e]};
    ok = laplacian_from_edges(static_cast<size_t>(n), edges, lap);
  }
  if (!ok) return 0;
  std::copy(lap.begin(), lap.end(), out);
  return 1;
}

int32_t orc_density_start_epoch(int32_t n_epochs, float dens_frac) {
  return density_start(n_epochs, dens_frac);
}

void orc_local_radii(const int64_t* idx, const float* dist, int64_t n, int32_t k,
                     float vs, float* out) {
  local_radii(idx, dist, n, k, vs, out);
}

void orc_embedding_radii(const float* y, int64_t n, int32_t dim, const int64_t* idx,
                         int32_t k, float vs, float* out) {
  embedding_radii(y, n, dim, idx, k, vs, out);
}

int32_t orc_fit(const orc_params* p, const float* x, int64_t n, int64_t d,
                const int64_t* knn_idx, const float* knn_dist, int64_t knn_cols,
                float* y_out, float* sigmas, float* rhos, float* rad_orig,
                float* rad_emb, orc_timings* t) {
  orc_timings tm;
  std::memset(&tm, 0, sizeof(tm));
  const double t_total = now_sec();
  const int32_t k = p->n_neighbors;
  const int nt = p->n_threads > 0 ? p->n_threads : omp_get_max_threads();
  vector<int64_t> idx(static_cast<size_t>(n) * k);
  vector<float> dist(static_cast<size_t>(n) * k);
  double s = now_sec();
  if (knn_idx == nullptr) {
    knn_impl(x, n, x, n, d, k, p->metric, nt, idx.data(), dist.data(), p->metric_p);
  }
  tm.knn_sec = now_sec() - s;
  const int32_t n_epochs = p->n_epochs > 0 ? p->n_epochs : (n <= 10000 ? 500 : 200);
  s = now_sec();
  float a, b;
  orc_find_ab_params(p->spread, p->min_dist, &a, &b);
  tm.curve_params_sec = now_sec() - s;
  s = now_sec();
  if (knn_idx != nullptr) {
    /* trim to n_neighbors (lib.rs:1968-2040; validation is the product's) */
    for (int64_t i = 0; i < n; ++i)
      for (int32_t c = 0; c < k; ++c) {
        idx[i * k + c] = knn_idx[i * knn_cols + c];
        dist[i * k + c] = knn_dist[i * knn_cols + c];
      }
  }
  tm.knn_validate_trim_sec = now_sec() - s;
  s = now_sec();
  vector<float> sig(n), rho(n);
  smooth_knn(dist.data(), n, k, static_cast<float>(k), p->local_connectivity,
             DEFAULT_BANDWIDTH, rows_include_self(idx.data(), dist.data(), n, k),
             sig.data(), rho.data());
  tm.smooth_knn_sec = now_sec() - s;
  s = now_sec();
  const bool dens_out = p->densmap || p->output_dens;
  vector<float> rorig;
  if (dens_out) {
    rorig.resize(n);
    local_radii(idx.data(), dist.data(), n, k, p->dens_var_shift, rorig.data());
  }
  const bool dens_active = p->densmap && p->dens_lambda > 0.0f;
  tm.density_original_sec = now_sec() - s;
  s = now_sec();
  auto directed = membership(idx.data(), dist.data(), sig.data(), rho.data(), n, k);
  tm.membership_sec = now_sec() - s;
  s = now_sec();
  auto edges = symmetrize(directed, p->set_op_mix_ratio);
  tm.symmetrize_sec = now_sec() - s;
  s = now_sec();
  prune(edges, n_epochs);
  tm.prune_sec = now_sec() - s;
  s = now_sec();
  vector<float> y;
  const int dim = p->n_components;
  if (!(p->init == 1 && spectral_init(x, n, d, dim, edges, p->random_seed, y)))
    y = random_init(n, dim, p->random_seed, -10.0f, 10.0f);
  tm.init_sec = now_sec() - s;
  s = now_sec();
  {
    const size_t ne = edges.size();
    vector<int64_t> hs(ne), ts(ne);
    vector<float> ws(ne);
    for (size_t e = 0; e < ne; ++e) {
      hs[e] = edges[e].head;
      ts[e] = edges[e].tail;
      ws[e] = edges[e].weight;
    }
    optimize_layout(y.data(), n, dim, hs.data(), ts.data(), ws.data(),
                    static_cast<int64_t>(ne), n_epochs, a, b, p->learning_rate,
                    p->negative_sample_rate, p->repulsion_strength,
                    p->random_seed ^ 0x9E3779B97F4A7C15ULL,
                    dens_active ? rorig.data() : nullptr,
                    density_start(n_epochs, p->dens_frac), p->dens_lambda,
                    p->dens_var_shift);
  }
  tm.optimize_sec = now_sec() - s;
  s = now_sec();
  if (dens_out && rad_emb)
    embedding_radii(y.data(), n, dim, idx.data(), k, p->dens_var_shift, rad_emb);
  tm.density_embedding_sec = now_sec() - s;
  s = now_sec();
  std::copy(y.begin(), y.end(), y_out);
  tm.output_copy_sec = now_sec() - s;
  if (sigmas) std::copy(sig.begin(), sig.end(), sigmas);
  if (rhos) std::copy(rho.begin(), rho.end(), rhos);
  if (dens_out && rad_orig) std::copy(rorig.begin(), rorig.end(), rad_orig);
  tm.total_sec = now_sec() - t_total;
  if (t) *t = tm;
  return 0;
}

int32_t orc_transform(const orc_params* p, const float* x_train, int64_t n, int64_t d,
                      const float* y_train, float a, float b, const float* q, int64_t m,
                      float* out) {
  if (m == 0) return 0;
  const int nt = p->n_threads > 0 ? p->n_threads : omp_get_max_threads();
  const int32_t k = std::min<int64_t>(p->n_neighbors, n);
  vector<int64_t> idx(static_cast<size_t>(m) * k);
  vector<float> dist(static_cast<size_t>(m) * k);
  knn_impl(q, m, x_train, n, d, k, p->metric, nt, idx.data(), dist.data(), p->metric_p);
  return orc_transform_with_knn(p, n, y_train, a, b, m, idx.data(), dist.data(), k, out);
}

/* transform_from_knn (lib.rs:1738-1780) on a query kNN [m x k] among the n
 * fitted rows (the sparse transforms, lib.rs:1655-1737, reach it this way). */
int32_t orc_transform_with_knn(const orc_params* p, int64_t n, const float* y_train, float a,
                               float b, int64_t m, const int64_t* idx, const float* dist,
                               int32_t k, float* out) {
  if (m == 0) return 0;
  const int nt = p->n_threads > 0 ? p->n_threads : omp_get_max_threads();
  const int32_t n_epochs = p->n_epochs > 0 ? p->n_epochs : (m <= 10000 ? 100 : 30);
  vector<float> sig(m), rho(m);
  const float lc = std::max(p->local_connectivity - 1.0f, 0.0f);
  smooth_knn(dist, m, k, static_cast<float>(k), lc, DEFAULT_BANDWIDTH, false, sig.data(),
             rho.data());
  auto edges = membership_bipartite(idx, dist, sig.data(), rho.data(), m, k);
  prune(edges, n_epochs);
  const int dim = p->n_components;
  init_graph_transform(m, dim, edges, y_train, out);
  optimize_layout_transform(out, m, dim, y_train, n, edges, n_epochs, a, b,
                            p->learning_rate / 4.0f, p->negative_sample_rate,
                            p->repulsion_strength, p->random_seed ^ 0xA24BAED40B53C117ULL, nt);
  return 0;
}

/* inverse_transform_rows (lib.rs:1800-1915); optimize = 0 stops after
 * init_inverse_points. */
int32_t orc_inverse_transform(const orc_params* p, const float* x_train, int64_t n, int64_t d,
                              const float* y_train, const float* sigmas, const float* rhos,
                              float a, float b, const float* q, int64_t m, int32_t optimize,
                              float* out) {
  if (m == 0) return 0;
  const int nt = p->n_threads > 0 ? p->n_threads : omp_get_max_threads();
  const int dim = p->n_components;
  const int32_t k = inverse_neighbor_count(p->n_neighbors, n);
  const int32_t n_epochs = p->n_epochs > 0 ? std::max(p->n_epochs / 3, 1) : (m <= 10000 ? 100 : 30);
  vector<int64_t> idx(static_cast<size_t>(m) * k);
  vector<float> dist(static_cast<size_t>(m) * k);
  knn_impl(q, m, y_train, n, dim, k, 0, nt, idx.data(), dist.data());
  vector<float> raw, norm;
  vector<char> fixed;
  inverse_weights(dist.data(), m, k, a, b, raw, norm, fixed);
  init_inverse_points(q, m, dim, idx.data(), norm, fixed, k, y_train, x_train, d, out);
  if (!optimize) return 0;
  vector<Edge> edges;
  for (int64_t i = 0; i < m; ++i)
    for (int32_t j = 0; j < k; ++j)
      if (raw[i * k + j] > 0.0f) edges.push_back({i, idx[i * k + j], raw[i * k + j]});
  optimize_layout_inverse(out, m, x_train, n, d, edges, fixed, sigmas, rhos, n_epochs,
                          p->learning_rate / 4.0f, p->negative_sample_rate,
                          p->repulsion_strength, p->random_seed ^ 0xD1B54A32D192ED03ULL);
  return 0;
}

}  // extern "C"
