// Host orchestrator and C ABI (include/umapx.h).
//
// Mirrors UmapModel (rust_umap/src/lib.rs:570-1966): validation with the
// reference's messages, the fixed stage order of
// build_dense_fit_artifacts_with_target (lib.rs:1367-1464), the fitted state
// kept for transform, and transform_rows / transform_from_knn
// (lib.rs:1604-1780).  All stage data stays resident on the GPU between
// stages; only the small sequential pieces (curve fit, Ritz/eigen solves,
// random init) run on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/umapx.h"
#include "common.hpp"
#include "graph.hpp"
#include "hostmath.hpp"
#include "inverse.hpp"
#include "knn.hpp"
#include "mlp.hpp"
#include "sgd.hpp"
#include "spectral.hpp"

namespace umapx {

void check_hip(hipError_t e, const char* what, const char* file, int line) {
  if (e == hipSuccess) return;
  std::string msg = std::string("device error: ") + hipGetErrorString(e) + " (" + what + " at " +
                    file + ":" + std::to_string(line) + ")";
  throw Error(kDeviceError, msg);
}

void retain_pool_memory() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  hipMemPool_t pool;
  HIP_CHECK(hipDeviceGetDefaultMemPool(&pool, dev));
  uint64_t keep = UINT64_MAX;
  HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
}

namespace {
thread_local std::string g_last_error;
struct LastErrorDetail {
  int32_t code = 0;
  int64_t row = -1, expected = -1, got = -1;
};
thread_local LastErrorDetail g_last_detail;
}  // namespace

int32_t set_last_error(const Error& e) {
  g_last_error = e.what();
  g_last_detail = LastErrorDetail{e.code, e.row, e.expected, e.got};
  return e.code;
}

void maybe_inject_fault(const char* stage) {
  static const char* env = std::getenv("UMAPX_FAULT_INJECT");
  if (env && std::strcmp(env, stage) == 0)
    throw Error(kDeviceError, std::string("device error: injected fault in stage ") + stage);
}

namespace {

double now_sec() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

std::string fmt_f32(float v) {
  // Rust's Display for f32: shortest round-trip digits, never exponent form.
  char buf[128];
  auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::fixed);
  return std::string(buf, r.ptr);
}

// The reference's Debug names of Metric (lib.rs:89-101).
std::string metric_debug(int m, float p) {
  switch (m) {
    case UMAPX_METRIC_EUCLIDEAN: return "Euclidean";
    case UMAPX_METRIC_COSINE: return "Cosine";
    case UMAPX_METRIC_MANHATTAN: return "Manhattan";
    case UMAPX_METRIC_CHEBYSHEV: return "Chebyshev";
    case UMAPX_METRIC_MINKOWSKI: {
      char buf[64];
      std::snprintf(buf, sizeof(buf), "Minkowski { p: %g%s }", static_cast<double>(p),
                    std::floor(p) == p ? ".0" : "");
      return buf;
    }
    case UMAPX_METRIC_CORRELATION: return "Correlation";
    case UMAPX_METRIC_CANBERRA: return "Canberra";
    case UMAPX_METRIC_BRAYCURTIS: return "BrayCurtis";
    default: return "Unsupported";
  }
}

// validate_metric (lib.rs:2230-2239) plus the codes this library knows.
void validate_metric(int metric, float p) {
  if (metric < UMAPX_METRIC_EUCLIDEAN || metric > UMAPX_METRIC_BRAYCURTIS)
    throw_invalid("unknown metric code " + std::to_string(metric));
  if (metric == UMAPX_METRIC_MINKOWSKI && (!std::isfinite(p) || p <= 0.0f))
    throw_invalid("minkowski metric requires finite p > 0");
}

// validate_data (lib.rs:2113-2138) on host rows.
void validate_data_host(const float* x, int64_t n, int64_t d) {
  if (n == 0) throw Error(kEmptyData, "input data is empty");
  if (n < 2) throw Error(kNeedTwoSamples, "need at least two samples");
  if (d == 0) throw_invalid("input rows must have at least one feature");
  for (int64_t r = 0; r < n; ++r)
    for (int64_t c = 0; c < d; ++c)
      if (!std::isfinite(x[r * d + c]))
        throw_invalid("input contains non-finite value at row " + std::to_string(r));
}

__global__ void first_nonfinite_kernel(const float* __restrict__ x, int64_t total, int64_t d,
                                       unsigned long long* __restrict__ first_row) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  if (!isfinite(x[i])) atomicMin(first_row, static_cast<unsigned long long>(i / d));
}

// Rejects the first row of x [n x d] (device) holding a non-finite value.
void check_finite_device(const float* x, int64_t n, int64_t d, hipStream_t s) {
  if (n == 0 || d == 0) return;
  DevBuf<unsigned long long> first(1);
  const unsigned long long init = std::numeric_limits<unsigned long long>::max();
  HIP_CHECK(hipMemcpyAsync(first.get(), &init, sizeof(init), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(first_nonfinite_kernel, dim3(blocks_for(n * d, 256)), dim3(256), 0, s, x,
                     n * d, d, first.get());
  KERNEL_CHECK();
  unsigned long long row = init;
  HIP_CHECK(hipMemcpyAsync(&row, first.get(), sizeof(row), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (row != init) throw_invalid("input contains non-finite value at row " + std::to_string(row));
}

void validate_data_device(const float* x, int64_t n, int64_t d, hipStream_t s) {
  if (n == 0) throw Error(kEmptyData, "input data is empty");
  if (n < 2) throw Error(kNeedTwoSamples, "need at least two samples");
  if (d == 0) throw_invalid("input rows must have at least one feature");
  check_finite_device(x, n, d, s);
}

// validate_params (lib.rs:2140-2228) for the dense path.
void validate_params(const umapx_params& p, int64_t n_samples) {
  validate_metric(p.metric, p.metric_p);
  if (p.n_neighbors < 2) throw_invalid("n_neighbors must be >= 2");
  if (p.n_neighbors >= n_samples)
    throw_invalid("n_neighbors (" + std::to_string(p.n_neighbors) +
                  ") must be smaller than number of samples (" + std::to_string(n_samples) + ")");
  if (p.n_components == 0 || p.n_components < 0) throw_invalid("n_components must be >= 1");
  // checked before any device work: the GPU layout kernels hold at most 8
  // coordinates per row
  if (p.n_components > 8) throw_invalid("n_components must be <= 8 on the GPU path");
  if (!std::isfinite(p.learning_rate) || p.learning_rate <= 0.0f)
    throw_invalid("learning_rate must be finite and > 0");
  if (!std::isfinite(p.min_dist) || p.min_dist < 0.0f) throw_invalid("min_dist must be finite and >= 0");
  if (!std::isfinite(p.spread) || p.spread <= 0.0f) throw_invalid("spread must be finite and > 0");
  if (p.min_dist > p.spread) throw_invalid("min_dist must be <= spread");
  if (!std::isfinite(p.set_op_mix_ratio) || !(p.set_op_mix_ratio >= 0.0f && p.set_op_mix_ratio <= 1.0f))
    throw_invalid("set_op_mix_ratio must be finite and in [0, 1]");
  if (!std::isfinite(p.repulsion_strength) || p.repulsion_strength < 0.0f)
    throw_invalid("repulsion_strength must be finite and >= 0");
  if (p.negative_sample_rate <= 0) throw_invalid("negative_sample_rate must be >= 1");
  if (!std::isfinite(p.local_connectivity) || p.local_connectivity < 0.0f)
    throw_invalid("local_connectivity must be finite and >= 0");
  if (p.approx_knn_candidates <= 0) throw_invalid("approx_knn_candidates must be >= 1");
  if (p.approx_knn_iters <= 0) throw_invalid("approx_knn_iters must be >= 1");
  if (p.approx_knn_threshold < 0) throw_invalid("approx_knn_threshold must be >= 0");
  if (!std::isfinite(p.dens_lambda) || p.dens_lambda < 0.0f) throw_invalid("dens_lambda must be finite and >= 0");
  if (!std::isfinite(p.dens_frac) || !(p.dens_frac >= 0.0f && p.dens_frac <= 1.0f))
    throw_invalid("dens_frac must be finite and in [0, 1]");
  if (!std::isfinite(p.dens_var_shift) || p.dens_var_shift <= 0.0f)
    throw_invalid("dens_var_shift must be finite and > 0");
  if (p.knn_precision != UMAPX_KNN_FP32 && p.knn_precision != UMAPX_KNN_BF16_RERANK)
    throw_invalid("knn_precision must be fp32 or bf16_rerank");
  if (p.rerank_factor < 1) throw_invalid("rerank_factor must be >= 1");
}

// validate_sparse_metric and validate_sparse_density_params
// (lib.rs:2241-2258), plus what the GPU path holds for sparse fits.
void validate_sparse_params(const umapx_params& p) {
  if (p.metric == UMAPX_METRIC_CANBERRA || p.metric == UMAPX_METRIC_BRAYCURTIS)
    throw_invalid(metric_debug(p.metric, p.metric_p) +
                  " metric is currently supported for dense input only");
  if (p.densmap || p.output_dens)
    throw_invalid("densmap and output_dens are currently supported for dense input only");
  if (p.transform_mode == UMAPX_TRANSFORM_LEARNED)
    throw_invalid("the learned transform supports dense input only");
}

// should_use_approximate_knn (lib.rs:2285-2311, constants lib.rs:24-26):
// auto mode switches to the approximate rounds only past 4 x the threshold
// (16 features or fewer: the same multiplier); a threshold of 0 is the
// forced "approximate" mode.
bool should_use_approximate_knn(const umapx_params& p, int64_t n_samples, int64_t n_features) {
  constexpr int64_t kLowDimMaxFeatures = 16, kLowDimMultiplier = 4, kGeneralMultiplier = 4;
  const int64_t threshold = std::max<int64_t>(p.approx_knn_threshold, 0);
  if (!p.use_approximate_knn || n_samples <= threshold) return false;
  // the GPU's approximate rounds hold at most 32 neighbours: auto mode keeps
  // the exact kNN above that (a forced approximate run reports the limit)
  if (p.n_neighbors > 32 && threshold > 0) return false;
  if (threshold > 0) {
    const int64_t mult = n_features <= kLowDimMaxFeatures ? kLowDimMultiplier : kGeneralMultiplier;
    if (n_samples <= threshold * mult) return false;
  }
  return true;
}

int32_t default_epochs(const umapx_params& p, int64_t n) {
  return p.n_epochs > 0 ? p.n_epochs : (n <= 10000 ? 500 : 200);
}

// density_start_epoch (lib.rs:5054-5055)
int32_t density_start(int32_t n_epochs, float frac) {
  const int64_t active = static_cast<int64_t>(std::ceil(static_cast<float>(n_epochs) * frac));
  return static_cast<int32_t>(std::max<int64_t>(0, n_epochs - std::max<int64_t>(active, 1)));
}

// validate_precomputed_array_shapes + value checks + i64 conversion +
// validate_and_trim_precomputed_knn (lib.rs:1968-2111).
void validate_precomputed(const int64_t* idx, int64_t ir, int64_t ic, const float* dist, int64_t dr,
                          int64_t dc, int64_t n, int k, bool check_values,
                          std::vector<int32_t>& out_idx, std::vector<float>& out_dist) {
  if (ir != dr || ic != dc) throw_invalid("knn_indices and knn_dists must have identical shapes");
  if (ir != n) throw_invalid("knn row count must match data row count");
  if (ic < k) throw_invalid("knn columns must be >= n_neighbors (" + std::to_string(k) + ")");
  if (check_values)
    for (int64_t i = 0; i < dr * dc; ++i) {
      if (!std::isfinite(dist[i])) throw_invalid("knn_dists must contain only finite values");
      if (dist[i] < 0.0f) throw_invalid("knn_dists must be non-negative");
    }
  for (int64_t i = 0; i < ir * ic; ++i)
    if (idx[i] < 0) throw_invalid("knn indices must be non-negative integers");
  out_idx.resize(static_cast<size_t>(n) * k);
  out_dist.resize(static_cast<size_t>(n) * k);
  std::unordered_set<int64_t> seen;
  seen.reserve(2 * k);
  for (int64_t r = 0; r < n; ++r) {
    seen.clear();
    float prev = -std::numeric_limits<float>::infinity();
    for (int c = 0; c < k; ++c) {
      const int64_t j = idx[r * ic + c];
      const float dv = dist[r * dc + c];
      if (j >= n)
        throw_invalid("precomputed knn index out of range at row " + std::to_string(r) + ", col " +
                      std::to_string(c));
      if (!std::isfinite(dv) || dv < 0.0f)
        throw_invalid("precomputed knn distance must be finite and >= 0 at row " + std::to_string(r) +
                      ", col " + std::to_string(c));
      if (!seen.insert(j).second)
        throw_invalid("precomputed knn row " + std::to_string(r) + " contains duplicate index " +
                      std::to_string(j));
      if (c > 0 && dv + 1e-5f < prev)
        throw_invalid("precomputed knn distances must be non-decreasing within each row; row " +
                      std::to_string(r) + " has dist[" + std::to_string(c) + "]=" + fmt_f32(dv) +
                      " < dist[" + std::to_string(c - 1) + "]=" + fmt_f32(prev));
      prev = dv;
      out_idx[r * k + c] = static_cast<int32_t>(j);
      out_dist[r * k + c] = dv;
    }
  }
}

struct StageClock {
  hipStream_t s;
  double t0;
  explicit StageClock(hipStream_t st) : s(st), t0(now_sec()) {}
  double lap() {
    HIP_CHECK(hipStreamSynchronize(s));
    const double t = now_sec();
    const double dt = t - t0;
    t0 = t;
    return dt;
  }
};

}  // namespace
}  // namespace umapx

using namespace umapx;

struct umapx_handle {
  // Declared first so that it is destroyed last: the buffers below are
  // stream-ordered allocations on this stream and are freed on it.
  struct StreamOwner {
    hipStream_t s = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    ~StreamOwner() {
      if (s) (void)hipStreamSynchronize(s);
      for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
      if (s) (void)hipStreamDestroy(s);
    }
  } owner;
  hipStream_t& stream = owner.s;
  hipEvent_t (&ev)[4] = owner.ev;
  umapx_params p{};
  float a = 0.f, b = 0.f;
  bool fitted = false;
  int64_t n = 0, d = 0;
  int k_fit = 0;
  DevBuf<float> x, y, sig, rho, rad_orig, rad_emb, knn_dist;
  bool sparse = false;  // fitted on CSR rows (held in csr instead of x)
  CsrDevice csr;
  DevBuf<int32_t> knn_idx;
  KnnWorkspace kws;
  GraphWorkspace gws;
  SgdWorkspace sws;
  SpectralWorkspace spws;
  Graph graph;
  MlpModel mlp;
  double learned_kernel_ms = 0.0;  // HIP-event time of the last learned transform's kernels
  umapx_timings timings{};
  bool count_schedule = false;  // set by the profile entry point

  void ensure_stream() {
    if (!stream) {
      HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
      retain_pool_memory();
    }
  }

  std::pair<float, float> resolve_ab() {
    if (!(a > 0.f && b > 0.f)) find_ab_params(p.spread, p.min_dist, &a, &b);
    return {a, b};
  }

  // build_dense_fit_artifacts_with_target (lib.rs:1367-1464) on device data.
  // d_x stays valid for the whole call; pre_idx/pre_dist are device kNN when
  // the caller supplied it (already validated and trimmed to k columns).
  // csr_in: a sparse fit (build_sparse_fit_artifacts, lib.rs:1491-1550); d_x is
  // then null, the kNN is the sparse one and the rows are kept as CSR.
  void fit_device(const float* d_x, int64_t n_, int64_t d_, const int32_t* pre_idx,
                  const float* pre_dist, bool validated_on_host, double t_start,
                  CsrDevice* csr_in = nullptr) {
    ensure_stream();
    hipStream_t s = stream;
    AllocStreamScope alloc_scope(s);
    umapx_timings t{};
    StageClock clk(s);
    if (!validated_on_host) validate_data_device(d_x, n_, d_, s);
    validate_params(p, n_);
    t.validation_sec = clk.lap() + (clk.t0 - t_start > 0 ? 0.0 : 0.0);
    const int k = p.n_neighbors;
    DevBuf<int32_t> idx;
    DevBuf<float> dist;
    KnnStats kst;
    kst.cand_start = ev[0];
    kst.cand_stop = ev[1];
    if (csr_in) {  // exact_nearest_neighbors over the CSR rows (sparse.rs:605-626)
      idx.alloc(static_cast<size_t>(n_) * k);
      dist.alloc(static_cast<size_t>(n_) * k);
      csr_knn_device(*csr_in, *csr_in, true, k, p.metric, p.metric_p, idx.get(), dist.get(), s);
    } else if (pre_idx == nullptr) {
      idx.alloc(static_cast<size_t>(n_) * k);
      dist.alloc(static_cast<size_t>(n_) * k);
      const bool approx = should_use_approximate_knn(p, n_, d_);
      if (approx)
        knn_approx_device(d_x, n_, d_, k, p.metric, p.metric_p, p.approx_knn_candidates,
                          p.approx_knn_iters, p.random_seed ^ 0xC0FEFEED1234ABCDULL, idx.get(),
                          dist.get(), kws, s);
      else
        knn_device(d_x, n_, d_x, n_, d_, k, p.metric, p.knn_precision, p.rerank_factor, idx.get(),
                   dist.get(), kws, s, &kst, p.metric_p);
      t.used_approximate_knn = approx ? 1 : 0;
    }
    t.knn_sec = clk.lap();
    if (kst.recorded) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, kst.cand_start, kst.cand_stop));
      t.knn_kernel_ms = ms;
      t.knn_flops = kst.flops;
    }
    const int32_t n_epochs = default_epochs(p, n_);
    resolve_ab();
    t.curve_params_sec = clk.lap();
    const int32_t* kidx = pre_idx ? pre_idx : idx.get();
    const float* kdist = pre_dist ? pre_dist : dist.get();
    t.knn_validate_trim_sec = clk.lap();
    DevBuf<float> sg(n_), rh(n_);
    const bool include_self = knn_include_self_device(kidx, kdist, n_, k, gws, s);
    smooth_knn_dist_device(kdist, n_, k, static_cast<float>(k), p.local_connectivity, 1.0f,
                           include_self, p.smooth_fp64 != 0, sg.get(), rh.get(), gws, s);
    t.smooth_knn_sec = clk.lap();
    const bool dens_out = p.densmap || p.output_dens;
    DevBuf<float> ro;
    if (dens_out) {
      ro.alloc(n_);
      local_radii_device(kidx, kdist, n_, k, p.dens_var_shift, ro.get(), s);
    }
    const bool dens_active = p.densmap && p.dens_lambda > 0.0f;
    t.density_original_sec = clk.lap();
    Graph g;
    fuzzy_graph_device(kidx, kdist, sg.get(), rh.get(), n_, k, p.set_op_mix_ratio, n_epochs, g, gws, s);
    t.membership_sec = clk.lap();
    t.symmetrize_sec = 0.0;
    t.prune_sec = 0.0;
    DevBuf<float> yy(static_cast<size_t>(n_) * p.n_components);
    SpectralReport rep;
    bool spectral_ok = false;
    if (p.init == UMAPX_INIT_SPECTRAL)
      spectral_ok = spectral_init_device(d_x, n_, d_, p.n_components, g, p.random_seed, yy.get(),
                                         spws, s, &rep, csr_in);
    if (!spectral_ok) {
      auto r = random_init(n_, p.n_components, p.random_seed, -10.0f, 10.0f);
      HIP_CHECK(hipMemcpyAsync(yy.get(), r.data(), 4 * r.size(), hipMemcpyHostToDevice, s));
      HIP_CHECK(hipStreamSynchronize(s));
    }
    t.init_sec = clk.lap();
    SgdArgs sa;
    sa.n_epochs = n_epochs;
    sa.a = a;
    sa.b = b;
    sa.learning_rate = p.learning_rate;
    sa.negative_sample_rate = p.negative_sample_rate;
    sa.repulsion_strength = p.repulsion_strength;
    sa.seed = p.random_seed ^ 0x9E3779B97F4A7C15ULL;
    sa.symmetric = true;  // the fuzzy union: sorted rows, every edge has its reverse
    if (dens_active) {
      sa.radii = ro.get();
      sa.dens_start_epoch = density_start(n_epochs, p.dens_frac);
      sa.dens_lambda = p.dens_lambda;
      sa.dens_var_shift = p.dens_var_shift;
    }
    SgdStats sst;
    sst.start = ev[2];
    sst.stop = ev[3];
    optimize_layout_device(yy.get(), n_, p.n_components, g, sa, sws, s, &sst);
    t.optimize_sec = clk.lap();
    if (sst.recorded) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, sst.start, sst.stop));
      t.sgd_kernel_ms = ms;
    }
    if (count_schedule) {
      // outside optimize_sec; reads only the schedule the epochs used
      schedule_totals_device(sws, g.n_edges, n_epochs, p.negative_sample_rate, s,
                             &t.sgd_positive_updates, &t.sgd_negative_samples);
      clk.lap();
    }
    DevBuf<float> re;
    if (dens_out) {
      re.alloc(n_);
      embedding_radii_device(yy.get(), n_, p.n_components, kidx, k, p.dens_var_shift, re.get(), s);
    }
    t.density_embedding_sec = clk.lap();
    t.output_copy_sec = 0.0;
    // store_dense_fit_state / store_sparse_fit_state (lib.rs:1551-1590)
    if (csr_in) {
      csr = std::move(*csr_in);
      x.release();
    } else {
      x.alloc(static_cast<size_t>(n_) * d_);
      HIP_CHECK(hipMemcpyAsync(x.get(), d_x, 4 * n_ * d_, hipMemcpyDeviceToDevice, s));
      csr = CsrDevice{};
    }
    sparse = csr_in != nullptr;
    knn_idx.alloc(static_cast<size_t>(n_) * k);
    knn_dist.alloc(static_cast<size_t>(n_) * k);
    HIP_CHECK(hipMemcpyAsync(knn_idx.get(), kidx, 4 * n_ * k, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipMemcpyAsync(knn_dist.get(), kdist, 4 * n_ * k, hipMemcpyDeviceToDevice, s));
    y = std::move(yy);
    sig = std::move(sg);
    rho = std::move(rh);
    rad_orig = std::move(ro);
    rad_emb = std::move(re);
    graph = std::move(g);
    n = n_;
    d = d_;
    k_fit = k;
    fitted = true;
    mlp.trained = false;
    t.store_state_sec = clk.lap();
    t.n_edges = graph.n_edges;
    t.n_epochs = n_epochs;
    t.init_used_spectral = spectral_ok ? 1 : 0;
    t.n_graph_components = static_cast<int32_t>(rep.n_components);
    if (p.transform_mode == UMAPX_TRANSFORM_LEARNED) {
      mlp_train(mlp, x.get(), n, d, y.get(), p.n_components, graph, a, b, p, s);
      HIP_CHECK(hipStreamSynchronize(s));
    }
    t.total_sec = now_sec() - t_start;
    timings = t;
  }

  void fit_host(const float* hx, int64_t n_, int64_t d_, const int32_t* pre_idx_h,
                const float* pre_dist_h, float* y_out) {
    const double t0 = now_sec();
    validate_data_host(hx, n_, d_);
    validate_params(p, n_);
    DevBuf<float> dx(static_cast<size_t>(n_) * d_);
    ensure_stream();
    HIP_CHECK(hipMemcpyAsync(dx.get(), hx, 4 * n_ * d_, hipMemcpyHostToDevice, stream));
    DevBuf<int32_t> di;
    DevBuf<float> dd;
    if (pre_idx_h) {
      di.alloc(static_cast<size_t>(n_) * p.n_neighbors);
      dd.alloc(static_cast<size_t>(n_) * p.n_neighbors);
      HIP_CHECK(hipMemcpyAsync(di.get(), pre_idx_h, 4 * di.size(), hipMemcpyHostToDevice, stream));
      HIP_CHECK(hipMemcpyAsync(dd.get(), pre_dist_h, 4 * dd.size(), hipMemcpyHostToDevice, stream));
    }
    fit_device(dx.get(), n_, d_, pre_idx_h ? di.get() : nullptr, pre_idx_h ? dd.get() : nullptr, true, t0);
    if (y_out) {
      const double s0 = now_sec();
      HIP_CHECK(hipMemcpyAsync(y_out, y.get(), 4 * n * p.n_components, hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      timings.output_copy_sec = now_sec() - s0;
      timings.total_sec = now_sec() - t0;
    }
  }

  // fit_sparse_csr / fit_transform_sparse_csr (lib.rs:998-1036) on host CSR
  // arrays (SparseCsrMatrix::new's checks first, as the reference's caller
  // builds the matrix before the fit).
  void fit_csr_host(int64_t rows, int64_t cols, const int64_t* indptr, const int64_t* indices,
                    const float* data, int64_t nnz, float* y_out) {
    const double t0 = now_sec();
    if (!indptr || (nnz > 0 && (!indices || !data))) throw_invalid("null argument");
    csr_validate_host(rows, cols, indptr, rows + 1, indices, data, nnz);
    if (rows == 0) throw Error(kEmptyData, "input data is empty");
    if (rows < 2) throw Error(kNeedTwoSamples, "need at least two samples");
    validate_params(p, rows);
    validate_sparse_params(p);
    ensure_stream();
    CsrDevice c;
    {
      AllocStreamScope alloc_scope(stream);
      csr_upload(c, rows, cols, indptr, indices, data, stream);
    }
    fit_device(nullptr, rows, cols, nullptr, nullptr, true, t0, &c);
    if (y_out) {
      const double s0 = now_sec();
      HIP_CHECK(hipMemcpyAsync(y_out, y.get(), 4 * n * p.n_components, hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      timings.output_copy_sec = now_sec() - s0;
      timings.total_sec = now_sec() - t0;
    }
  }

  // transform_rows + transform_from_knn (lib.rs:1604-1685, 1738-1780).
  void transform_plain_device(const float* d_q, int64_t m, float* d_out) {
    ensure_stream();
    hipStream_t s = stream;
    AllocStreamScope alloc_scope(s);
    const int k = static_cast<int>(std::min<int64_t>(p.n_neighbors, n));
    DevBuf<int32_t> idx(static_cast<size_t>(m) * k);
    DevBuf<float> dist(static_cast<size_t>(m) * k);
    if (sparse)  // dense queries against the sparse fit (sparse.rs:753-900)
      dense_to_csr_knn_device(d_q, m, csr, k, p.metric, p.metric_p, idx.get(), dist.get(), s);
    else
      knn_device(d_q, m, x.get(), n, d, k, p.metric, p.knn_precision, p.rerank_factor, idx.get(),
                 dist.get(), kws, s, nullptr, p.metric_p);
    transform_from_knn_device(idx.get(), dist.get(), m, k, d_out);
  }

  // transform_sparse_csr (lib.rs:1687-1737): CSR queries on a sparse fit.
  void transform_csr_host(int64_t m, int64_t cols, const int64_t* indptr, const int64_t* indices,
                          const float* data, int64_t nnz, float* out) {
    if (!fitted) throw Error(kNotFitted, "model is not fitted yet");
    if (cols != d)
      throw Error(kFeatureMismatch,
                  "feature mismatch: expected " + std::to_string(d) + ", got " + std::to_string(cols),
                  -1, d, cols);
    if (!sparse) throw_invalid("sparse query transform is supported only for sparse-fitted models");
    if (m == 0) return;
    csr_validate_host(m, cols, indptr, m + 1, indices, data, nnz);
    ensure_stream();
    hipStream_t s = stream;
    AllocStreamScope alloc_scope(s);
    CsrDevice qc;
    csr_upload(qc, m, cols, indptr, indices, data, s);
    const int k = static_cast<int>(std::min<int64_t>(p.n_neighbors, n));
    DevBuf<int32_t> idx(static_cast<size_t>(m) * k);
    DevBuf<float> dist(static_cast<size_t>(m) * k);
    csr_knn_device(qc, csr, false, k, p.metric, p.metric_p, idx.get(), dist.get(), s);
    DevBuf<float> dout(static_cast<size_t>(m) * p.n_components);
    transform_from_knn_device(idx.get(), dist.get(), m, k, dout.get());
    HIP_CHECK(hipMemcpyAsync(out, dout.get(), 4 * m * p.n_components, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }

  // transform_from_knn (lib.rs:1738-1780) on a device query kNN [m x k].
  void transform_from_knn_device(const int32_t* d_idx, const float* d_dist, int64_t m, int k,
                                 float* d_out) {
    hipStream_t s = stream;
    const int32_t n_epochs = p.n_epochs > 0 ? std::max(p.n_epochs, 1) : (m <= 10000 ? 100 : 30);
    DevBuf<float> sg(m), rh(m);
    const float lc = std::max(p.local_connectivity - 1.0f, 0.0f);
    smooth_knn_dist_device(d_dist, m, k, static_cast<float>(k), lc, 1.0f, false,
                           p.smooth_fp64 != 0, sg.get(), rh.get(), gws, s);
    Graph g;
    bipartite_graph_device(d_idx, d_dist, sg.get(), rh.get(), m, k, n_epochs, g, gws, s);
    init_graph_transform_device(m, p.n_components, g, y.get(), d_out, s);
    SgdArgs sa;
    sa.n_epochs = n_epochs;
    sa.a = a;
    sa.b = b;
    sa.learning_rate = p.learning_rate / 4.0f;
    sa.negative_sample_rate = p.negative_sample_rate;
    sa.repulsion_strength = p.repulsion_strength;
    sa.seed = p.random_seed ^ 0xA24BAED40B53C117ULL;
    sa.max_row_edges = k;  // the bipartite graph keeps at most k edges per query
    SgdWorkspace ws;
    optimize_transform_device(d_out, m, p.n_components, y.get(), n, g, sa, ws, s);
    HIP_CHECK(hipStreamSynchronize(s));
  }

  void check_query(int64_t m, int64_t dq) {
    if (!fitted) throw Error(kNotFitted, "model is not fitted yet");
    if (m > 0 && dq != d)
      throw Error(kFeatureMismatch,
                  "feature mismatch: expected " + std::to_string(d) + ", got " + std::to_string(dq),
                  -1, d, dq);
  }

  void transform_host(const float* q, int64_t m, int64_t dq, float* out, int mode) {
    check_query(m, dq);
    if (m == 0) return;
    for (int64_t r = 0; r < m; ++r)
      for (int64_t c = 0; c < dq; ++c)
        if (!std::isfinite(q[r * dq + c]))
          throw_invalid("input contains non-finite value at row " + std::to_string(r));
    ensure_stream();
    DevBuf<float> dqb(static_cast<size_t>(m) * dq), dout(static_cast<size_t>(m) * p.n_components);
    HIP_CHECK(hipMemcpyAsync(dqb.get(), q, 4 * m * dq, hipMemcpyHostToDevice, stream));
    run_transform(dqb.get(), m, dout.get(), mode);
    HIP_CHECK(hipMemcpyAsync(out, dout.get(), 4 * m * p.n_components, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  // inverse_transform_rows (lib.rs:1800-1915) for device rows d_q [m x n_components].
  void inverse_device(const float* d_q, int64_t m, float* d_out) {
    ensure_stream();
    AllocStreamScope alloc_scope(stream);
    InverseArgs ia;
    ia.k = static_cast<int>(std::max<int64_t>(std::min<int64_t>(p.n_neighbors, n), 1));
    ia.n_epochs = p.n_epochs > 0 ? std::max(p.n_epochs / 3, 1) : (m <= 10000 ? 100 : 30);
    ia.a = a;
    ia.b = b;
    ia.learning_rate = p.learning_rate / 4.0f;
    ia.negative_sample_rate = p.negative_sample_rate;
    ia.repulsion_strength = p.repulsion_strength;
    ia.seed = p.random_seed ^ 0xD1B54A32D192ED03ULL;
    inverse_transform_device(d_q, m, p.n_components, y.get(), x.get(), n, d, sig.get(), rho.get(),
                             ia, d_out, kws, stream);
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  // lib.rs:1803-1847: fitted model, then per-row width and finiteness.
  void check_embedded(int64_t m, int64_t dq) {
    if (!fitted) throw Error(kNotFitted, "model is not fitted yet");
    if (sparse) throw_invalid("inverse_transform is not supported for sparse-trained models yet");
    if (m > 0 && dq != p.n_components)
      throw Error(kEmbeddingDimensionMismatch,
                  "embedding dimension mismatch: expected " + std::to_string(p.n_components) +
                      ", got " + std::to_string(dq),
                  -1, p.n_components, dq);
  }

  void inverse_host(const float* q, int64_t m, int64_t dq, float* out) {
    check_embedded(m, dq);
    if (m == 0) return;
    for (int64_t r = 0; r < m; ++r)
      for (int64_t c = 0; c < dq; ++c)
        if (!std::isfinite(q[r * dq + c]))
          throw_invalid("input contains non-finite value at row " + std::to_string(r));
    ensure_stream();
    AllocStreamScope alloc_scope(stream);
    DevBuf<float> dqb(static_cast<size_t>(m) * dq), dout(static_cast<size_t>(m) * d);
    HIP_CHECK(hipMemcpyAsync(dqb.get(), q, 4 * m * dq, hipMemcpyHostToDevice, stream));
    inverse_device(dqb.get(), m, dout.get());
    HIP_CHECK(hipMemcpyAsync(out, dout.get(), 4 * m * d, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  void run_transform(const float* d_q, int64_t m, float* d_out, int mode) {
    AllocStreamScope alloc_scope(stream);
    if (mode == UMAPX_TRANSFORM_LEARNED) {
      if (sparse) throw_invalid("the learned transform supports dense input only");
      if (!mlp.trained) {
        mlp_train(mlp, x.get(), n, d, y.get(), p.n_components, graph, a, b, p, stream);
      }
      hipEvent_t ev[2] = {nullptr, nullptr};
      HIP_CHECK(hipEventCreate(&ev[0]));
      struct EventGuard {
        hipEvent_t* e;
        ~EventGuard() {
          for (int i = 0; i < 2; ++i)
            if (e[i]) (void)hipEventDestroy(e[i]);
        }
      } ev_guard{ev};
      HIP_CHECK(hipEventCreate(&ev[1]));
      HIP_CHECK(hipEventRecord(ev[0], stream));
      mlp_forward_rows(mlp, d_q, m, d_out, stream);
      HIP_CHECK(hipEventRecord(ev[1], stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
      learned_kernel_ms = ms;
    } else {
      transform_plain_device(d_q, m, d_out);
    }
  }
};

namespace {

template <typename F>
int32_t guarded(F&& f) {
  try {
    f();
    return UMAPX_OK;
  } catch (const Error& e) {
    return set_last_error(e);
  } catch (const std::exception& e) {
    return set_last_error(Error(kDeviceError, std::string("device error: ") + e.what()));
  }
}

void require_handle(const umapx_handle* h) {
  if (!h) throw_invalid("null handle");
}

void idx64_from32(const std::vector<int32_t>& a, int64_t* out) {
  for (size_t i = 0; i < a.size(); ++i) out[i] = a[i];
}

}  // namespace

extern "C" {

void umapx_default_params(umapx_params* o) {
  std::memset(o, 0, sizeof(*o));
  o->n_neighbors = 15;
  o->n_components = 2;
  o->n_epochs = 0;
  o->metric = UMAPX_METRIC_EUCLIDEAN;
  o->learning_rate = 1.0f;
  o->min_dist = 0.1f;
  o->spread = 1.0f;
  o->local_connectivity = 1.0f;
  o->set_op_mix_ratio = 1.0f;
  o->repulsion_strength = 1.0f;
  o->negative_sample_rate = 5;
  o->init = UMAPX_INIT_SPECTRAL;
  o->random_seed = 42;
  o->use_approximate_knn = 1;
  o->approx_knn_candidates = 50;
  o->approx_knn_iters = 14;
  o->approx_knn_threshold = 4096;
  o->densmap = 0;
  o->dens_lambda = 2.0f;
  o->dens_frac = 0.3f;
  o->dens_var_shift = 0.1f;
  o->output_dens = 0;
  o->knn_precision = UMAPX_KNN_FP32;
  o->rerank_factor = 2;
  o->smooth_fp64 = 1;
  o->transform_mode = UMAPX_TRANSFORM_PLAIN;
  o->mlp_hidden = 256;
  o->mlp_steps = 1500;
  o->mlp_batch = 4096;
  o->mlp_negatives = 5;
  o->mlp_learning_rate = 1e-3f;
  o->mlp_embedding_weight = 1.0f;
  o->metric_p = 2.0f;
}

const char* umapx_last_error(void) { return g_last_error.c_str(); }

int32_t umapx_last_error_detail(int64_t* row, int64_t* expected, int64_t* got) {
  if (row) *row = g_last_detail.row;
  if (expected) *expected = g_last_detail.expected;
  if (got) *got = g_last_detail.got;
  return g_last_detail.code;
}

int32_t umapx_uses_approximate_knn(const umapx_params* p, int64_t n_samples, int64_t n_features) {
  return p && should_use_approximate_knn(*p, n_samples, n_features) ? 1 : 0;
}

const char* umapx_version(void) { return "umapx 0.1.0 (gfx950)"; }

int32_t umapx_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

int32_t umapx_create(const umapx_params* params, umapx_handle** out) {
  return guarded([&] {
    if (!out) throw_invalid("null output handle");
    auto* h = new umapx_handle();
    if (params)
      h->p = *params;
    else
      umapx_default_params(&h->p);
    *out = h;
  });
}

void umapx_destroy(umapx_handle* h) { delete h; }

int32_t umapx_fit(umapx_handle* h, const float* x, int64_t n, int64_t d) {
  return guarded([&] {
    require_handle(h);
    h->fit_host(x, n, d, nullptr, nullptr, nullptr);
  });
}

int32_t umapx_fit_transform(umapx_handle* h, const float* x, int64_t n, int64_t d, float* y_out) {
  return guarded([&] {
    require_handle(h);
    h->fit_host(x, n, d, nullptr, nullptr, y_out);
  });
}

int32_t umapx_fit_transform_device(umapx_handle* h, const float* d_x, int64_t n, int64_t d,
                                   float* d_y_out) {
  return guarded([&] {
    require_handle(h);
    const double t0 = now_sec();
    h->fit_device(d_x, n, d, nullptr, nullptr, false, t0);
    if (d_y_out) {
      HIP_CHECK(hipMemcpyAsync(d_y_out, h->y.get(), 4 * n * h->p.n_components, hipMemcpyDeviceToDevice,
                               h->stream));
      HIP_CHECK(hipStreamSynchronize(h->stream));
    }
    h->timings.total_sec = now_sec() - t0;
  });
}

int32_t umapx_fit_transform_with_knn(umapx_handle* h, const float* x, int64_t n, int64_t d,
                                     const int64_t* knn_idx, int64_t idx_rows, int64_t idx_cols,
                                     const float* knn_dist, int64_t dist_rows, int64_t dist_cols,
                                     int32_t knn_metric, int32_t validate_precomputed_values,
                                     float* y_out) {
  return guarded([&] {
    require_handle(h);
    if (knn_metric != h->p.metric)
      throw_invalid("precomputed knn metric (" + metric_debug(knn_metric, h->p.metric_p) +
                    ") must match model metric (" + metric_debug(h->p.metric, h->p.metric_p) + ")");
    validate_data_host(x, n, d);
    validate_params(h->p, n);
    std::vector<int32_t> ti;
    std::vector<float> td;
    validate_precomputed(knn_idx, idx_rows, idx_cols, knn_dist, dist_rows, dist_cols, n,
                         h->p.n_neighbors, validate_precomputed_values != 0, ti, td);
    h->fit_host(x, n, d, ti.data(), td.data(), y_out);
  });
}

int32_t umapx_profile_fit_transform(umapx_handle* h, const float* x, int64_t n, int64_t d,
                                    float* y_out, umapx_timings* t) {
  return guarded([&] {
    require_handle(h);
    struct Flag {
      bool& f;
      ~Flag() { f = false; }
    } reset{h->count_schedule};
    h->count_schedule = true;
    h->fit_host(x, n, d, nullptr, nullptr, y_out);
    if (t) *t = h->timings;
  });
}

int32_t umapx_fit_transform_csr(umapx_handle* h, int64_t n_rows, int64_t n_cols,
                                const int64_t* indptr, const int64_t* indices, const float* data,
                                int64_t nnz, float* y_out) {
  return guarded([&] {
    require_handle(h);
    h->fit_csr_host(n_rows, n_cols, indptr, indices, data, nnz, y_out);
  });
}

int32_t umapx_transform_csr(umapx_handle* h, int64_t m, int64_t n_cols, const int64_t* indptr,
                            const int64_t* indices, const float* data, int64_t nnz, float* out) {
  return guarded([&] {
    require_handle(h);
    if (m > 0 && (!indptr || !out || (nnz > 0 && (!indices || !data))))
      throw_invalid("null argument");
    h->transform_csr_host(m, n_cols, indptr, indices, data, nnz, out);
  });
}

int32_t umapx_is_sparse_fit(const umapx_handle* h, int32_t* sparse) {
  return guarded([&] {
    require_handle(h);
    if (!h->fitted) throw Error(kNotFitted, "model is not fitted yet");
    *sparse = h->sparse ? 1 : 0;
  });
}

int32_t umapx_get_timings(const umapx_handle* h, umapx_timings* t) {
  return guarded([&] {
    require_handle(h);
    *t = h->timings;
  });
}

int32_t umapx_transform(umapx_handle* h, const float* q, int64_t m, int64_t d, float* out) {
  return guarded([&] {
    require_handle(h);
    h->transform_host(q, m, d, out, h->p.transform_mode);
  });
}

int32_t umapx_transform_plain(umapx_handle* h, const float* q, int64_t m, int64_t d, float* out) {
  return guarded([&] {
    require_handle(h);
    h->transform_host(q, m, d, out, UMAPX_TRANSFORM_PLAIN);
  });
}

int32_t umapx_transform_learned(umapx_handle* h, const float* q, int64_t m, int64_t d, float* out) {
  return guarded([&] {
    require_handle(h);
    h->transform_host(q, m, d, out, UMAPX_TRANSFORM_LEARNED);
  });
}

int32_t umapx_transform_device(umapx_handle* h, const float* d_q, int64_t m, int64_t d, float* d_out) {
  return guarded([&] {
    require_handle(h);
    h->check_query(m, d);
    if (m == 0) return;
    h->ensure_stream();
    check_finite_device(d_q, m, d, h->stream);
    h->run_transform(d_q, m, d_out, h->p.transform_mode);
  });
}

int32_t umapx_inverse_transform(umapx_handle* h, const float* q, int64_t m, int64_t dim,
                                float* out) {
  return guarded([&] {
    require_handle(h);
    if (m > 0 && (!q || !out)) throw_invalid("null argument");
    h->inverse_host(q, m, dim, out);
  });
}

int32_t umapx_inverse_transform_device(umapx_handle* h, const float* d_q, int64_t m, int64_t dim,
                                       float* d_out) {
  return guarded([&] {
    require_handle(h);
    h->check_embedded(m, dim);
    if (m == 0) return;
    h->ensure_stream();
    check_finite_device(d_q, m, dim, h->stream);
    h->inverse_device(d_q, m, d_out);
  });
}

int32_t umapx_inverse_init(const float* x_train, int64_t n, int64_t d, const float* y_train,
                           int32_t dim, int32_t k, float a, float b, const float* q, int64_t m,
                           float* out) {
  return guarded([&] {
    if (m == 0) return;
    if (!x_train || !y_train || !q || !out) throw_invalid("null argument");
    if (n < 1 || d < 1 || dim < 1) throw_invalid("need n >= 1, d >= 1 and dim >= 1");
    umapx_handle::StreamOwner so;
    HIP_CHECK(hipStreamCreateWithFlags(&so.s, hipStreamNonBlocking));
    {
      AllocStreamScope alloc_scope(so.s);
      DevBuf<float> dx(static_cast<size_t>(n) * d), dy(static_cast<size_t>(n) * dim),
          dq(static_cast<size_t>(m) * dim), dout(static_cast<size_t>(m) * d);
      HIP_CHECK(hipMemcpyAsync(dx.get(), x_train, 4 * n * d, hipMemcpyHostToDevice, so.s));
      HIP_CHECK(hipMemcpyAsync(dy.get(), y_train, 4 * n * dim, hipMemcpyHostToDevice, so.s));
      HIP_CHECK(hipMemcpyAsync(dq.get(), q, 4 * m * dim, hipMemcpyHostToDevice, so.s));
      InverseArgs ia;
      ia.k = static_cast<int>(std::max<int64_t>(std::min<int64_t>(k, n), 1));
      ia.n_epochs = 0;
      ia.a = a;
      ia.b = b;
      KnnWorkspace kws;
      inverse_transform_device(dq.get(), m, dim, dy.get(), dx.get(), n, d, nullptr, nullptr, ia,
                               dout.get(), kws, so.s);
      HIP_CHECK(hipMemcpyAsync(out, dout.get(), 4 * m * d, hipMemcpyDeviceToHost, so.s));
      HIP_CHECK(hipStreamSynchronize(so.s));
    }
  });
}

int32_t umapx_mlp_param_count(int32_t d, int32_t hidden, int32_t dim, int64_t* count) {
  return guarded([&] {
    if (!count) throw_invalid("count must not be null");
    if (d < 1 || hidden < 64 || hidden % 64 != 0 || dim < 1 || dim > 8)
      throw_invalid("need d >= 1, hidden a positive multiple of 64 and 1 <= dim <= 8");
    *count = MlpShape::make(d, hidden, dim).total;
  });
}

int32_t umapx_mlp_batch_grad(int32_t d, int32_t hidden, int32_t dim, const float* params,
                             const float* mean, const float* inv_std, const float* x_rows,
                             const float* y_rows, int32_t n_edges, int32_t n_neg, float a,
                             float b, float mse_weight, float* grads, float* loss, float* z_out) {
  return guarded([&] {
    if (!params || !mean || !inv_std || !x_rows || !y_rows || !grads || !loss)
      throw_invalid("null argument");
    mlp_batch_grad(d, hidden, dim, params, mean, inv_std, x_rows, y_rows, n_edges, n_neg, a, b,
                   mse_weight, grads, loss, z_out);
  });
}

int32_t umapx_mlp_forward(int32_t d, int32_t hidden, int32_t dim, const float* params,
                          const float* mean, const float* inv_std, const float* x, int64_t rows,
                          int32_t fused, float* z) {
  return guarded([&] {
    if (!params || !mean || !inv_std || !x || !z) throw_invalid("null argument");
    mlp_forward_host(d, hidden, dim, params, mean, inv_std, x, rows, fused != 0, z);
  });
}

int32_t umapx_get_mlp_loss(const umapx_handle* h, double* first, double* last, int32_t* steps) {
  return guarded([&] {
    require_handle(h);
    if (!h->mlp.trained) throw Error(kNotFitted, "the learned transform has not been trained");
    if (first) *first = h->mlp.first_loss;
    if (last) *last = h->mlp.last_loss;
    if (steps) *steps = h->mlp.steps_done;
  });
}

int32_t umapx_get_learned_kernel_ms(const umapx_handle* h, double* ms) {
  return guarded([&] {
    require_handle(h);
    if (!ms) throw_invalid("null argument");
    *ms = h->learned_kernel_ms;
  });
}

int32_t umapx_transform_learned_device(umapx_handle* h, const float* d_q, int64_t m, int64_t d,
                                       float* d_out) {
  return guarded([&] {
    require_handle(h);
    h->check_query(m, d);
    if (m == 0) return;
    h->ensure_stream();
    h->run_transform(d_q, m, d_out, UMAPX_TRANSFORM_LEARNED);
  });
}

int32_t umapx_n_samples(const umapx_handle* h, int64_t* n) {
  return guarded([&] {
    require_handle(h);
    if (!h->fitted) throw Error(kNotFitted, "model is not fitted yet");
    *n = h->n;
  });
}

int32_t umapx_n_features(const umapx_handle* h, int64_t* d) {
  return guarded([&] {
    require_handle(h);
    if (!h->fitted) throw Error(kNotFitted, "model is not fitted yet");
    *d = h->d;
  });
}

int32_t umapx_get_embedding(const umapx_handle* h, float* out) {
  return guarded([&] {
    require_handle(h);
    if (!h->fitted) throw Error(kNotFitted, "model is not fitted yet");
    HIP_CHECK(hipMemcpy(out, h->y.get(), 4 * h->n * h->p.n_components, hipMemcpyDeviceToHost));
  });
}

int32_t umapx_get_ab(const umapx_handle* h, float* a, float* b) {
  return guarded([&] {
    require_handle(h);
    if (!(h->a > 0.f && h->b > 0.f)) throw Error(kNotFitted, "model is not fitted yet");
    *a = h->a;
    *b = h->b;
  });
}

int32_t umapx_get_radii(const umapx_handle* h, float* orig, float* emb) {
  return guarded([&] {
    require_handle(h);
    if (!h->fitted) throw Error(kNotFitted, "model is not fitted yet");
    if (!h->rad_orig.get() || !h->rad_emb.get()) throw_invalid("radii were not computed (set densmap or output_dens)");
    if (orig) HIP_CHECK(hipMemcpy(orig, h->rad_orig.get(), 4 * h->n, hipMemcpyDeviceToHost));
    if (emb) HIP_CHECK(hipMemcpy(emb, h->rad_emb.get(), 4 * h->n, hipMemcpyDeviceToHost));
  });
}

int32_t umapx_find_ab_params(float spread, float min_dist, float* a, float* b) {
  return guarded([&] { find_ab_params(spread, min_dist, a, b); });
}

}  // extern "C"

namespace {

int32_t knn_host(const float* q, int64_t m, const float* r, int64_t n, int64_t d, int32_t k,
                 int32_t metric, float metric_p, int32_t precision, int32_t rerank_factor,
                 int64_t* idx, float* dist) {
  return guarded([&] {
    if (k < 1) throw_invalid("n_neighbors must be >= 1");
    validate_metric(metric, metric_p);
    if (m == 0) return;
    const bool self = (q == r) && (m == n);
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct SG {
      hipStream_t s;
      ~SG() { (void)hipStreamDestroy(s); }
    } sg{s};
    DevBuf<float> dr(static_cast<size_t>(n) * d);
    HIP_CHECK(hipMemcpyAsync(dr.get(), r, 4 * n * d, hipMemcpyHostToDevice, s));
    DevBuf<float> dq;
    const float* qp = dr.get();
    if (!self) {
      dq.alloc(static_cast<size_t>(m) * d);
      HIP_CHECK(hipMemcpyAsync(dq.get(), q, 4 * m * d, hipMemcpyHostToDevice, s));
      qp = dq.get();
    }
    DevBuf<int32_t> di(static_cast<size_t>(m) * k);
    DevBuf<float> dd(static_cast<size_t>(m) * k);
    KnnWorkspace ws;
    knn_device(qp, m, dr.get(), n, d, k, metric, precision, rerank_factor, di.get(), dd.get(), ws, s,
               nullptr, metric_p);
    std::vector<int32_t> hi(static_cast<size_t>(m) * k);
    HIP_CHECK(hipMemcpyAsync(hi.data(), di.get(), 4 * hi.size(), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(dist, dd.get(), 4 * hi.size(), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    idx64_from32(hi, idx);
  });
}

// The sparse kNN stages on host buffers: upload, run on a private stream,
// copy back [m x k] int64 indices and f32 distances.
template <typename Run>
int32_t csr_knn_host(int64_t m, int32_t k, int64_t* idx, float* dist, Run run) {
  return guarded([&] {
    if (k < 1) throw_invalid("n_neighbors must be >= 1");
    if (m == 0) return;
    if (!idx || !dist) throw_invalid("null argument");
    umapx_handle::StreamOwner so;
    HIP_CHECK(hipStreamCreateWithFlags(&so.s, hipStreamNonBlocking));
    std::vector<int32_t> hi(static_cast<size_t>(m) * k);
    {
      AllocStreamScope alloc_scope(so.s);
      DevBuf<int32_t> di(hi.size());
      DevBuf<float> dd(hi.size());
      run(di.get(), dd.get(), so.s);
      HIP_CHECK(hipMemcpyAsync(hi.data(), di.get(), 4 * hi.size(), hipMemcpyDeviceToHost, so.s));
      HIP_CHECK(hipMemcpyAsync(dist, dd.get(), 4 * hi.size(), hipMemcpyDeviceToHost, so.s));
      HIP_CHECK(hipStreamSynchronize(so.s));
    }
    idx64_from32(hi, idx);
  });
}
}  // namespace

extern "C" {

int32_t umapx_knn_to_reference(const float* q, int64_t m, const float* r, int64_t n, int64_t d,
                               int32_t k, int32_t metric, int32_t precision, int32_t rerank_factor,
                               int64_t* idx, float* dist) {
  return knn_host(q, m, r, n, d, k, metric, 2.0f, precision, rerank_factor, idx, dist);
}

int32_t umapx_knn_metric(const float* q, int64_t m, const float* r, int64_t n, int64_t d,
                         int32_t k, int32_t metric, float metric_p, int64_t* idx, float* dist) {
  return knn_host(q, m, r, n, d, k, metric, metric_p, UMAPX_KNN_FP32, 2, idx, dist);
}

int32_t umapx_knn_approx(const float* x, int64_t n, int64_t d, int32_t k, int32_t metric,
                         float metric_p, int32_t candidate_pool, int32_t n_iters, uint64_t seed,
                         int64_t* idx, float* dist) {
  return guarded([&] {
    if (!x || !idx || !dist) throw_invalid("null argument");
    if (n < 2 || d < 1) throw_invalid("need at least two samples and one feature");
    umapx_handle::StreamOwner so;
    HIP_CHECK(hipStreamCreateWithFlags(&so.s, hipStreamNonBlocking));
    std::vector<int32_t> hi(static_cast<size_t>(n) * k);
    {
      AllocStreamScope alloc_scope(so.s);
      DevBuf<float> dx(static_cast<size_t>(n) * d);
      DevBuf<int32_t> di(hi.size());
      DevBuf<float> dd(hi.size());
      HIP_CHECK(hipMemcpyAsync(dx.get(), x, 4 * n * d, hipMemcpyHostToDevice, so.s));
      KnnWorkspace ws;
      validate_metric(metric, metric_p);
      knn_approx_device(dx.get(), n, d, k, metric, metric_p, candidate_pool, n_iters, seed,
                        di.get(), dd.get(), ws, so.s);
      HIP_CHECK(hipMemcpyAsync(hi.data(), di.get(), 4 * hi.size(), hipMemcpyDeviceToHost, so.s));
      HIP_CHECK(hipMemcpyAsync(dist, dd.get(), 4 * hi.size(), hipMemcpyDeviceToHost, so.s));
      HIP_CHECK(hipStreamSynchronize(so.s));
    }
    idx64_from32(hi, idx);
  });
}


int32_t umapx_knn_csr(int64_t n, int64_t n_cols, const int64_t* indptr, const int64_t* indices,
                      const float* data, int64_t nnz, int32_t k, int32_t metric, float metric_p,
                      int64_t* idx, float* dist) {
  return csr_knn_host(n, k, idx, dist, [&](int32_t* di, float* dd, hipStream_t s) {
    validate_metric(metric, metric_p);
    csr_validate_host(n, n_cols, indptr, n + 1, indices, data, nnz);
    CsrDevice c;
    csr_upload(c, n, n_cols, indptr, indices, data, s);
    csr_knn_device(c, c, true, k, metric, metric_p, di, dd, s);
    HIP_CHECK(hipStreamSynchronize(s));
  });
}

int32_t umapx_knn_csr_query(int64_t m, const int64_t* q_indptr, const int64_t* q_indices,
                            const float* q_data, int64_t q_nnz, int64_t n,
                            const int64_t* r_indptr, const int64_t* r_indices,
                            const float* r_data, int64_t r_nnz, int64_t n_cols, int32_t k,
                            int32_t metric, float metric_p, int64_t* idx, float* dist) {
  return csr_knn_host(m, k, idx, dist, [&](int32_t* di, float* dd, hipStream_t s) {
    validate_metric(metric, metric_p);
    csr_validate_host(m, n_cols, q_indptr, m + 1, q_indices, q_data, q_nnz);
    csr_validate_host(n, n_cols, r_indptr, n + 1, r_indices, r_data, r_nnz);
    CsrDevice qc, rc;
    csr_upload(qc, m, n_cols, q_indptr, q_indices, q_data, s);
    csr_upload(rc, n, n_cols, r_indptr, r_indices, r_data, s);
    csr_knn_device(qc, rc, false, k, metric, metric_p, di, dd, s);
    HIP_CHECK(hipStreamSynchronize(s));
  });
}

int32_t umapx_knn_dense_to_csr(const float* q, int64_t m, int64_t n, const int64_t* r_indptr,
                               const int64_t* r_indices, const float* r_data, int64_t r_nnz,
                               int64_t n_cols, int32_t k, int32_t metric, float metric_p,
                               int64_t* idx, float* dist) {
  return csr_knn_host(m, k, idx, dist, [&](int32_t* di, float* dd, hipStream_t s) {
    validate_metric(metric, metric_p);
    if (!q) throw_invalid("null argument");
    csr_validate_host(n, n_cols, r_indptr, n + 1, r_indices, r_data, r_nnz);
    CsrDevice rc;
    csr_upload(rc, n, n_cols, r_indptr, r_indices, r_data, s);
    DevBuf<float> dq(static_cast<size_t>(m) * n_cols);
    HIP_CHECK(hipMemcpyAsync(dq.get(), q, 4 * m * n_cols, hipMemcpyHostToDevice, s));
    dense_to_csr_knn_device(dq.get(), m, rc, k, metric, metric_p, di, dd, s);
    HIP_CHECK(hipStreamSynchronize(s));
  });
}

int32_t umapx_knn(const float* x, int64_t n, int64_t d, int32_t k, int32_t metric, int32_t precision,
                  int32_t rerank_factor, int64_t* idx, float* dist) {
  return umapx_knn_to_reference(x, n, x, n, d, k, metric, precision, rerank_factor, idx, dist);
}

int32_t umapx_knn_rows_include_self(const int64_t* idx, const float* dist, int64_t n, int32_t k,
                                    int32_t* out) {
  return guarded([&] {
    if (n == 0 || k == 0) {
      *out = 0;
      return;
    }
    std::vector<int32_t> i32(idx, idx + n * k);
    DevBuf<int32_t> di(i32.size());
    DevBuf<float> dd(i32.size());
    HIP_CHECK(hipMemcpy(di.get(), i32.data(), 4 * i32.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dd.get(), dist, 4 * i32.size(), hipMemcpyHostToDevice));
    GraphWorkspace ws;
    *out = knn_include_self_device(di.get(), dd.get(), n, k, ws, nullptr) ? 1 : 0;
  });
}

int32_t umapx_smooth_knn_dist(const float* dist, int64_t n, int32_t k, float k_value,
                              float local_connectivity, float bandwidth, int32_t include_self,
                              int32_t fp64_sums, float* sigmas, float* rhos) {
  return guarded([&] {
    if (n == 0) return;
    DevBuf<float> dd(static_cast<size_t>(n) * k), ds(n), dr(n);
    HIP_CHECK(hipMemcpy(dd.get(), dist, 4 * n * k, hipMemcpyHostToDevice));
    GraphWorkspace ws;
    smooth_knn_dist_device(dd.get(), n, k, k_value, local_connectivity, bandwidth, include_self != 0,
                           fp64_sums != 0, ds.get(), dr.get(), ws, nullptr);
    HIP_CHECK(hipMemcpy(sigmas, ds.get(), 4 * n, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(rhos, dr.get(), 4 * n, hipMemcpyDeviceToHost));
  });
}

int32_t umapx_fuzzy_graph(const int64_t* idx, const float* dist, const float* sigmas,
                          const float* rhos, int64_t n, int32_t k, float set_op_mix_ratio,
                          int32_t prune_epochs, int64_t cap, int64_t* heads, int64_t* tails,
                          float* weights, int64_t* n_edges) {
  return guarded([&] {
    std::vector<int32_t> i32(idx, idx + n * k);
    for (auto v : i32)
      if (v < 0 || v >= n) throw_invalid("knn index out of range");
    DevBuf<int32_t> di(i32.size());
    DevBuf<float> dd(i32.size()), ds(n), dr(n);
    HIP_CHECK(hipMemcpy(di.get(), i32.data(), 4 * i32.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dd.get(), dist, 4 * i32.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(ds.get(), sigmas, 4 * n, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dr.get(), rhos, 4 * n, hipMemcpyHostToDevice));
    GraphWorkspace ws;
    Graph g;
    fuzzy_graph_device(di.get(), dd.get(), ds.get(), dr.get(), n, k, set_op_mix_ratio, prune_epochs, g,
                       ws, nullptr);
    *n_edges = g.n_edges;
    if (g.n_edges > cap) throw_invalid("edge capacity too small");
    std::vector<int32_t> hh(g.n_edges), tt(g.n_edges);
    HIP_CHECK(hipMemcpy(hh.data(), g.head.get(), 4 * g.n_edges, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(tt.data(), g.tail.get(), 4 * g.n_edges, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(weights, g.weight.get(), 4 * g.n_edges, hipMemcpyDeviceToHost));
    idx64_from32(hh, heads);
    idx64_from32(tt, tails);
  });
}

}  // extern "C"

namespace {
// Builds a device Graph from host head-sorted edges.
void graph_from_host(int64_t n, const int64_t* heads, const int64_t* tails, const float* w,
                     int64_t ne, Graph& g) {
  std::vector<int32_t> rp(n + 1, 0), hh(ne), tt(ne);
  float mx = 0.f;
  for (int64_t e = 0; e < ne; ++e) {
    if (heads[e] < 0 || heads[e] >= n || tails[e] < 0 || tails[e] >= n)
      throw_invalid("edge endpoint out of range");
    if (e > 0 && heads[e] < heads[e - 1]) throw_invalid("edges must be sorted by head");
    rp[heads[e] + 1] += 1;
    hh[e] = static_cast<int32_t>(heads[e]);
    tt[e] = static_cast<int32_t>(tails[e]);
    mx = std::max(mx, w[e]);
  }
  for (int64_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
  g.n = n;
  g.n_edges = ne;
  g.max_weight = mx;
  g.row_ptr.alloc(n + 1);
  g.head.alloc(std::max<int64_t>(ne, 1));
  g.tail.alloc(std::max<int64_t>(ne, 1));
  g.weight.alloc(std::max<int64_t>(ne, 1));
  HIP_CHECK(hipMemcpy(g.row_ptr.get(), rp.data(), 4 * (n + 1), hipMemcpyHostToDevice));
  if (ne > 0) {
    HIP_CHECK(hipMemcpy(g.head.get(), hh.data(), 4 * ne, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(g.tail.get(), tt.data(), 4 * ne, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(g.weight.get(), w, 4 * ne, hipMemcpyHostToDevice));
  }
}
}  // namespace

extern "C" {

int32_t umapx_initialize_embedding(const float* x, int64_t n, int64_t d, int32_t dim,
                                   const int64_t* heads, const int64_t* tails, const float* weights,
                                   int64_t n_edges, int32_t init, uint64_t seed, float* y,
                                   int32_t* used_spectral) {
  return guarded([&] {
    Graph g;
    graph_from_host(n, heads, tails, weights, n_edges, g);
    DevBuf<float> dx(static_cast<size_t>(n) * d), dy(static_cast<size_t>(n) * dim);
    HIP_CHECK(hipMemcpy(dx.get(), x, 4 * n * d, hipMemcpyHostToDevice));
    bool ok = false;
    if (init == UMAPX_INIT_SPECTRAL) {
      hipStream_t s;
      HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      SpectralWorkspace ws;
      try {
        ok = spectral_init_device(dx.get(), n, d, dim, g, seed, dy.get(), ws, s, nullptr);
      } catch (...) {
        (void)hipStreamDestroy(s);
        throw;
      }
      HIP_CHECK(hipStreamSynchronize(s));
      (void)hipStreamDestroy(s);
    }
    if (ok) {
      HIP_CHECK(hipMemcpy(y, dy.get(), 4 * n * dim, hipMemcpyDeviceToHost));
    } else {
      auto r = random_init(n, dim, seed, -10.0f, 10.0f);
      std::copy(r.begin(), r.end(), y);
    }
    if (used_spectral) *used_spectral = ok ? 1 : 0;
  });
}

int32_t umapx_connected_components(int64_t n, const int64_t* heads, const int64_t* tails,
                                   const float* weights, int64_t n_edges, int64_t* labels,
                                   int64_t* n_components) {
  return guarded([&] {
    Graph g;
    graph_from_host(n, heads, tails, weights, n_edges, g);
    SpectralWorkspace ws;
    std::vector<int32_t> lab;
    connected_components_device(g, lab, ws, nullptr);
    std::vector<int64_t> id(n, -1);
    int64_t nc = 0;
    for (int64_t v = 0; v < n; ++v) {
      if (id[lab[v]] < 0) id[lab[v]] = nc++;
      labels[v] = id[lab[v]];
    }
    *n_components = nc;
  });
}

int32_t umapx_optimize_layout(float* y, int64_t n, int32_t dim, const int64_t* heads,
                              const int64_t* tails, const float* weights, int64_t n_edges,
                              int32_t n_epochs, float a, float b, float learning_rate,
                              int32_t negative_sample_rate, float repulsion_strength, uint64_t seed,
                              const float* dens_radii, int32_t dens_start_epoch, float dens_lambda,
                              float dens_var_shift, int32_t* pos_counts, int32_t* neg_counts) {
  return guarded([&] {
    Graph g;
    graph_from_host(n, heads, tails, weights, n_edges, g);
    DevBuf<float> dy(static_cast<size_t>(n) * dim), dr;
    HIP_CHECK(hipMemcpy(dy.get(), y, 4 * n * dim, hipMemcpyHostToDevice));
    SgdArgs sa;
    sa.n_epochs = n_epochs;
    sa.a = a;
    sa.b = b;
    sa.learning_rate = learning_rate;
    sa.negative_sample_rate = negative_sample_rate;
    sa.repulsion_strength = repulsion_strength;
    sa.seed = seed;
    if (dens_radii) {
      dr.alloc(n);
      HIP_CHECK(hipMemcpy(dr.get(), dens_radii, 4 * n, hipMemcpyHostToDevice));
      sa.radii = dr.get();
      sa.dens_start_epoch = dens_start_epoch;
      sa.dens_lambda = dens_lambda;
      sa.dens_var_shift = dens_var_shift;
    }
    DevBuf<int32_t> pc, nc;
    if (pos_counts || neg_counts) {
      pc.alloc(std::max<int64_t>(n_edges, 1));
      nc.alloc(std::max<int64_t>(n_edges, 1));
      HIP_CHECK(hipMemset(pc.get(), 0, 4 * std::max<int64_t>(n_edges, 1)));
      HIP_CHECK(hipMemset(nc.get(), 0, 4 * std::max<int64_t>(n_edges, 1)));
      sa.pos_counts = pc.get();
      sa.neg_counts = nc.get();
    }
    SgdWorkspace ws;
    optimize_layout_device(dy.get(), n, dim, g, sa, ws, nullptr, nullptr);
    HIP_CHECK(hipDeviceSynchronize());
    HIP_CHECK(hipMemcpy(y, dy.get(), 4 * n * dim, hipMemcpyDeviceToHost));
    if (pos_counts) HIP_CHECK(hipMemcpy(pos_counts, pc.get(), 4 * n_edges, hipMemcpyDeviceToHost));
    if (neg_counts) HIP_CHECK(hipMemcpy(neg_counts, nc.get(), 4 * n_edges, hipMemcpyDeviceToHost));
  });
}

int32_t umapx_local_radii(const int64_t* idx, const float* dist, int64_t n, int32_t k, float var_shift,
                          float* out) {
  return guarded([&] {
    std::vector<int32_t> i32(idx, idx + n * k);
    DevBuf<int32_t> di(i32.size());
    DevBuf<float> dd(i32.size()), dout(n);
    HIP_CHECK(hipMemcpy(di.get(), i32.data(), 4 * i32.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dd.get(), dist, 4 * i32.size(), hipMemcpyHostToDevice));
    local_radii_device(di.get(), dd.get(), n, k, var_shift, dout.get(), nullptr);
    HIP_CHECK(hipMemcpy(out, dout.get(), 4 * n, hipMemcpyDeviceToHost));
  });
}

int32_t umapx_embedding_radii(const float* y, int64_t n, int32_t dim, const int64_t* idx, int32_t k,
                              float var_shift, float* out) {
  return guarded([&] {
    std::vector<int32_t> i32(idx, idx + n * k);
    DevBuf<int32_t> di(i32.size());
    DevBuf<float> dy(static_cast<size_t>(n) * dim), dout(n);
    HIP_CHECK(hipMemcpy(di.get(), i32.data(), 4 * i32.size(), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dy.get(), y, 4 * n * dim, hipMemcpyHostToDevice));
    embedding_radii_device(dy.get(), n, dim, di.get(), k, var_shift, dout.get(), nullptr);
    HIP_CHECK(hipMemcpy(out, dout.get(), 4 * n, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
