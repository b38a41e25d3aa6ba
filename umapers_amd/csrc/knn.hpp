// Exact kNN on the GPU (candidate MFMA tiles + exact f32 re-rank).
#pragma once

#include "common.hpp"

namespace umapx {

// Reusable device scratch for knn_device (grown on demand).
// A GEMM library handle owned by one workspace (so concurrent handles on
// different streams never share one); created on first use by the wide-row
// path, released with the workspace.
struct BlasHandle {
  void* h = nullptr;
  BlasHandle() = default;
  BlasHandle(const BlasHandle&) = delete;
  BlasHandle& operator=(const BlasHandle&) = delete;
  ~BlasHandle();
};

struct KnnWorkspace {
  BlasHandle blas;
  DevBuf<double> colsum;
  DevBuf<unsigned char> qprep, rprep;
  DevBuf<float> rnorm2, qnorm, rnorm;  // rnorm2: -||y||^2/2 of the compute rows
  DevBuf<int32_t> cand;
  DevBuf<uint2> pairs;  // per-split candidate lists when the reference rows are split
  DevBuf<float> msum;   // per-row sums / sums of squares (correlation)
  DevBuf<float> scores; // score block of the wide-row path, or the metric
                        // kernel's query tiles of rows wider than 512
};

// Events around the candidate (MFMA) kernel for live roofline figures.
struct KnnStats {
  hipEvent_t cand_start = nullptr, cand_stop = nullptr;
  double flops = 0.0;  // 2 * m * n * d, all ordered pairs
  bool recorded = false;
};

int knn_padded_dim(int64_t d);

// Elements of each KnnWorkspace buffer that knn_device fills for euclidean /
// cosine rows of at most 256 features (all zero for other shapes); self: q
// and r are the same rows.
struct KnnScratchNeed {
  size_t colsum = 0, rprep = 0, rnorm2 = 0, qprep = 0, cand = 0, pairs = 0, rnorm = 0, qnorm = 0;
};
KnnScratchNeed knn_scratch_need(int64_t m, int64_t n, int64_t d, int k, int metric,
                                int precision, int rerank_factor, bool self);

// Candidates for rows wider than 256 features: score blocks Q R^T by GEMM,
// then the kc largest q.r + nr_half[r] per query (kc 32 or 64).  qp / rp are
// the prepared [m x dpad] / [n x dpad] rows (BF16 when bf16, else FP32); r
// the original [n x d] reference rows (to find identical rows).
void knn_wide_candidates(const void* qp, int64_t m, const void* rp, const float* nr_half,
                         const float* r, int64_t n, int64_t d, int dpad, bool bf16, int kc,
                         int32_t* cand, KnnWorkspace& ws, hipStream_t s);
int knn_candidate_count(int k, int rerank_factor);

// Top-k of every query row q[m x d] against r[n x d] (device pointers).  When
// q == r and m == n this is the self-kNN of the fit (self included, as the
// reference does).  metric: 0 euclidean, 1 cosine.  precision: 0 FP32
// candidates, 1 BF16 candidates (both re-ranked exactly in f32).
// metric: 0 euclidean, 1 cosine (MFMA candidates + exact re-rank); 2
// manhattan, 3 chebyshev, 4 minkowski (metric_p), 5 correlation, 6 canberra,
// 7 braycurtis (exact distances on the vector ALUs, knn_metric_device).
void knn_device(const float* q, int64_t m, const float* r, int64_t n, int64_t d, int k,
                int metric, int precision, int rerank_factor, int32_t* out_idx,
                float* out_dist, KnnWorkspace& ws, hipStream_t s, KnnStats* stats,
                float metric_p = 2.0f);

// approximate_nearest_neighbors (lib.rs:2932-3168, 3213-3239): the
// reference's neighbours-of-neighbours rounds from random pools under any
// metric code, drawn from the reference's RNG stream; out_idx / out_dist
// [n x k].
void knn_approx_device(const float* x, int64_t n, int64_t d, int k, int metric, float metric_p,
                       int candidate_pool, int n_iters, uint64_t seed, int32_t* out_idx,
                       float* out_dist, KnnWorkspace& ws, hipStream_t s);

// Exact top-k under metric codes 2-7, distances in the reference's f32
// arithmetic (lib.rs:2494-2698), ties by index.
void knn_metric_device(const float* q, int64_t m, const float* r, int64_t n, int64_t d, int k,
                       int metric, float p, int32_t* out_idx, float* out_dist, KnnWorkspace& ws,
                       hipStream_t s);

// CSR rows on the device as the reference's SparseCsrMatrix holds them
// (rust_umap/src/sparse.rs:9-107): int64 row offsets, int32 column indices
// strictly increasing within a row, f32 values, and the per-row squared norm,
// L2 norm and sum in the reference's nonzero order.
struct CsrDevice {
  int64_t rows = 0, cols = 0, nnz = 0;
  DevBuf<int64_t> ptr;
  DevBuf<int32_t> ind;
  DevBuf<float> val, sq, l2, sum;
};
// SparseCsrMatrix::new's checks (sparse.rs:20-90) on host arrays; throws
// InvalidParameter with the reference's message.
void csr_validate_host(int64_t rows, int64_t cols, const int64_t* indptr, int64_t indptr_len,
                       const int64_t* indices, const float* data, int64_t nnz);
// Uploads validated host CSR arrays and forms the row summaries.
void csr_upload(CsrDevice& c, int64_t rows, int64_t cols, const int64_t* indptr,
                const int64_t* indices, const float* data, hipStream_t s);
// Exact top-k of every CSR query row against the CSR reference rows, each
// distance in the reference's sparse arithmetic (sparse.rs:220-603); self:
// q is the fit's own rows and row i is at distance 0.0 from itself
// (exact_nearest_neighbors, sparse.rs:605-626), else no self case
// (exact_nearest_neighbors_sparse_query, :628-657).  metric codes 0-5; k <= 64.
void csr_knn_device(const CsrDevice& q, const CsrDevice& r, bool self, int k, int metric,
                    float p, int32_t* out_idx, float* out_dist, hipStream_t s);
// Dense query rows q [m x r.cols] against CSR reference rows
// (exact_nearest_neighbors_dense_query, sparse.rs:659-900).
void dense_to_csr_knn_device(const float* q, int64_t m, const CsrDevice& r, int k, int metric,
                             float p, int32_t* out_idx, float* out_dist, hipStream_t s);
}  // namespace umapx
