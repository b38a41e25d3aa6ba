/*
 * umapx — C ABI of the MI355X-native UMAP fit / transform library.
 *
 * The reference (umapers 1.1.4) crosses its language boundary through PyO3
 * (umap_rs/src/lib.rs) into the Rust core (rust_umap/src/lib.rs).  This header
 * is the equivalent boundary as plain C: caller-owned pointers and sizes, no
 * framework types.  Entry points ending in _device take device pointers that
 * are already resident in GPU memory; all others take host pointers and copy.
 *
 * Every function returns UMAPX_OK (0) or an error code; umapx_last_error()
 * gives the message of the last failure on the calling thread.  The codes
 * mirror UmapError (rust_umap/src/lib.rs:198-244) plus UMAPX_ERR_DEVICE for
 * HIP runtime failures; the binding maps UMAPX_ERR_NOT_FITTED and
 * UMAPX_ERR_DEVICE to RuntimeError and the rest to ValueError, as
 * umap_rs/src/lib.rs:147-157 does.
 */
#ifndef UMAPX_H
#define UMAPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UMAPX_OK 0
#define UMAPX_ERR_EMPTY_DATA 1
#define UMAPX_ERR_NEED_TWO_SAMPLES 2
#define UMAPX_ERR_INCONSISTENT_DIMENSIONS 3
#define UMAPX_ERR_FEATURE_MISMATCH 4
#define UMAPX_ERR_EMBEDDING_DIMENSION_MISMATCH 5
#define UMAPX_ERR_NOT_FITTED 6
#define UMAPX_ERR_INVALID_PARAMETER 7
#define UMAPX_ERR_DEVICE 8

#define UMAPX_METRIC_EUCLIDEAN 0
#define UMAPX_METRIC_COSINE 1
#define UMAPX_METRIC_MANHATTAN 2
#define UMAPX_METRIC_CHEBYSHEV 3
#define UMAPX_METRIC_MINKOWSKI 4  /* p in umapx_params.metric_p */
#define UMAPX_METRIC_CORRELATION 5
#define UMAPX_METRIC_CANBERRA 6
#define UMAPX_METRIC_BRAYCURTIS 7

#define UMAPX_INIT_RANDOM 0
#define UMAPX_INIT_SPECTRAL 1

#define UMAPX_KNN_FP32 0
#define UMAPX_KNN_BF16_RERANK 1

#define UMAPX_TRANSFORM_PLAIN 0
#define UMAPX_TRANSFORM_LEARNED 1

/* UmapParams (rust_umap/src/lib.rs:128-195) plus the GPU path's own knobs. */
typedef struct umapx_params {
  int32_t n_neighbors;          /* 15 */
  int32_t n_components;         /* 2 */
  int32_t n_epochs;             /* <= 0: 500 if n <= 10000 else 200 */
  int32_t metric;               /* UMAPX_METRIC_* */
  float learning_rate;          /* 1.0 */
  float min_dist;               /* 0.1 */
  float spread;                 /* 1.0 */
  float local_connectivity;     /* 1.0 */
  float set_op_mix_ratio;       /* 1.0 */
  float repulsion_strength;     /* 1.0 */
  int32_t negative_sample_rate; /* 5 */
  int32_t init;                 /* UMAPX_INIT_SPECTRAL */
  uint64_t random_seed;         /* 42 */
  int32_t use_approximate_knn;  /* 1 (default): the approximate rounds past 4 x
                                   approx_knn_threshold rows, as the reference's auto mode
                                   (threshold 0: always); 0: the exact kNN */
  int32_t approx_knn_candidates;
  int32_t approx_knn_iters;
  int32_t approx_knn_threshold;
  int32_t densmap;
  float dens_lambda;            /* 2.0 */
  float dens_frac;              /* 0.3 */
  float dens_var_shift;         /* 0.1 */
  int32_t output_dens;
  /* GPU path */
  int32_t knn_precision;        /* UMAPX_KNN_FP32 */
  int32_t rerank_factor;        /* candidates per row = rerank_factor * k (2) */
  int32_t smooth_fp64;          /* 1: FP64 bisection sums (default), 0: f32 */
  int32_t transform_mode;       /* UMAPX_TRANSFORM_PLAIN */
  /* learned transform (MLP encoder d -> hidden -> hidden -> n_components) */
  int32_t mlp_hidden;           /* 256 */
  int32_t mlp_steps;            /* optimizer steps (1500) */
  int32_t mlp_batch;            /* edges per step (4096) */
  int32_t mlp_negatives;        /* negatives per edge (5) */
  float mlp_learning_rate;      /* Adam step size (1e-3) */
  float mlp_embedding_weight;   /* weight of the MSE to the fitted embedding (1.0) */
  float metric_p;               /* p of UMAPX_METRIC_MINKOWSKI (2.0) */
} umapx_params;

/* FitTimingBreakdown (rust_umap/src/lib.rs:51-68) plus GPU kernel figures. */
typedef struct umapx_timings {
  double validation_sec, knn_sec, curve_params_sec, knn_validate_trim_sec,
      smooth_knn_sec, density_original_sec, membership_sec, symmetrize_sec,
      target_intersection_sec, prune_sec, init_sec, optimize_sec,
      density_embedding_sec, output_copy_sec, store_state_sec, total_sec;
  double knn_kernel_ms;   /* candidate (MFMA) kernel, HIP events */
  double knn_flops;       /* 2 * n * n * d of that kernel */
  double sgd_kernel_ms;   /* all SGD epoch launches, HIP events */
  int64_t n_edges;        /* edges after prune */
  int32_t n_epochs;
  int32_t used_approximate_knn; /* 1 when the approximate rounds built the kNN */
  int32_t init_used_spectral;
  int32_t n_graph_components;
  /* umapx_profile_fit_transform only (0 otherwise): positive updates and
   * negative samples over all epochs, from the exact firing schedule */
  int64_t sgd_positive_updates;
  int64_t sgd_negative_samples;
} umapx_timings;

typedef struct umapx_handle umapx_handle;

void umapx_default_params(umapx_params* out);
const char* umapx_last_error(void);
/* The structured fields of the calling thread's last failure, so that
 * UmapError (rust_umap/src/lib.rs:198-216) can be rebuilt exactly: row /
 * expected / got of InconsistentDimensions, expected / got of
 * FeatureMismatch and EmbeddingDimensionMismatch, -1 where the variant has
 * no such field.  Returns that failure's code (UMAPX_OK before any). */
int32_t umapx_last_error_detail(int64_t* row, int64_t* expected, int64_t* got);
/* should_use_approximate_knn (rust_umap/src/lib.rs:2285-2311): 1 when a fit
 * of n_samples x n_features rows under these parameters builds its kNN with
 * the approximate rounds (auto mode past 4 x approx_knn_threshold rows;
 * threshold 0 forces them), 0 when it runs the exact kNN. */
int32_t umapx_uses_approximate_knn(const umapx_params* params, int64_t n_samples,
                                   int64_t n_features);
const char* umapx_version(void);
/* Number of visible GPUs (0 when none); never fails. */
int32_t umapx_device_count(void);

/* UmapModel::new (rust_umap/src/lib.rs:584-598); validates nothing yet. */
int32_t umapx_create(const umapx_params* params, umapx_handle** out);
void umapx_destroy(umapx_handle* h);

/* UmapModel::fit_dense / fit_transform_dense (lib.rs:712-794). */
int32_t umapx_fit(umapx_handle* h, const float* x, int64_t n, int64_t d);
int32_t umapx_fit_transform(umapx_handle* h, const float* x, int64_t n, int64_t d,
                            float* y_out);
/* Same with x and y_out in device memory (x stays resident, nothing is
 * copied over PCIe). */
int32_t umapx_fit_transform_device(umapx_handle* h, const float* d_x, int64_t n, int64_t d,
                                   float* d_y_out);
/* fit_transform_with_knn_metric_dense_i64_flat (lib.rs:1159-1212). */
int32_t umapx_fit_transform_with_knn(umapx_handle* h, const float* x, int64_t n, int64_t d,
                                     const int64_t* knn_idx, int64_t idx_rows, int64_t idx_cols,
                                     const float* knn_dist, int64_t dist_rows, int64_t dist_cols,
                                     int32_t knn_metric, int32_t validate_precomputed,
                                     float* y_out);
/* profile_fit_transform_dense (lib.rs:796-962). */
int32_t umapx_profile_fit_transform(umapx_handle* h, const float* x, int64_t n, int64_t d,
                                    float* y_out, umapx_timings* t);
/* fit_transform_sparse_csr (lib.rs:1015-1036; fit_sparse_csr :998-1013 when
 * y_out is NULL) on CSR rows: indptr [n_rows + 1], column indices strictly
 * increasing within a row, f32 data, nnz = indptr[n_rows].  The CSR checks
 * of SparseCsrMatrix::new (rust_umap/src/sparse.rs:20-90) come first, then
 * validate_sparse_metric / validate_sparse_density_params (lib.rs:2241-2258):
 * canberra / braycurtis, densMAP and output_dens are dense-only.  The kNN is
 * the reference's sparse exact kNN (sparse.rs:605-626, its arithmetic bit
 * for bit except minkowski); n_neighbors <= 64. */
int32_t umapx_fit_transform_csr(umapx_handle* h, int64_t n_rows, int64_t n_cols,
                                const int64_t* indptr, const int64_t* indices, const float* data,
                                int64_t nnz, float* y_out);
/* transform_sparse_csr (lib.rs:1687-1737): CSR queries on a sparse fit.
 * umapx_transform on a sparse fit takes dense queries through the
 * reference's dense-query sparse kNN (sparse.rs:753-900, lib.rs:1655-1672). */
int32_t umapx_transform_csr(umapx_handle* h, int64_t m, int64_t n_cols, const int64_t* indptr,
                            const int64_t* indices, const float* data, int64_t nnz, float* out);
/* 1 when the handle was fitted on CSR rows (no counterpart in the reference,
 * whose model keeps training_data_sparse, lib.rs:1571-1590). */
int32_t umapx_is_sparse_fit(const umapx_handle* h, int32_t* sparse);

/* Timings of the last fit on this handle. */
int32_t umapx_get_timings(const umapx_handle* h, umapx_timings* t);

/* transform_dense (lib.rs:1593-1685): plain kNN + SGD transform, or the
 * learned encoder when transform_mode is UMAPX_TRANSFORM_LEARNED. */
int32_t umapx_transform(umapx_handle* h, const float* q, int64_t m, int64_t d, float* out);
int32_t umapx_transform_device(umapx_handle* h, const float* d_q, int64_t m, int64_t d,
                               float* d_out);
/* Force one transform form regardless of transform_mode. */
int32_t umapx_transform_plain(umapx_handle* h, const float* q, int64_t m, int64_t d,
                              float* out);
int32_t umapx_transform_learned(umapx_handle* h, const float* q, int64_t m, int64_t d,
                                float* out);
int32_t umapx_transform_learned_device(umapx_handle* h, const float* d_q, int64_t m, int64_t d,
                                       float* d_out);

/* inverse_transform_dense (lib.rs:1789-1915): embedded rows [m x n_components]
 * back to the training feature space [m x n_features]: kNN among the fitted
 * embedding, the local linear initialisation (lib.rs:4838-4935) and the
 * feature-space SGD (lib.rs:5547-5645). */
int32_t umapx_inverse_transform(umapx_handle* h, const float* q, int64_t m, int64_t dim,
                                float* out);
int32_t umapx_inverse_transform_device(umapx_handle* h, const float* d_q, int64_t m,
                                       int64_t dim, float* d_out);

/* Fitted state (UmapModel accessors, lib.rs:600-634). */
int32_t umapx_n_samples(const umapx_handle* h, int64_t* n);
int32_t umapx_n_features(const umapx_handle* h, int64_t* d);
int32_t umapx_get_embedding(const umapx_handle* h, float* out);
int32_t umapx_get_ab(const umapx_handle* h, float* a, float* b);
/* radii_original / radii_embedding (only when densmap or output_dens). */
int32_t umapx_get_radii(const umapx_handle* h, float* orig, float* emb);

/* ------------------------------------------------------------------------
 * Stage-level entry points (one per reference stage; host pointers).
 * ------------------------------------------------------------------------ */

/* find_ab_params (lib.rs:5669-5746). */
int32_t umapx_find_ab_params(float spread, float min_dist, float* a, float* b);

/* exact_nearest_neighbors (lib.rs:2911-2928): self-kNN, self included. */
int32_t umapx_knn(const float* x, int64_t n, int64_t d, int32_t k, int32_t metric,
                  int32_t precision, int32_t rerank_factor, int64_t* idx, float* dist);
/* exact_nearest_neighbors_to_reference (lib.rs:3346-3399). */
int32_t umapx_knn_to_reference(const float* q, int64_t m, const float* r, int64_t n,
                               int64_t d, int32_t k, int32_t metric, int32_t precision,
                               int32_t rerank_factor, int64_t* idx, float* dist);
/* No counterpart in the reference.  umapx_knn_to_reference (and, with
 * d_q == d_r and m == n, umapx_knn) for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Inputs d_q [m x d] / d_r [n x d] and outputs d_idx /
 * d_dist [m x k] are caller-owned device buffers; every temporary comes from
 * one static device array of UMAPX_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends.  Euclidean or cosine, 1 to 256 features; a shape whose
 * temporaries exceed the array is rejected with UMAPX_ERR_INVALID_PARAMETER.
 * Work is queued on `stream` (a hipStream_t; NULL is the null stream) and
 * not waited for.  All calls on one GPU share the array: calls that overlap
 * on one GPU race. */
#define UMAPX_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)
int32_t umapx_knn_to_reference_device_noalloc(const float* d_q, int64_t m, const float* d_r,
                                              int64_t n, int64_t d, int32_t k, int32_t metric,
                                              int32_t precision, int32_t rerank_factor,
                                              int64_t* d_idx, float* d_dist, void* stream);
/* approximate_nearest_neighbors (lib.rs:2932-3168, 3213-3281): the
 * reference's neighbours-of-neighbours rounds from random pools under any
 * UMAPX_METRIC_* code, on the reference's RNG stream (seed = random_seed ^
 * 0xC0FEFEED1234ABCD in a fit, lib.rs:2326). */
int32_t umapx_knn_approx(const float* x, int64_t n, int64_t d, int32_t k, int32_t metric,
                         float metric_p, int32_t candidate_pool, int32_t n_iters, uint64_t seed,
                         int64_t* idx, float* dist);
/* Sparse exact kNN stages (rust_umap/src/sparse.rs) on host CSR buffers;
 * metric codes 0-5 (canberra / braycurtis are dense-only), k <= 64:
 * umapx_knn_csr: exact_nearest_neighbors (:605-626), the fit's kNN, every
 *   row at distance 0.0 from itself;
 * umapx_knn_csr_query: exact_nearest_neighbors_sparse_query (:628-657);
 * umapx_knn_dense_to_csr: exact_nearest_neighbors_dense_query (:753-900),
 *   dense q [m x n_cols] against CSR rows. */
int32_t umapx_knn_csr(int64_t n, int64_t n_cols, const int64_t* indptr, const int64_t* indices,
                      const float* data, int64_t nnz, int32_t k, int32_t metric, float metric_p,
                      int64_t* idx, float* dist);
int32_t umapx_knn_csr_query(int64_t m, const int64_t* q_indptr, const int64_t* q_indices,
                            const float* q_data, int64_t q_nnz, int64_t n,
                            const int64_t* r_indptr, const int64_t* r_indices,
                            const float* r_data, int64_t r_nnz, int64_t n_cols, int32_t k,
                            int32_t metric, float metric_p, int64_t* idx, float* dist);
int32_t umapx_knn_dense_to_csr(const float* q, int64_t m, int64_t n, const int64_t* r_indptr,
                               const int64_t* r_indices, const float* r_data, int64_t r_nnz,
                               int64_t n_cols, int32_t k, int32_t metric, float metric_p,
                               int64_t* idx, float* dist);
/* exact_nearest_neighbors_by / _to_reference_by (lib.rs:2891-2928, 3346-3399)
 * under any UMAPX_METRIC_* code with minkowski's p (self included when q == r
 * and m == n). */
int32_t umapx_knn_metric(const float* q, int64_t m, const float* r, int64_t n, int64_t d,
                         int32_t k, int32_t metric, float metric_p, int64_t* idx, float* dist);

/* knn_rows_include_self (lib.rs:3402-3417): writes 0/1. */
int32_t umapx_knn_rows_include_self(const int64_t* idx, const float* dist, int64_t n, int32_t k,
                                    int32_t* out);
/* smooth_knn_dist (lib.rs:3424-3523). */
int32_t umapx_smooth_knn_dist(const float* dist, int64_t n, int32_t k, float k_value,
                              float local_connectivity, float bandwidth, int32_t include_self,
                              int32_t fp64_sums, float* sigmas, float* rhos);
/* compute_membership_strengths + symmetrize_fuzzy_graph (+ prune_edges when
 * prune_epochs > 0) (lib.rs:3525-3649, 3804-3822).  Writes at most cap
 * head-sorted edges and their count. */
int32_t umapx_fuzzy_graph(const int64_t* idx, const float* dist, const float* sigmas,
                          const float* rhos, int64_t n, int32_t k, float set_op_mix_ratio,
                          int32_t prune_epochs, int64_t cap, int64_t* heads, int64_t* tails,
                          float* weights, int64_t* n_edges);
/* initialize_embedding (lib.rs:3841-3853) over head-sorted edges; writes
 * 1 to used_spectral when the spectral layout succeeded. */
int32_t umapx_initialize_embedding(const float* x, int64_t n, int64_t d, int32_t dim,
                                   const int64_t* heads, const int64_t* tails,
                                   const float* weights, int64_t n_edges, int32_t init,
                                   uint64_t seed, float* y, int32_t* used_spectral);
/* connected_components_from_edges (lib.rs:4394-4437): component id per node,
 * components numbered by their smallest member. */
int32_t umapx_connected_components(int64_t n, const int64_t* heads, const int64_t* tails,
                                   const float* weights, int64_t n_edges, int64_t* labels,
                                   int64_t* n_components);
/* optimize_layout_training (lib.rs:5322-5441) on head-sorted edges; y is
 * updated in place.  dens_radii may be NULL (no densMAP term).  pos_counts /
 * neg_counts, when not NULL, receive each edge's sample and negative counts. */
int32_t umapx_optimize_layout(float* y, int64_t n, int32_t dim, const int64_t* heads,
                              const int64_t* tails, const float* weights, int64_t n_edges,
                              int32_t n_epochs, float a, float b, float learning_rate,
                              int32_t negative_sample_rate, float repulsion_strength,
                              uint64_t seed, const float* dens_radii, int32_t dens_start_epoch,
                              float dens_lambda, float dens_var_shift, int32_t* pos_counts,
                              int32_t* neg_counts);
/* ------------------------------------------------------------------------
 * Learned transform (the reference's nearest analog is ParametricUmapModel,
 * rust_umap/src/parametric.rs:89-534; the encoder and its loss are this
 * library's own, see csrc/mlp.hip).
 * ------------------------------------------------------------------------ */

/* Number of f32 parameters of the encoder d -> hidden -> hidden -> dim; the
 * flat layout is W1 [dp x hidden], b1, W2 [hidden x hidden], b2,
 * W3 [hidden x dim], b3 with dp = d rounded up to a multiple of 32. */
int32_t umapx_mlp_param_count(int32_t d, int32_t hidden, int32_t dim, int64_t* count);
/* One forward / backward pass of the encoder's training loss on a given
 * batch: rows are n_edges heads, n_edges tails, then n_edges * n_neg
 * negatives (negative t of edge e at row 2 n_edges + e n_neg + t); x_rows are
 * raw features (standardised with mean / inv_std), y_rows the fitted
 * embedding.  Writes the parameter gradients, the loss and the outputs z. */
int32_t umapx_mlp_batch_grad(int32_t d, int32_t hidden, int32_t dim, const float* params,
                             const float* mean, const float* inv_std, const float* x_rows,
                             const float* y_rows, int32_t n_edges, int32_t n_neg, float a,
                             float b, float mse_weight, float* grads, float* loss, float* z_out);
/* The encoder's forward pass of given parameters over host rows x
 * [rows x d] into z [rows x dim]: fused = 0 runs the unfused layer chain
 * that training uses, fused = 1 the single inference kernel that
 * umapx_transform_learned runs (mlp_hidden 64, 128, 192 or 256, d % 4 == 0,
 * d <= 1024).  For checking the two against each other and a host forward. */
int32_t umapx_mlp_forward(int32_t d, int32_t hidden, int32_t dim, const float* params,
                          const float* mean, const float* inv_std, const float* x, int64_t rows,
                          int32_t fused, float* z);
/* HIP-event time (ms, on the handle's stream) of the encoder kernels of the
 * last learned transform; training excluded. */
int32_t umapx_get_learned_kernel_ms(const umapx_handle* h, double* ms);
/* Training loss of the learned transform at its first and last step. */
int32_t umapx_get_mlp_loss(const umapx_handle* h, double* first, double* last, int32_t* steps);

/* local_radii_from_knn / embedding_radii_from_knn (lib.rs:5057-5108). */
int32_t umapx_local_radii(const int64_t* idx, const float* dist, int64_t n, int32_t k,
                          float var_shift, float* out);
int32_t umapx_embedding_radii(const float* y, int64_t n, int32_t dim, const int64_t* idx,
                              int32_t k, float var_shift, float* out);

/* The first two steps of inverse_transform_rows (lib.rs:1856-1890) and
 * init_inverse_points (lib.rs:4838-4935) without the SGD: the exact kNN of q
 * [m x dim] among y_train [n x dim] (k = min(k, n)), the weights
 * 1 / (1 + a d^2b) and the local linear estimate in the space of x_train
 * [n x d], written to out [m x d]. */
int32_t umapx_inverse_init(const float* x_train, int64_t n, int64_t d, const float* y_train,
                           int32_t dim, int32_t k, float a, float b, const float* q, int64_t m,
                           float* out);

#ifdef __cplusplus
}
#endif
#endif /* UMAPX_H */
