/*
 * TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the umapers 1.1.4 UMAP fit / transform path
 * (rust_umap/src/lib.rs).  It is the correctness oracle for the HIP library
 * and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Every function cites the reference file:line it follows.  Parity status of
 * each part is listed in the header comment of umap_oracle.cpp.
 */
#ifndef UMAP_ORACLE_H
#define UMAP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors UmapParams (rust_umap/src/lib.rs:128-195) for the dense path. */
typedef struct orc_params {
  int32_t n_neighbors;
  int32_t n_components;
  int32_t n_epochs;            /* <= 0: default (500 if n <= 10000 else 200) */
  int32_t metric;              /* 0 euclidean, 1 cosine, 2 manhattan, 3 chebyshev,
                                  4 minkowski, 5 correlation, 6 canberra,
                                  7 braycurtis */
  float learning_rate;
  float min_dist;
  float spread;
  float local_connectivity;
  float set_op_mix_ratio;
  float repulsion_strength;
  int32_t negative_sample_rate;
  int32_t init;                /* 0 random, 1 spectral */
  uint64_t random_seed;
  int32_t densmap;
  float dens_lambda;
  float dens_frac;
  float dens_var_shift;
  int32_t output_dens;
  int32_t n_threads;           /* threads for the stages rayon runs in parallel */
  float metric_p;              /* minkowski p */
} orc_params;

/* Same keys as FitTimingBreakdown (rust_umap/src/lib.rs:51-68). */
typedef struct orc_timings {
  double validation_sec, knn_sec, curve_params_sec, knn_validate_trim_sec,
      smooth_knn_sec, density_original_sec, membership_sec, symmetrize_sec,
      target_intersection_sec, prune_sec, init_sec, optimize_sec,
      density_embedding_sec, output_copy_sec, store_state_sec, total_sec;
} orc_timings;

void orc_find_ab_params(float spread, float min_dist, float* a, float* b);

void orc_knn(const float* x, int64_t n, int64_t d, int32_t k, int32_t metric,
             int32_t n_threads, int64_t* idx, float* dist);
/* Exact kNN of q [m x d] among r [n x d] under any metric code (self
 * included when q == r); metric_p is minkowski's p. */
void orc_knn_metric(const float* q, int64_t m, const float* r, int64_t n, int64_t d,
                    int32_t k, int32_t metric, float metric_p, int32_t n_threads,
                    int64_t* idx, float* dist);
/* approximate_nearest_neighbors (lib.rs:2932-3168, 3213-3281) under any
 * metric code: the k entries of each sample's table. */
void orc_knn_approx(const float* x, int64_t n, int64_t d, int32_t k, int32_t metric,
                    float metric_p, int32_t candidate_pool, int32_t n_iters, uint64_t seed,
                    int32_t n_threads, int64_t* idx, float* dist);
void orc_knn_to_reference(const float* q, int64_t m, const float* r, int64_t n,
                          int64_t d, int32_t k, int32_t metric,
                          int32_t n_threads, int64_t* idx, float* dist);
/* Single-pair distances with the reference's arithmetic. */
float orc_sq_euclidean(const float* x, const float* y, int64_t d);
float orc_cosine_distance(const float* x, const float* y, int64_t d);

int32_t orc_knn_rows_include_self(const int64_t* idx, const float* dist,
                                  int64_t n, int32_t k);
void orc_smooth_knn_dist(const float* dist, int64_t n, int32_t k, float k_value,
                         float local_connectivity, float bandwidth,
                         int32_t include_self, float* sigmas, float* rhos);
/* normalized_laplacian_from_edges (lib.rs:3966-4023), or with from_affinity
 * the dense max-symmetrised affinity route (lib.rs:3928-3964): n x n, row-major.
 * Returns 0 when the graph has no positive weight. */
int32_t orc_normalized_laplacian(int64_t n, const int64_t* heads, const int64_t* tails,
                                 const float* w, int64_t ne, int32_t from_affinity,
                                 double* out);
/* The same with the bisection's psum in f64 (the GPU path's default). */
void orc_smooth_knn_dist_fp64(const float* dist, int64_t n, int32_t k, float k_value,
                              float local_connectivity, float bandwidth,
                              int32_t include_self, float* sigmas, float* rhos);
/* Membership + fuzzy union (+ prune when n_epochs > 0) into head-sorted
 * edges.  Returns the edge count, or -1 when cap is too small. */
int64_t orc_fuzzy_graph(const int64_t* idx, const float* dist,
                        const float* sigmas, const float* rhos, int64_t n,
                        int32_t k, float set_op_mix_ratio, int32_t n_epochs,
                        int64_t cap, int64_t* heads, int64_t* tails, float* w);
int64_t orc_connected_components(int64_t n, const int64_t* heads,
                                 const int64_t* tails, int64_t n_edges,
                                 int64_t* labels);
/* Returns 1 on spectral success, 0 when it fell back to random init. */
int32_t orc_initialize_embedding(const float* x, int64_t n, int64_t d,
                                 int32_t dim, const int64_t* heads,
                                 const int64_t* tails, const float* w,
                                 int64_t n_edges, int32_t init, uint64_t seed,
                                 float* y);
void orc_make_epochs_per_sample(const float* w, int64_t n_edges,
                                int32_t n_epochs, float* eps);
/* Exact per-edge schedule: positive and negative sample counts over all
 * epochs, plus per-epoch totals of active edges and negatives. */
void orc_schedule_counts(const float* w, int64_t n_edges, int32_t n_epochs,
                         int32_t negative_sample_rate, int32_t* pos_counts,
                         int32_t* neg_counts, int64_t* active_per_epoch,
                         int64_t* neg_per_epoch);
void orc_optimize_layout(float* y, int64_t n, int32_t dim,
                         const int64_t* heads, const int64_t* tails,
                         const float* w, int64_t n_edges, int32_t n_epochs,
                         float a, float b, float learning_rate,
                         int32_t negative_sample_rate, float repulsion_strength,
                         uint64_t seed, const float* dens_radii,
                         int32_t dens_start_epoch, float dens_lambda,
                         float dens_var_shift);
int32_t orc_density_start_epoch(int32_t n_epochs, float dens_frac);
void orc_local_radii(const int64_t* idx, const float* dist, int64_t n, int32_t k,
                     float var_shift, float* out);
void orc_embedding_radii(const float* y, int64_t n, int32_t dim,
                         const int64_t* idx, int32_t k, float var_shift,
                         float* out);

/* Whole dense fit (lib.rs:1367-1464 via profile_fit_transform_dense :796-962).
 * knn_idx/knn_dist may be NULL (exact kNN is then computed); otherwise they
 * hold [n x knn_cols] precomputed neighbours.  Returns 0 on success. */
int32_t orc_fit(const orc_params* p, const float* x, int64_t n, int64_t d,
                const int64_t* knn_idx, const float* knn_dist, int64_t knn_cols,
                float* y_out, float* sigmas, float* rhos, float* rad_orig,
                float* rad_emb, orc_timings* t);
/* Plain transform (lib.rs:1604-1685, 1738-1780). */
int32_t orc_transform(const orc_params* p, const float* x_train, int64_t n,
                      int64_t d, const float* y_train, float a, float b,
                      const float* q, int64_t m, float* out);
/* transform_from_knn (lib.rs:1738-1780) on a query kNN [m x k] among the n
 * fitted rows, e.g. from the sparse kNN below. */
int32_t orc_transform_with_knn(const orc_params* p, int64_t n, const float* y_train, float a,
                               float b, int64_t m, const int64_t* idx, const float* dist,
                               int32_t k, float* out);
/* Sparse (CSR) exact kNN, sparse_oracle.cpp (rust_umap/src/sparse.rs):
 * indptr [rows + 1], strictly increasing column indices per row, f32 data.
 * orc_knn_csr: every row against every row, the row itself at 0.0 (:605-626);
 * orc_knn_csr_query: CSR queries against CSR rows (:628-657);
 * orc_knn_dense_to_csr: dense queries [m x n_cols] against CSR rows
 * (:753-900).  Metric codes as above (canberra / braycurtis are dense-only). */
void orc_knn_csr(int64_t n, int64_t n_cols, const int64_t* indptr, const int64_t* indices,
                 const float* data, int32_t k, int32_t metric, float p, int32_t n_threads,
                 int64_t* out_idx, float* out_dist);
void orc_knn_csr_query(int64_t m, const int64_t* q_indptr, const int64_t* q_indices,
                       const float* q_data, int64_t n, const int64_t* r_indptr,
                       const int64_t* r_indices, const float* r_data, int64_t n_cols, int32_t k,
                       int32_t metric, float p, int32_t n_threads, int64_t* out_idx,
                       float* out_dist);
void orc_knn_dense_to_csr(const float* q, int64_t m, int64_t n, const int64_t* r_indptr,
                          const int64_t* r_indices, const float* r_data, int64_t n_cols,
                          int32_t k, int32_t metric, float p, int32_t n_threads,
                          int64_t* out_idx, float* out_dist);
/* SparseCsrMatrix::new's checks (sparse.rs:20-90): 0 if valid, else 1 with
 * the reference's message in msg. */
int32_t orc_csr_check(int64_t n_rows, int64_t n_cols, const int64_t* indptr, int64_t indptr_len,
                      const int64_t* indices, int64_t n_indices, const float* data,
                      int64_t n_data, char* msg, int64_t cap);
/* Inverse transform (lib.rs:1800-1915): embedded rows [m x n_components]
 * back to the feature space [m x d]; optimize = 0 stops after the local
 * linear initialisation (init_inverse_points, lib.rs:4838-4935). */
int32_t orc_inverse_transform(const orc_params* p, const float* x_train, int64_t n,
                              int64_t d, const float* y_train, const float* sigmas,
                              const float* rhos, float a, float b, const float* q,
                              int64_t m, int32_t optimize, float* out);

#ifdef __cplusplus
}
#endif
#endif
