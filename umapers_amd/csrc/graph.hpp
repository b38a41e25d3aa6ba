// Fuzzy simplicial set on the GPU: smooth_knn_dist, membership strengths,
// fuzzy union into head-sorted CSR, prune, and device scans.
#pragma once

#include "common.hpp"

namespace umapx {

struct ScanWorkspace {
  DevBuf<int32_t> partial;  // per-block sums, recursively
  DevBuf<int32_t> partial2;
  DevBuf<int32_t> total;
};

// out[i] = sum(in[0..i)), returns the grand total (synchronises the stream).
int64_t exclusive_scan(const int32_t* in, int32_t* out, int64_t n, ScanWorkspace& ws,
                       hipStream_t s);

// Device CSR graph with per-edge head ids (head-sorted, then tail-sorted).
struct Graph {
  int64_t n = 0;
  int64_t n_edges = 0;
  DevBuf<int32_t> row_ptr;  // n + 1
  DevBuf<int32_t> head;     // E
  DevBuf<int32_t> tail;     // E
  DevBuf<float> weight;     // E
  float max_weight = 0.f;
  bool symmetric = false;  // every edge (h, t) has its reverse (t, h): the fit's union
};

struct GraphWorkspace {
  DevBuf<float> wdir;       // n*k directed weights
  DevBuf<int32_t> indeg, in_off, fill, in_src;
  DevBuf<float> in_w;
  DevBuf<int32_t> tmp_col, ucount, kcount, tmp_off, serial_c;
  DevBuf<float> serial_w;
  DevBuf<float> tmp_w;
  DevBuf<int32_t> flag;
  DevBuf<int32_t> long_rows;  // rows with more than 64 union entries, and their counts
  DevBuf<double> dsum;
  DevBuf<uint32_t> umax;
  ScanWorkspace scan;
};

// knn_rows_include_self (lib.rs:3402-3417) over device kNN.
bool knn_include_self_device(const int32_t* idx, const float* dist, int64_t n, int k,
                             GraphWorkspace& ws, hipStream_t s);

// smooth_knn_dist (lib.rs:3424-3523).  fp64_sums selects FP64 bisection sums
// (default) instead of the reference's f32 sums.
void smooth_knn_dist_device(const float* dist, int64_t n, int k, float k_value,
                            float local_connectivity, float bandwidth, bool include_self,
                            bool fp64_sums, float* sigmas, float* rhos, GraphWorkspace& ws,
                            hipStream_t s);

// compute_membership_strengths + symmetrize_fuzzy_graph + prune_edges
// (lib.rs:3525-3560, 3597-3649, 3804-3822).  prune_epochs <= 0 skips prune.
void fuzzy_graph_device(const int32_t* idx, const float* dist, const float* sigmas,
                        const float* rhos, int64_t n, int k, float set_op_mix_ratio,
                        int prune_epochs, Graph& g, GraphWorkspace& ws, hipStream_t s);

// compute_membership_strengths_bipartite + prune_edges for transform
// (lib.rs:3562-3595, 3804-3822): edges grouped by query head.
void bipartite_graph_device(const int32_t* idx, const float* dist, const float* sigmas,
                            const float* rhos, int64_t m, int k, int prune_epochs, Graph& g,
                            GraphWorkspace& ws, hipStream_t s);

// make_epochs_per_sample (lib.rs:4937-4960) with the graph's max weight.
void epochs_per_sample_device(const float* w, int64_t n_edges, float max_weight, int n_epochs,
                              float* eps, hipStream_t s);

// local_radii_from_knn / embedding_radii_from_knn (lib.rs:5057-5108).
void local_radii_device(const int32_t* idx, const float* dist, int64_t n, int k, float var_shift,
                        float* out, hipStream_t s);
void embedding_radii_device(const float* y, int64_t n, int dim, const int32_t* idx, int k,
                            float var_shift, float* out, hipStream_t s);

}  // namespace umapx
