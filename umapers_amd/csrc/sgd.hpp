// Layout optimisation on the GPU: edge-parallel Hogwild SGD epochs with the
// reference's exact per-edge f32 schedule, and the per-query transform SGD.
#pragma once

#include "common.hpp"
#include "graph.hpp"

namespace umapx {

struct SgdArgs {
  int n_epochs = 200;
  float a = 1.f, b = 1.f;
  float learning_rate = 1.f;
  int negative_sample_rate = 5;
  float repulsion_strength = 1.f;
  uint64_t seed = 0;
  // densMAP term (lib.rs:5110-5134); radii == nullptr disables it
  const float* radii = nullptr;
  int dens_start_epoch = 0;
  float dens_lambda = 0.f;
  float dens_var_shift = 0.1f;
  // the edge list is the fit's fuzzy union: every edge (h, t) has exactly one
  // reverse (t, h) of equal weight.  A 2-D tail move is then made by the
  // reverse edge's own thread instead of a float atomic.
  bool symmetric = false;
  // transform walk: the largest number of edges of one query row (its kNN
  // width), 0 when unknown; rows of at most 32 edges take the LDS-staged walk
  int max_row_edges = 0;
  // optional per-edge sample counters (tests of the exact schedule)
  int32_t* pos_counts = nullptr;
  int32_t* neg_counts = nullptr;
};

struct SgdWorkspace {
  DevBuf<float> eps, eons, eonns;
  DevBuf<int32_t> chunk_row, chunk_start;  // (head, first edge) of each walked chunk
};

struct SgdStats {
  hipEvent_t start = nullptr, stop = nullptr;
  bool recorded = false;
};

// optimize_layout_training(_flat) (lib.rs:5137-5441) over a head-sorted graph.
void optimize_layout_device(float* y, int64_t n, int dim, const Graph& g, const SgdArgs& args,
                            SgdWorkspace& ws, hipStream_t s, SgdStats* stats);

// Positive updates and negative samples summed over all epochs of the last
// optimize_layout_device on this workspace (the exact schedule, replayed).
void schedule_totals_device(const SgdWorkspace& ws, int64_t ne, int n_epochs,
                            int negative_sample_rate, hipStream_t s, int64_t* positive_updates,
                            int64_t* negative_samples);

// optimize_layout_transform (lib.rs:5444-5544): only query rows move.
void optimize_transform_device(float* yq, int64_t m, int dim, const float* y_base, int64_t n_base,
                               const Graph& g, const SgdArgs& args, SgdWorkspace& ws,
                               hipStream_t s);

// init_graph_transform (lib.rs:4793-4836).
void init_graph_transform_device(int64_t m, int dim, const Graph& g, const float* y_base,
                                 float* yq, hipStream_t s);

}  // namespace umapx
