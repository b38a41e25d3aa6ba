// Learned transform: an MLP encoder d -> H -> H -> n_components (ReLU), BF16
// weights and activations with FP32 accumulation and FP32 master weights,
// trained on the UMAP cross-entropy over the fitted fuzzy graph plus a
// squared error to the fitted embedding, with HIP forward / backward kernels
// (BF16 MFMA GEMMs for the two wide layers) and Adam.
//
// The reference's analog is ParametricUmapModel (rust_umap/src/parametric.rs:
// a d -> H (tanh) -> out network regressed on a teacher UMAP embedding,
// :89-534); this encoder keeps the teacher target as the squared-error term
// and adds the UMAP cross-entropy over the teacher's fuzzy graph.
#pragma once

#include "../../include/umapx.h"
#include "common.hpp"
#include "graph.hpp"

namespace umapx {

// Flat parameter layout (f32): W1 [dp x h], b1 [h], W2 [h x h], b2 [h],
// W3 [h x dim], b3 [dim]; dp = d rounded up to a multiple of 32 (rows of W1
// past d stay zero).
struct MlpShape {
  int d = 0, dp = 0, h = 0, dim = 0;
  int64_t off_w1 = 0, off_b1 = 0, off_w2 = 0, off_b2 = 0, off_w3 = 0, off_b3 = 0, total = 0;
  static MlpShape make(int d, int h, int dim);
};

struct MlpModel {
  bool trained = false;
  MlpShape shape;
  DevBuf<float> params, grads, adam_m, adam_v;
  DevBuf<uint16_t> wh;  // BF16 copies of W1, W2, W3 (same offsets as params)
  DevBuf<float> mean, inv_std;  // input standardisation over the training rows
  // the BF16 weights packed for the fused inference kernel (see mlp.hip)
  bool packed = false;
  DevBuf<uint16_t> w1p, w2p, w3p;
  double last_loss = 0.0, first_loss = 0.0;
  int steps_done = 0;
};

void mlp_train(MlpModel& m, const float* x, int64_t n, int64_t d, const float* y, int dim,
               const Graph& g, float a, float b, const umapx_params& p, hipStream_t s);
void mlp_forward_rows(const MlpModel& m, const float* x, int64_t rows, float* out, hipStream_t s);

// Forward pass of given parameters over host rows x [rows x d] into host z
// [rows x dim], for tests: fused = false runs the unfused chain (the
// training forward), fused = true the fused inference kernel (an error where
// the fused kernel does not apply).
void mlp_forward_host(int d, int h, int dim, const float* params, const float* mean,
                      const float* inv_std, const float* x, int64_t rows, bool fused, float* z);

// One forward / backward pass on a given batch (host arrays), for gradient
// checks: rows are laid out as B heads, B tails, then B * n_neg negatives
// (negative t of edge e at 2B + e * n_neg + t).
void mlp_batch_grad(int d, int h, int dim, const float* params, const float* mean,
                    const float* inv_std, const float* x_rows, const float* y_rows, int n_edges,
                    int n_neg, float a, float b, float mse_weight, float* grads, float* loss,
                    float* z_out);

}  // namespace umapx
