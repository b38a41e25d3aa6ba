// Learned transform: an MLP encoder d -> H -> H -> n_components (ReLU), BF16
// weights and activations with FP32 accumulation and FP32 master weights,
// trained on the UMAP cross-entropy over the fitted fuzzy graph with HIP
// forward / backward kernels.
#pragma once

#include "../../include/umapx.h"
#include "common.hpp"
#include "graph.hpp"

namespace umapx {

struct MlpModel {
  bool trained = false;
  int d = 0, h = 0, out = 0;
  // FP32 master weights, row-major [in x out]
  DevBuf<float> w1, b1, w2, b2, w3, b3;
  // BF16 copies used by the forward / backward GEMMs
  DevBuf<uint16_t> w1h, w2h, w3h;
  // input standardisation
  DevBuf<float> mean, inv_std;
  double last_loss = 0.0;
  int steps_done = 0;
};

void mlp_train(MlpModel& m, const float* x, int64_t n, int64_t d, const float* y, int dim,
               const Graph& g, float a, float b, const umapx_params& p, hipStream_t s);
void mlp_forward_rows(const MlpModel& m, const float* x, int64_t rows, float* out, hipStream_t s);

}  // namespace umapx
