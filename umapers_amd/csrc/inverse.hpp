// Inverse transform on the GPU: embedded points back to the training feature
// space (inverse_transform_rows, lib.rs:1800-1915).
#pragma once

#include "common.hpp"
#include "knn.hpp"

namespace umapx {

struct InverseArgs {
  int k = 15;             // inverse_neighbor_count (lib.rs:3420-3422)
  int n_epochs = 100;     // 0 stops after the local linear initialisation
  float a = 1.f, b = 1.f;
  float learning_rate = 0.25f;  // the model's learning rate / 4
  int negative_sample_rate = 5;
  float repulsion_strength = 1.f;
  uint64_t seed = 0;
};

// Largest feature count the inverse SGD holds in registers (up to 40 per
// lane of a 64-lane group).
constexpr int64_t kInverseMaxFeatures = 2560;

// q [m x dim] embedded rows (device), y_train [n x dim] the fitted embedding,
// x_train [n x d] the training rows, sigmas / rhos [n] the fit's; out [m x d].
// sigmas / rhos may be null when n_epochs == 0.
void inverse_transform_device(const float* q, int64_t m, int dim, const float* y_train,
                              const float* x_train, int64_t n, int64_t d, const float* sigmas,
                              const float* rhos, const InverseArgs& args, float* out,
                              KnnWorkspace& kws, hipStream_t s);

}  // namespace umapx
