#include "mlp.hpp"

namespace umapx {

void mlp_train(MlpModel&, const float*, int64_t, int64_t, const float*, int, const Graph&, float,
               float, const umapx_params&, hipStream_t) {
  throw_invalid("learned transform is not available in this build");
}

void mlp_forward_rows(const MlpModel&, const float*, int64_t, float*, hipStream_t) {
  throw_invalid("learned transform is not available in this build");
}

}  // namespace umapx
