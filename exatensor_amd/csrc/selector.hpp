// Learned plan selector: model object shared by the training entry points and
// the contraction planner.
#pragma once

#include <vector>

#include "mlp.hpp"

namespace xt {

struct MlpModel {
  MlpShape s;
  std::vector<float> params;      // layout of mlp_layer_offsets
  std::vector<float> mean, stdv;  // input normalization: (x - mean) / stdv
  float ymean = 0.f, ystd = 1.f;  // output de-normalization: y * ystd + ymean
  // raw features -> normalized inputs
  void normalize(const float *x, float *out) const
  {
    for (int i = 0; i < s.nin; ++i) out[i] = (x[i] - mean[i]) / stdv[i];
  }
  float predict_host(const float *x) const;
};

// Install a copy of m as the planner's selector (nullptr removes it).
void selector_install(const MlpModel *m);
bool selector_installed();

}  // namespace xt
