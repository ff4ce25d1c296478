// Multilayer perceptron of the learned contraction-plan selector.
#pragma once

#include "common.hpp"

namespace xt {

constexpr int kMlpMaxWidth = 64;
constexpr int kMlpMaxHidden = 3;

struct MlpShape {
  int nin = 0;     // inputs (<= kMlpMaxWidth)
  int hidden = 1;  // hidden ReLU layers (1 .. kMlpMaxHidden)
  int width = 64;  // units per hidden layer (<= kMlpMaxWidth)
};

// Device work buffers of one batch size.
struct MlpWork {
  int rows = 0;
  float *pre[kMlpMaxHidden] = {};   // pre-activations, rows*width each
  float *post[kMlpMaxHidden] = {};  // activations, rows*width each
  float *dz[2] = {};                // gradient ping-pong, rows*width each
  float *out = nullptr;             // rows
  float *part = nullptr;            // per-block loss partials
  int part_cap = 0;
};

void mlp_layer_offsets(const MlpShape &s, int l, size_t *w_off, size_t *b_off, int *nin, int *nout);
size_t mlp_num_params(const MlpShape &s);
int mlp_forward(const MlpShape &s, const float *params, const float *x, int rows, float *const *pre,
                float *const *post, float *out, hipStream_t st);
int mlp_backward(const MlpShape &s, const float *params, const float *x, const float *y, int rows, MlpWork &ws,
                 float *grad, float *loss, hipStream_t st);
int mlp_adam(float *p, const float *g, float *m, float *v, int n, float lr_t, float b1, float b2, float eps,
             hipStream_t st);
int mlp_gather(const float *x, const float *y, const int *idx, float *xb, float *yb, int rows, int nin,
               hipStream_t st);

// Host-side forward of one input (plan-time inference, no device work).
float mlp_forward_host(const MlpShape &s, const float *params, const float *x);

}  // namespace xt
