// Multilayer perceptron of the learned contraction-plan selector: forward,
// backward and Adam kernels (FP32), gfx950.
//
// The network is small (inputs <= 64, 1-3 hidden ReLU layers of <= 64 units,
// one linear output), so every kernel maps one thread to one output element
// and loops over the reduction dimension in a fixed order: results are
// deterministic (no atomics) and independent of the launch shape.
//
// Parameter layout (shared with the host-side forward in selector.cpp): for
// each layer in order, W[in][out] (row-major, out fastest) followed by b[out].
#include "mlp.hpp"

namespace xt {
namespace {

constexpr int kThreads = 256;

inline int nblocks(long long n) { return (int)((n + kThreads - 1) / kThreads); }

// y[r][o] = act(sum_i x[r][i] * W[i][o] + b[o]);  z keeps the pre-activation.
__global__ void k_dense_fwd(const float *__restrict__ x, const float *__restrict__ W, const float *__restrict__ b,
                            float *__restrict__ z, float *__restrict__ y, int rows, int nin, int nout, int relu)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rows * nout) return;
  const int r = (int)(t / nout), o = (int)(t % nout);
  const float *xr = x + (long long)r * nin;
  float s = b[o];
  for (int i = 0; i < nin; ++i) s = __fmaf_rn(xr[i], W[i * nout + o], s);
  if (z) z[t] = s;
  y[t] = relu ? fmaxf(s, 0.f) : s;
}

// Mean squared error of a single-output network: g[r] = 2 (p[r] - y[r]) / n,
// per-block partial losses into part[blockIdx.x] (reduced on the host side
// of the launcher by k_sum).
__global__ void k_mse(const float *__restrict__ p, const float *__restrict__ y, float *__restrict__ g,
                      float *__restrict__ part, int n)
{
  __shared__ float red[kThreads];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (t < n) {
    const float d = p[t] - y[t];
    g[t] = 2.f * d / (float)n;
    l = d * d;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out[0] = sum(part[0..n)) / denom (one block, fixed order).
__global__ void k_sum(const float *__restrict__ part, int n, float denom, float *__restrict__ out)
{
  __shared__ float red[kThreads];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += kThreads) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = kThreads / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] / denom;
}

// dW[i][o] = sum_r x[r][i] * dz[r][o];  db[o] = sum_r dz[r][o]
__global__ void k_dense_bwd_w(const float *__restrict__ x, const float *__restrict__ dz, float *__restrict__ dW,
                              float *__restrict__ db, int rows, int nin, int nout)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nin * nout) {
    const int i = t / nout, o = t % nout;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s = __fmaf_rn(x[(long long)r * nin + i], dz[(long long)r * nout + o], s);
    dW[t] = s;
  } else if (t < nin * nout + nout) {
    const int o = t - nin * nout;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += dz[(long long)r * nout + o];
    db[o] = s;
  }
}

// dzp[r][i] = (zp[r][i] > 0) * sum_o dz[r][o] * W[i][o]   (ReLU of the previous layer)
__global__ void k_dense_bwd_x(const float *__restrict__ dz, const float *__restrict__ W, const float *__restrict__ zp,
                              float *__restrict__ dzp, int rows, int nin, int nout)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rows * nin) return;
  const int r = (int)(t / nin), i = (int)(t % nin);
  const float *dr = dz + (long long)r * nout;
  float s = 0.f;
  for (int o = 0; o < nout; ++o) s = __fmaf_rn(dr[o], W[i * nout + o], s);
  dzp[t] = zp[t] > 0.f ? s : 0.f;
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int n, float lr_t, float b1, float b2, float eps)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr_t * mt / (sqrtf(vt) + eps);
}

// xb[r][:] = x[idx[r]][:], yb[r] = y[idx[r]]
__global__ void k_gather_rows(const float *__restrict__ x, const float *__restrict__ y, const int *__restrict__ idx,
                              float *__restrict__ xb, float *__restrict__ yb, int rows, int nin)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rows * nin) return;
  const int r = (int)(t / nin), i = (int)(t % nin);
  const long long src = idx[r];
  xb[t] = x[src * nin + i];
  if (i == 0) yb[r] = y[src];
}

}  // namespace

// Offsets of layer l's W and b inside the flat parameter vector.
void mlp_layer_offsets(const MlpShape &s, int l, size_t *w_off, size_t *b_off, int *nin, int *nout)
{
  size_t off = 0;
  for (int q = 0; q <= l; ++q) {
    const int a = q == 0 ? s.nin : s.width;
    const int c = q == s.hidden ? 1 : s.width;
    if (q == l) {
      *w_off = off;
      *b_off = off + (size_t)a * c;
      *nin = a;
      *nout = c;
      return;
    }
    off += (size_t)a * c + c;
  }
}

size_t mlp_num_params(const MlpShape &s)
{
  size_t w, b;
  int a, c;
  mlp_layer_offsets(s, s.hidden, &w, &b, &a, &c);
  return b + (size_t)c;
}

// Forward over `rows` inputs already on the device. acts must hold
// (hidden + 1) buffers of rows*width floats (pre-activations) and the same
// number of post-activation buffers; out receives rows outputs.
int mlp_forward(const MlpShape &s, const float *params, const float *x, int rows, float *const *pre,
                float *const *post, float *out, hipStream_t st)
{
  const float *in = x;
  for (int l = 0; l <= s.hidden; ++l) {
    size_t wo, bo;
    int a, c;
    mlp_layer_offsets(s, l, &wo, &bo, &a, &c);
    const bool last = l == s.hidden;
    float *y = last ? out : post[l];
    hipLaunchKernelGGL(k_dense_fwd, dim3(nblocks((long long)rows * c)), dim3(kThreads), 0, st, in, params + wo,
                       params + bo, last ? nullptr : pre[l], y, rows, a, c, last ? 0 : 1);
    in = y;
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

// Loss (mean squared error) and its gradient with respect to every
// parameter for a batch on the device. Work buffers come from ws.
int mlp_backward(const MlpShape &s, const float *params, const float *x, const float *y, int rows, MlpWork &ws,
                 float *grad, float *loss, hipStream_t st)
{
  int e = mlp_forward(s, params, x, rows, ws.pre, ws.post, ws.out, st);
  if (e) return e;
  const int nb = nblocks(rows);
  if (nb > ws.part_cap) return TALSH_INVALID_ARGS;
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(kThreads), 0, st, ws.out, y, ws.dz[0], ws.part, rows);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(kThreads), 0, st, ws.part, nb, (float)rows, loss);
  // dz of the output layer lives in ws.dz[0]; alternate buffers going down
  int cur = 0;
  for (int l = s.hidden; l >= 0; --l) {
    size_t wo, bo;
    int a, c;
    mlp_layer_offsets(s, l, &wo, &bo, &a, &c);
    const float *in = l == 0 ? x : ws.post[l - 1];
    hipLaunchKernelGGL(k_dense_bwd_w, dim3(nblocks((long long)a * c + c)), dim3(kThreads), 0, st, in, ws.dz[cur],
                       grad + wo, grad + bo, rows, a, c);
    if (l > 0) {
      hipLaunchKernelGGL(k_dense_bwd_x, dim3(nblocks((long long)rows * a)), dim3(kThreads), 0, st, ws.dz[cur],
                         params + wo, ws.pre[l - 1], ws.dz[cur ^ 1], rows, a, c);
      cur ^= 1;
    }
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

int mlp_adam(float *p, const float *g, float *m, float *v, int n, float lr_t, float b1, float b2, float eps,
             hipStream_t st)
{
  hipLaunchKernelGGL(k_adam, dim3(nblocks(n)), dim3(kThreads), 0, st, p, g, m, v, n, lr_t, b1, b2, eps);
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

int mlp_gather(const float *x, const float *y, const int *idx, float *xb, float *yb, int rows, int nin,
               hipStream_t st)
{
  hipLaunchKernelGGL(k_gather_rows, dim3(nblocks((long long)rows * nin)), dim3(kThreads), 0, st, x, y, idx, xb, yb,
                     rows, nin);
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

}  // namespace xt
