// C-ABI of the learned contraction-plan selector: the MLP model object, its
// training on the GPU (k_mlp.hip kernels), persistence, candidate features and
// installation into the contraction planner.
//
// No counterpart exists in the reference; it learns what the fixed heuristics
// of TALSH/talshc.cpp:366-488 and TALSH/tensor_algebra_gpu_nvidia.cu:1188-1272
// decide by rule.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "contract_plan.hpp"
#include "kernels.hpp"
#include "runtime_impl.hpp"
#include "selector.hpp"
#include "talsh_amd.h"

namespace xt {
namespace {

uint64_t splitmix64(uint64_t &s)
{
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Device-side state for batches of up to `rows` rows.
struct DevBuf {
  std::vector<void *> ptrs;
  ~DevBuf()
  {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T *get(size_t n)
  {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return static_cast<T *>(p);
  }
};

bool alloc_work(DevBuf &db, const MlpShape &s, int rows, MlpWork &w)
{
  w.rows = rows;
  const size_t rw = (size_t)rows * s.width;
  for (int l = 0; l < s.hidden; ++l) {
    if (!(w.pre[l] = db.get<float>(rw)) || !(w.post[l] = db.get<float>(rw))) return false;
  }
  if (!(w.dz[0] = db.get<float>(rw)) || !(w.dz[1] = db.get<float>(rw)) || !(w.out = db.get<float>(rows)))
    return false;
  w.part_cap = (rows + 255) / 256;
  return (w.part = db.get<float>(w.part_cap)) != nullptr;
}

struct GpuScope {
  bool ok = false;
  hipStream_t s = nullptr;
  GpuScope()
  {
    Engine &E = engine();
    if (!E.on || E.gpus.empty()) return;
    if (hipSetDevice(E.gpus[0]->ordinal) != hipSuccess) return;
    s = E.gpus[0]->streams[0];
    ok = true;
  }
};

std::vector<float> normalized(const MlpModel &m, const float *x, long long n)
{
  std::vector<float> out((size_t)n * m.s.nin);
  for (long long r = 0; r < n; ++r) m.normalize(x + r * m.s.nin, out.data() + r * m.s.nin);
  return out;
}

}  // namespace
}  // namespace xt

using namespace xt;

extern "C" {

int talshAmdMlpCreate(int nin, int hidden_layers, int width, unsigned long long seed, void **mlp)
{
  if (!mlp || nin < 1 || nin > kMlpMaxWidth || hidden_layers < 1 || hidden_layers > kMlpMaxHidden || width < 1 ||
      width > kMlpMaxWidth)
    return TALSH_INVALID_ARGS;
  MlpModel *m = new MlpModel();
  m->s.nin = nin;
  m->s.hidden = hidden_layers;
  m->s.width = width;
  m->params.assign(mlp_num_params(m->s), 0.f);
  m->mean.assign(nin, 0.f);
  m->stdv.assign(nin, 1.f);
  // He-uniform weights, zero biases
  uint64_t st = seed;
  for (int l = 0; l <= m->s.hidden; ++l) {
    size_t wo, bo;
    int a, c;
    mlp_layer_offsets(m->s, l, &wo, &bo, &a, &c);
    const double lim = std::sqrt((l == m->s.hidden ? 3.0 : 6.0) / a);
    for (size_t i = 0; i < (size_t)a * c; ++i) {
      const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
      m->params[wo + i] = (float)((2.0 * u - 1.0) * lim);
    }
  }
  *mlp = m;
  return 0;
}

int talshAmdMlpDestroy(void *mlp)
{
  delete static_cast<MlpModel *>(mlp);
  return 0;
}

int talshAmdMlpNumParams(void *mlp, int *nparams)
{
  if (!mlp || !nparams) return TALSH_INVALID_ARGS;
  *nparams = (int)static_cast<MlpModel *>(mlp)->params.size();
  return 0;
}

int talshAmdMlpGetParams(void *mlp, float *params)
{
  if (!mlp || !params) return TALSH_INVALID_ARGS;
  const MlpModel *m = static_cast<MlpModel *>(mlp);
  std::memcpy(params, m->params.data(), m->params.size() * sizeof(float));
  return 0;
}

int talshAmdMlpSetParams(void *mlp, const float *params)
{
  if (!mlp || !params) return TALSH_INVALID_ARGS;
  MlpModel *m = static_cast<MlpModel *>(mlp);
  std::memcpy(m->params.data(), params, m->params.size() * sizeof(float));
  return 0;
}

int talshAmdMlpSetNormalization(void *mlp, const float *x_mean, const float *x_std, float y_mean, float y_std)
{
  if (!mlp || !x_mean || !x_std || !(y_std > 0.f)) return TALSH_INVALID_ARGS;
  MlpModel *m = static_cast<MlpModel *>(mlp);
  for (int i = 0; i < m->s.nin; ++i) {
    if (!(x_std[i] > 0.f)) return TALSH_INVALID_ARGS;
    m->mean[i] = x_mean[i];
    m->stdv[i] = x_std[i];
  }
  m->ymean = y_mean;
  m->ystd = y_std;
  return 0;
}

int talshAmdMlpGradient(void *mlp, const float *x, const float *y, int n, float *grad, float *loss)
{
  if (!mlp || !x || !y || !grad || !loss || n < 1) return TALSH_INVALID_ARGS;
  const MlpModel *m = static_cast<MlpModel *>(mlp);
  GpuScope g;
  if (!g.ok) return TALSH_NOT_INITIALIZED;
  const std::vector<float> xn = normalized(*m, x, n);
  std::vector<float> yn(n);
  for (int r = 0; r < n; ++r) yn[r] = (y[r] - m->ymean) / m->ystd;
  DevBuf db;
  MlpWork w;
  float *dx = db.get<float>(xn.size()), *dy = db.get<float>(n), *dp = db.get<float>(m->params.size()),
        *dg = db.get<float>(m->params.size()), *dl = db.get<float>(1);
  if (!dx || !dy || !dp || !dg || !dl || !alloc_work(db, m->s, n, w)) return TRY_LATER;
  bool ok = XT_HIP_OK(hipMemcpyAsync(dx, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, g.s)) &&
            XT_HIP_OK(hipMemcpyAsync(dy, yn.data(), (size_t)n * 4, hipMemcpyHostToDevice, g.s)) &&
            XT_HIP_OK(hipMemcpyAsync(dp, m->params.data(), m->params.size() * 4, hipMemcpyHostToDevice, g.s));
  if (!ok) return TALSH_FAILURE;
  int e = mlp_backward(m->s, dp, dx, dy, n, w, dg, dl, g.s);
  if (e) return e;
  ok = XT_HIP_OK(hipMemcpyAsync(grad, dg, m->params.size() * 4, hipMemcpyDeviceToHost, g.s)) &&
       XT_HIP_OK(hipMemcpyAsync(loss, dl, 4, hipMemcpyDeviceToHost, g.s)) && XT_HIP_OK(hipStreamSynchronize(g.s));
  return ok ? 0 : TALSH_FAILURE;
}

int talshAmdMlpPredict(void *mlp, const float *x, long long n, float *y)
{
  if (!mlp || !x || !y || n < 0) return TALSH_INVALID_ARGS;
  if (n == 0) return 0;
  const MlpModel *m = static_cast<MlpModel *>(mlp);
  GpuScope g;
  if (!g.ok) return TALSH_NOT_INITIALIZED;
  const int chunk = (int)std::min<long long>(n, 65536);
  DevBuf db;
  MlpWork w;
  float *dx = db.get<float>((size_t)chunk * m->s.nin), *dp = db.get<float>(m->params.size());
  if (!dx || !dp || !alloc_work(db, m->s, chunk, w)) return TRY_LATER;
  if (!XT_HIP_OK(hipMemcpyAsync(dp, m->params.data(), m->params.size() * 4, hipMemcpyHostToDevice, g.s)))
    return TALSH_FAILURE;
  std::vector<float> out(chunk);
  for (long long r0 = 0; r0 < n; r0 += chunk) {
    const int rows = (int)std::min<long long>(chunk, n - r0);
    const std::vector<float> xn = normalized(*m, x + r0 * m->s.nin, rows);
    if (!XT_HIP_OK(hipMemcpyAsync(dx, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, g.s))) return TALSH_FAILURE;
    int e = mlp_forward(m->s, dp, dx, rows, w.pre, w.post, w.out, g.s);
    if (e) return e;
    if (!XT_HIP_OK(hipMemcpyAsync(out.data(), w.out, (size_t)rows * 4, hipMemcpyDeviceToHost, g.s)) ||
        !XT_HIP_OK(hipStreamSynchronize(g.s)))
      return TALSH_FAILURE;
    for (int r = 0; r < rows; ++r) y[r0 + r] = out[r] * m->ystd + m->ymean;
  }
  return 0;
}

int talshAmdMlpTrain(void *mlp, const float *x, const float *y, long long n, int epochs, int batch, float lr,
                     unsigned long long seed, float *epoch_loss)
{
  if (!mlp || !x || !y || n < 2 || epochs < 1 || batch < 1 || batch > (1 << 16) || !(lr > 0.f))
    return TALSH_INVALID_ARGS;
  if (n > (1ll << 31) - 1) return TALSH_LIMIT_EXCEEDED;
  MlpModel *m = static_cast<MlpModel *>(mlp);
  GpuScope g;
  if (!g.ok) return TALSH_NOT_INITIALIZED;
  const int nin = m->s.nin;
  // normalization from the training set
  for (int i = 0; i < nin; ++i) {
    double s = 0, s2 = 0;
    for (long long r = 0; r < n; ++r) s += x[r * nin + i];
    const double mu = s / n;
    for (long long r = 0; r < n; ++r) s2 += (x[r * nin + i] - mu) * (x[r * nin + i] - mu);
    const double sd = std::sqrt(s2 / n);
    m->mean[i] = (float)mu;
    m->stdv[i] = sd > 1e-6 ? (float)sd : 1.f;
  }
  {
    double s = 0, s2 = 0;
    for (long long r = 0; r < n; ++r) s += y[r];
    const double mu = s / n;
    for (long long r = 0; r < n; ++r) s2 += (y[r] - mu) * (y[r] - mu);
    const double sd = std::sqrt(s2 / n);
    m->ymean = (float)mu;
    m->ystd = sd > 1e-6 ? (float)sd : 1.f;
  }
  const std::vector<float> xn = normalized(*m, x, n);
  std::vector<float> yn(n);
  for (long long r = 0; r < n; ++r) yn[r] = (y[r] - m->ymean) / m->ystd;
  const size_t np = m->params.size();
  const int nbatches = (int)((n + batch - 1) / batch);
  DevBuf db;
  MlpWork w;
  float *dx = db.get<float>(xn.size()), *dy = db.get<float>(n), *dp = db.get<float>(np), *dg = db.get<float>(np),
        *dm = db.get<float>(np), *dv = db.get<float>(np), *xb = db.get<float>((size_t)batch * nin),
        *yb = db.get<float>(batch), *dl = db.get<float>(nbatches);
  int *didx = db.get<int>(n);
  if (!dx || !dy || !dp || !dg || !dm || !dv || !xb || !yb || !dl || !didx || !alloc_work(db, m->s, batch, w))
    return TRY_LATER;
  bool ok = XT_HIP_OK(hipMemcpyAsync(dx, xn.data(), xn.size() * 4, hipMemcpyHostToDevice, g.s)) &&
            XT_HIP_OK(hipMemcpyAsync(dy, yn.data(), (size_t)n * 4, hipMemcpyHostToDevice, g.s)) &&
            XT_HIP_OK(hipMemcpyAsync(dp, m->params.data(), np * 4, hipMemcpyHostToDevice, g.s)) &&
            XT_HIP_OK(hipMemsetAsync(dm, 0, np * 4, g.s)) && XT_HIP_OK(hipMemsetAsync(dv, 0, np * 4, g.s));
  if (!ok) return TALSH_FAILURE;
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::mt19937_64 rng(seed);
  std::vector<float> bl(nbatches);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  long long step = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(idx.begin(), idx.end(), rng);
    if (!XT_HIP_OK(hipMemcpyAsync(didx, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, g.s))) return TALSH_FAILURE;
    for (int b = 0; b < nbatches; ++b) {
      const long long r0 = (long long)b * batch;
      const int rows = (int)std::min<long long>(batch, n - r0);
      int e = mlp_gather(dx, dy, didx + r0, xb, yb, rows, nin, g.s);
      if (!e) e = mlp_backward(m->s, dp, xb, yb, rows, w, dg, dl + b, g.s);
      ++step;
      const float lr_t = lr * std::sqrt(1.f - std::pow(b2, (float)step)) / (1.f - std::pow(b1, (float)step));
      if (!e) e = mlp_adam(dp, dg, dm, dv, (int)np, lr_t, b1, b2, eps, g.s);
      if (e) return e;
    }
    if (!XT_HIP_OK(hipMemcpyAsync(bl.data(), dl, (size_t)nbatches * 4, hipMemcpyDeviceToHost, g.s)) ||
        !XT_HIP_OK(hipStreamSynchronize(g.s)))
      return TALSH_FAILURE;
    if (epoch_loss) {
      double s = 0;
      for (int b = 0; b < nbatches; ++b) s += (double)bl[b] * std::min<long long>(batch, n - (long long)b * batch);
      epoch_loss[ep] = (float)(s / n);
    }
  }
  ok = XT_HIP_OK(hipMemcpyAsync(m->params.data(), dp, np * 4, hipMemcpyDeviceToHost, g.s)) &&
       XT_HIP_OK(hipStreamSynchronize(g.s));
  return ok ? 0 : TALSH_FAILURE;
}

int talshAmdMlpSave(void *mlp, const char *path)
{
  if (!mlp || !path) return TALSH_INVALID_ARGS;
  const MlpModel *m = static_cast<MlpModel *>(mlp);
  FILE *f = std::fopen(path, "w");
  if (!f) return TALSH_FAILURE;
  std::fprintf(f, "talsh_plan_mlp 1\n%d %d %d\n%.9g %.9g\n", m->s.nin, m->s.hidden, m->s.width, m->ymean, m->ystd);
  for (float v : m->mean) std::fprintf(f, "%.9g ", v);
  std::fprintf(f, "\n");
  for (float v : m->stdv) std::fprintf(f, "%.9g ", v);
  std::fprintf(f, "\n%zu\n", m->params.size());
  for (size_t i = 0; i < m->params.size(); ++i) std::fprintf(f, "%.9g%c", m->params[i], (i % 8 == 7) ? '\n' : ' ');
  std::fprintf(f, "\n");
  return std::fclose(f) == 0 ? 0 : TALSH_FAILURE;
}

int talshAmdMlpLoad(const char *path, void **mlp)
{
  if (!path || !mlp) return TALSH_INVALID_ARGS;
  FILE *f = std::fopen(path, "r");
  if (!f) return TALSH_FAILURE;
  char tag[32] = {};
  int ver = 0, nin = 0, hid = 0, wid = 0;
  float ym = 0, ys = 1;
  bool ok = std::fscanf(f, "%31s %d %d %d %d %f %f", tag, &ver, &nin, &hid, &wid, &ym, &ys) == 7 &&
            std::strcmp(tag, "talsh_plan_mlp") == 0 && ver == 1;
  void *h = nullptr;
  if (ok) ok = talshAmdMlpCreate(nin, hid, wid, 0, &h) == 0;
  MlpModel *m = static_cast<MlpModel *>(h);
  if (ok) {
    m->ymean = ym;
    m->ystd = ys;
    for (int i = 0; ok && i < nin; ++i) ok = std::fscanf(f, "%f", &m->mean[i]) == 1;
    for (int i = 0; ok && i < nin; ++i) ok = std::fscanf(f, "%f", &m->stdv[i]) == 1 && m->stdv[i] > 0.f;
    size_t np = 0;
    ok = ok && std::fscanf(f, "%zu", &np) == 1 && np == m->params.size();
    for (size_t i = 0; ok && i < np; ++i) ok = std::fscanf(f, "%f", &m->params[i]) == 1;
  }
  std::fclose(f);
  if (!ok) {
    delete m;
    return TALSH_INVALID_ARGS;
  }
  *mlp = m;
  return 0;
}

int talshAmdPlanFeatureCount() { return kPlanFeatures; }

int talshAmdPlanFeatures(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                         const int *rdims, int data_kind, int accumulative, int fast_math, int cap, int *variants,
                         int *tiles, float *features, int *ncand)
{
  if (!cptrn || !ncand || (cap > 0 && (!variants || !tiles || !features))) return TALSH_INVALID_ARGS;
  int digc[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, digc, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr != drank || lr != lrank || rr != rrank || dr == 0 || lr == 0 || rr == 0) return TALSH_INVALID_ARGS;
  std::vector<int> dig(digc, digc + lr + rr), dd(ddims, ddims + dr), ld(ldims, ldims + lr), rd(rdims, rdims + rr);
  Engine &E = engine();
  const int cus = (E.on && !E.gpus.empty()) ? E.gpus[0]->num_cus : 256;
  std::vector<PlanCandidate> c;
  const R4Mode mode = fast_math ? R4Mode::kBF16x3 : R4Mode::kNativeF32;
  const int e = plan_candidates(dig, dr, lr, rr, dd, ld, rd, data_kind, mode, accumulative != NOPE, cus, c);
  if (e) return e;
  *ncand = (int)c.size();
  for (int i = 0; i < (int)c.size() && i < cap; ++i) {
    variants[i] = c[i].variant;
    tiles[i] = c[i].tile;
    std::memcpy(features + (size_t)i * kPlanFeatures, c[i].f, sizeof(float) * kPlanFeatures);
  }
  return 0;
}

int talshAmdSelectorInstall(void *mlp)
{
  const MlpModel *m = static_cast<const MlpModel *>(mlp);
  if (m && m->s.nin != kPlanFeatures) return TALSH_INVALID_ARGS;
  selector_install(m);
  return 0;
}

int talshAmdSelectorChoice(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                           const int *rdims, int data_kind, int accumulative, int fast_math, int *variant, int *tile)
{
  if (!cptrn || !variant || !tile) return TALSH_INVALID_ARGS;
  int digc[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, digc, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr != drank || lr != lrank || rr != rrank || dr == 0 || lr == 0 || rr == 0) return TALSH_INVALID_ARGS;
  std::vector<int> dig(digc, digc + lr + rr), dd(ddims, ddims + dr), ld(ldims, ldims + lr), rd(rdims, rdims + rr);
  Engine &E = engine();
  const int cus = (E.on && !E.gpus.empty()) ? E.gpus[0]->num_cus : 256;
  const R4Mode mode = fast_math ? R4Mode::kBF16x3 : R4Mode::kNativeF32;
  if (!selector_choose(dig, dr, lr, rr, dd, ld, rd, data_kind, mode, accumulative != NOPE, cus, variant, tile)) {
    // the heuristic: fewest transpose bytes, default tile for the GEMM shape
    ContractPlan P;
    const int e = plan_contraction(dig, dr, lr, rr, dd, ld, rd, kind_size(data_kind), accumulative != NOPE, -1, P,
                                   nullptr);
    if (e) return e;
    *variant = 0;
    *tile = gemm_default_tile(data_kind, mode, P.swap ? P.n : P.m, P.swap ? P.m : P.n, P.k);
  }
  return 0;
}

}  // extern "C"
