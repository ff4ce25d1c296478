// Element-wise tensor-body kernels for gfx950: fill (talshTensorInit), scale
// (talshTensorScale), scaled accumulate (talshTensorAdd and the rank-0 operand
// case of a contraction), and the full contraction to a scalar.
//
// Replace gpu_array_init__/gpu_array_scale__/gpu_array_add__/
// gpu_array_dot_product__ (TALSH/tensor_algebra_gpu_nvidia.cu:337-683).
// Arithmetic uses explicitly rounded products and sums (no fused multiply-add)
// so results equal the CPU path's separate multiply and add bit for bit.
#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.hpp"

// Products and sums below are rounded separately: no contraction into FMA.
#pragma clang fp contract(off)

namespace xt {
namespace {

constexpr int kNT = 256;

// (written as plain operators under the no-contraction pragma above: the
// library's __fmul_rn-style helpers are defined before it and may be fused)
__device__ __forceinline__ float mul(float a, float b) { return a * b; }
__device__ __forceinline__ double mul(double a, double b) { return a * b; }
__device__ __forceinline__ float add(float a, float b) { return a + b; }
__device__ __forceinline__ double add(double a, double b) { return a + b; }
__device__ __forceinline__ float sub(float a, float b) { return a - b; }
__device__ __forceinline__ double sub(double a, double b) { return a - b; }

template <typename R>
struct Cx { R re, im; };

template <typename T>
__global__ __launch_bounds__(kNT) void k_fill(T *__restrict__ d, uint64_t n, T v)
{
  const uint64_t st = (uint64_t)gridDim.x * kNT;
  for (uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += st) d[i] = v;
}

template <typename R>
__global__ __launch_bounds__(kNT) void k_scale_real(R *__restrict__ d, uint64_t n, R f)
{
  const uint64_t st = (uint64_t)gridDim.x * kNT;
  for (uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += st) d[i] = mul(d[i], f);
}

template <typename R>
__global__ __launch_bounds__(kNT) void k_scale_cplx(Cx<R> *__restrict__ d, uint64_t n, R fr, R fi)
{
  const uint64_t st = (uint64_t)gridDim.x * kNT;
  for (uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += st) {
    const Cx<R> x = d[i];
    d[i] = Cx<R>{sub(mul(x.re, fr), mul(x.im, fi)), add(mul(x.re, fi), mul(x.im, fr))};
  }
}

template <typename R, bool ACC>
__global__ __launch_bounds__(kNT) void k_axpy_real(R *__restrict__ d, const R *__restrict__ l, uint64_t n, R f)
{
  const uint64_t st = (uint64_t)gridDim.x * kNT;
  for (uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += st) {
    const R t = mul(l[i], f);
    d[i] = ACC ? add(d[i], t) : t;
  }
}

template <typename R, bool ACC, bool CONJ>
__global__ __launch_bounds__(kNT) void k_axpy_cplx(Cx<R> *__restrict__ d, const Cx<R> *__restrict__ l, uint64_t n,
                                                    R fr, R fi)
{
  const uint64_t st = (uint64_t)gridDim.x * kNT;
  for (uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += st) {
    Cx<R> x = l[i];
    if (CONJ) x.im = -x.im;
    const R tr = sub(mul(x.re, fr), mul(x.im, fi)), ti = add(mul(x.re, fi), mul(x.im, fr));
    if (ACC) {
      const Cx<R> y = d[i];
      d[i] = Cx<R>{add(y.re, tr), add(y.im, ti)};
    } else {
      d[i] = Cx<R>{tr, ti};
    }
  }
}

// Two-pass dot product: per-workgroup partial sums in double, then one
// workgroup reduces them (deterministic order, no atomics).
template <typename R, bool CPLX>
__global__ __launch_bounds__(kNT) void k_dot_partial(const R *__restrict__ l, const R *__restrict__ r, uint64_t n,
                                                      bool lconj, bool rconj, double2 *__restrict__ part)
{
  __shared__ double sr[kNT], si[kNT];
  double ar = 0, ai = 0;
  const uint64_t st = (uint64_t)gridDim.x * kNT;
  for (uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += st) {
    if (CPLX) {
      const double xr = l[2 * i], yr = r[2 * i];
      const double xi = lconj ? -(double)l[2 * i + 1] : (double)l[2 * i + 1];
      const double yi = rconj ? -(double)r[2 * i + 1] : (double)r[2 * i + 1];
      ar += xr * yr - xi * yi;
      ai += xr * yi + xi * yr;
    } else {
      ar += (double)l[i] * (double)r[i];
    }
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int w = kNT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sr[0], si[0]);
}

template <typename R, bool CPLX>
__global__ __launch_bounds__(kNT) void k_dot_final(const double2 *__restrict__ part, int np, R *__restrict__ d,
                                                    double are, double aim, bool acc)
{
  __shared__ double sr[kNT], si[kNT];
  double ar = 0, ai = 0;
  for (int i = threadIdx.x; i < np; i += kNT) {
    ar += part[i].x;
    ai += part[i].y;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int w = kNT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double vr = sr[0] * are - si[0] * aim, vi = sr[0] * aim + si[0] * are;
    if (CPLX) {
      d[0] = (R)((acc ? (double)d[0] : 0.0) + vr);
      d[1] = (R)((acc ? (double)d[1] : 0.0) + vi);
    } else {
      d[0] = (R)((acc ? (double)d[0] : 0.0) + vr);
    }
  }
}

// D (+)= (alpha * s) * (conj?) L with the scalar s read from device memory
// (the rank-0 operand of a contraction, TALSH/tensor_algebra_gpu_nvidia.cu:6755-6795).
template <typename R, bool CPLX>
__global__ __launch_bounds__(kNT) void k_axpy_dev(R *__restrict__ d, const R *__restrict__ l,
                                                   const R *__restrict__ sp, uint64_t n, double are, double aim,
                                                   bool lconj, bool sconj, bool acc)
{
  double fr, fi;
  if (CPLX) {
    const double sr = sp[0], si = sconj ? -(double)sp[1] : (double)sp[1];
    fr = sr * are - si * aim;
    fi = sr * aim + si * are;
  } else {
    // real kinds take the real approximant of the complex factor s*alpha
    const double pr = (double)sp[0] * are, pi = (double)sp[0] * aim;
    fr = pr != 0.0 ? sqrt(pr * pr + pi * pi) * (pr > 0 ? 1.0 : -1.0) : pi;
    fi = 0.0;
  }
  const R f0 = (R)fr, f1 = (R)fi;
  const uint64_t st = (uint64_t)gridDim.x * kNT;
  for (uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += st) {
    if (CPLX) {
      const R xr = l[2 * i], xi = lconj ? -l[2 * i + 1] : l[2 * i + 1];
      const R tr = sub(mul(xr, f0), mul(xi, f1)), ti = add(mul(xr, f1), mul(xi, f0));
      d[2 * i] = acc ? add(d[2 * i], tr) : tr;
      d[2 * i + 1] = acc ? add(d[2 * i + 1], ti) : ti;
    } else {
      const R t = mul(l[i], f0);
      d[i] = acc ? add(d[i], t) : t;
    }
  }
}

int grid_for(uint64_t n)
{
  static int cus = [] {
    int dev = 0, c = 256;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0)
      c = p.multiProcessorCount;
    return c;
  }();
  const uint64_t b = (n + kNT - 1) / kNT;
  const uint64_t cap = (uint64_t)cus * 16;
  return (int)(b < cap ? (b ? b : 1) : cap);
}

// Real approximant of a complex factor for real data kinds: modulus with the
// sign of the real part, or the imaginary part when the real part is zero
// (the CPU path's rule, TALSH/tensor_algebra_cpu.F90:430-445).
double real_factor(double re, double im)
{
  if (re != 0.0) return std::sqrt(re * re + im * im) * (re > 0 ? 1.0 : -1.0);
  return im;
}

// ---- box copy: tensor slice and insert ----------------------------------------
// out[dbase + sum_i x_i ds_i] (+)= in[sbase + sum_i x_i ss_i] over the box
// x < ext (dimension 0 fastest, its strides 1 on both sides after fusion).
// One launch covers fewer than 2^31 elements (32-bit element decode); the host
// cuts larger boxes into such pieces. Boxes of up to kBoxDims dimensions past
// the first use the compact parameter block, others (up to the maximum rank)
// the wide one.
constexpr int kBoxDims = 8;
constexpr int kBoxDimsWide = MAX_TENSOR_RANK;
struct BoxDiv { uint32_t d, m, s; };
template <int MAXD>
struct BoxParams {
  int nd;                       // box dimensions after the first (row decode)
  BoxDiv e0;                    // extent of dimension 0
  BoxDiv ext[MAXD];             // extents of dimensions 1..nd
  int64_t ss[MAXD], ds[MAXD];
  int64_t sbase, dbase;
  uint64_t vol;
};

__device__ __forceinline__ uint32_t bdiv(uint32_t n, const BoxDiv &f)
{
  return f.d == 1 ? n : ((__umulhi(n, f.m) + n) >> f.s);
}

template <typename R, int NC, int MAXD>
__global__ __launch_bounds__(kNT) void k_box(const R *__restrict__ in, R *__restrict__ out, const BoxParams<MAXD> p,
                                             int acc)
{
  const uint64_t st = (uint64_t)gridDim.x * kNT;
  for (uint64_t g = (uint64_t)blockIdx.x * kNT + threadIdx.x; g < p.vol; g += st) {
    // g < 2^31 for 32-bit decode (the launcher cuts larger boxes)
    uint32_t r = bdiv((uint32_t)g, p.e0);
    const uint32_t x0 = (uint32_t)g - r * p.e0.d;
    int64_t so = p.sbase + x0, dof = p.dbase + x0;
    for (int q = 0; q < p.nd; ++q) {
      const uint32_t rq = bdiv(r, p.ext[q]);
      const uint32_t c = r - rq * p.ext[q].d;
      r = rq;
      so += (int64_t)c * p.ss[q];
      dof += (int64_t)c * p.ds[q];
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const R v = in[so * NC + c];
      out[dof * NC + c] = acc ? add(out[dof * NC + c], v) : v;
    }
  }
}

template <int MAXD>
int launch_box_piece(const void *in, void *out, int kind, const std::vector<int64_t> &e,
                     const std::vector<int64_t> &a, const std::vector<int64_t> &b, int64_t sbase, int64_t dbase,
                     uint64_t vol, bool acc, hipStream_t s)
{
  BoxParams<MAXD> p{};
  auto mk = [](int64_t d) { FastDiv f((uint32_t)d); return BoxDiv{f.d, f.m, f.s}; };
  p.e0 = mk(e[0]);
  p.nd = (int)e.size() - 1;
  for (int q = 0; q < p.nd; ++q) {
    p.ext[q] = mk(e[q + 1]);
    p.ss[q] = a[q + 1];
    p.ds[q] = b[q + 1];
  }
  p.sbase = sbase;
  p.dbase = dbase;
  p.vol = vol;
  const int blocks = (int)std::min<uint64_t>((vol + kNT - 1) / kNT, 65536);
  const int ac = acc ? 1 : 0;
  switch (kind) {
  case R4: hipLaunchKernelGGL((k_box<float, 1, MAXD>), dim3(blocks), dim3(kNT), 0, s, (const float *)in, (float *)out, p, ac); break;
  case R8: hipLaunchKernelGGL((k_box<double, 1, MAXD>), dim3(blocks), dim3(kNT), 0, s, (const double *)in, (double *)out, p, ac); break;
  case C4: hipLaunchKernelGGL((k_box<float, 2, MAXD>), dim3(blocks), dim3(kNT), 0, s, (const float *)in, (float *)out, p, ac); break;
  case C8: hipLaunchKernelGGL((k_box<double, 2, MAXD>), dim3(blocks), dim3(kNT), 0, s, (const double *)in, (double *)out, p, ac); break;
  default: return TALSH_INVALID_ARGS;
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

// Launches the box (e, a, b) from the given bases, cut along its outermost
// dimension (or along the row when the box is a single row) until every
// piece has fewer than 2^31 elements.
int launch_box_cut(const void *in, void *out, int kind, std::vector<int64_t> e, std::vector<int64_t> a,
                   std::vector<int64_t> b, int64_t sbase, int64_t dbase, bool acc, hipStream_t s)
{
  constexpr uint64_t kMax = (1ull << 31) - 1;
  uint64_t vol = 1;
  for (int64_t x : e) vol *= (uint64_t)x;
  if (vol <= kMax) {
    if ((int)e.size() - 1 <= kBoxDims) return launch_box_piece<kBoxDims>(in, out, kind, e, a, b, sbase, dbase, vol, acc, s);
    return launch_box_piece<kBoxDimsWide>(in, out, kind, e, a, b, sbase, dbase, vol, acc, s);
  }
  if (e.size() == 1) {  // one contiguous row: pieces of 2^30 elements
    const int64_t step = 1ll << 30;
    for (int64_t x = 0; x < e[0]; x += step) {
      const int64_t n = std::min<int64_t>(step, e[0] - x);
      const int err = launch_box_cut(in, out, kind, {n}, {1}, {1}, sbase + x, dbase + x, acc, s);
      if (err) return err;
    }
    return 0;
  }
  const int64_t outer = e.back(), as = a.back(), bs = b.back();
  const uint64_t inner = vol / (uint64_t)outer;
  if (inner <= kMax) {  // several outer slabs per piece
    const int64_t step = (int64_t)std::max<uint64_t>(1, kMax / inner);
    for (int64_t x = 0; x < outer; x += step) {
      std::vector<int64_t> ee = e;
      ee.back() = std::min<int64_t>(step, outer - x);
      const int err = launch_box_cut(in, out, kind, ee, a, b, sbase + x * as, dbase + x * bs, acc, s);
      if (err) return err;
    }
    return 0;
  }
  e.pop_back(); a.pop_back(); b.pop_back();
  for (int64_t x = 0; x < outer; ++x) {
    const int err = launch_box_cut(in, out, kind, e, a, b, sbase + x * as, dbase + x * bs, acc, s);
    if (err) return err;
  }
  return 0;
}

}  // namespace

int launch_box(const void *in, void *out, int kind, int rank, const int *box, const int64_t *sstr,
               const int64_t *dstr, int64_t sbase, int64_t dbase, bool acc, hipStream_t s)
{
  if (!kind_size(kind) || rank < 0 || rank > MAX_TENSOR_RANK) return TALSH_INVALID_ARGS;
  // drop unit extents; fuse dimension i+1 into i when both sides are contiguous across them
  // (a fused extent may pass 32 bits only in the contiguous row, which the
  // launcher cuts; other extents are decoded in 32 bits)
  std::vector<int64_t> e, a, b;
  for (int i = 0; i < rank; ++i) {
    if (box[i] <= 0) return TALSH_INVALID_ARGS;
    if (box[i] == 1) continue;
    const bool row = e.size() == 1 && a[0] == 1 && b[0] == 1;
    if (!e.empty() && a.back() * e.back() == sstr[i] && b.back() * e.back() == dstr[i] &&
        (row || e.back() * box[i] <= (int64_t)INT32_MAX)) {
      e.back() *= box[i];
    } else {
      e.push_back(box[i]);
      a.push_back(sstr[i]);
      b.push_back(dstr[i]);
    }
  }
  if (e.empty()) { e.push_back(1); a.push_back(1); b.push_back(1); }
  if (a[0] != 1 || b[0] != 1) {  // innermost box dimension not contiguous: keep it as a row dimension
    e.insert(e.begin(), 1);
    a.insert(a.begin(), 1);
    b.insert(b.begin(), 1);
  }
  if ((int)e.size() - 1 > kBoxDimsWide) return TALSH_LIMIT_EXCEEDED;
  return launch_box_cut(in, out, kind, e, a, b, sbase, dbase, acc, s);
}

int launch_fill(void *d, int kind, size_t vol, double re, double im, hipStream_t s)
{
  if (!vol) return 0;
  const int g = grid_for(vol);
  switch (kind) {
  case R4: hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kNT), 0, s, (float *)d, (uint64_t)vol, (float)real_factor(re, im)); break;
  case R8: hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kNT), 0, s, (double *)d, (uint64_t)vol, real_factor(re, im)); break;
  case C4: hipLaunchKernelGGL(k_fill<float2>, dim3(g), dim3(kNT), 0, s, (float2 *)d, (uint64_t)vol, make_float2((float)re, (float)im)); break;
  case C8: hipLaunchKernelGGL(k_fill<double2>, dim3(g), dim3(kNT), 0, s, (double2 *)d, (uint64_t)vol, make_double2(re, im)); break;
  default: return TALSH_INVALID_ARGS;
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

int launch_scale(void *d, int kind, size_t vol, double re, double im, hipStream_t s)
{
  if (!vol) return 0;
  const int g = grid_for(vol);
  switch (kind) {
  case R4: hipLaunchKernelGGL(k_scale_real<float>, dim3(g), dim3(kNT), 0, s, (float *)d, (uint64_t)vol, (float)real_factor(re, im)); break;
  case R8: hipLaunchKernelGGL(k_scale_real<double>, dim3(g), dim3(kNT), 0, s, (double *)d, (uint64_t)vol, real_factor(re, im)); break;
  case C4: hipLaunchKernelGGL(k_scale_cplx<float>, dim3(g), dim3(kNT), 0, s, (Cx<float> *)d, (uint64_t)vol, (float)re, (float)im); break;
  case C8: hipLaunchKernelGGL(k_scale_cplx<double>, dim3(g), dim3(kNT), 0, s, (Cx<double> *)d, (uint64_t)vol, re, im); break;
  default: return TALSH_INVALID_ARGS;
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

int launch_axpy(void *d, const void *l, int kind, size_t vol, double re, double im, bool conj, bool acc, hipStream_t s)
{
  if (!vol) return 0;
  const int g = grid_for(vol);
  const uint64_t n = vol;
  switch (kind) {
  case R4: {
    const float f = (float)real_factor(re, im);
    if (acc) hipLaunchKernelGGL((k_axpy_real<float, true>), dim3(g), dim3(kNT), 0, s, (float *)d, (const float *)l, n, f);
    else hipLaunchKernelGGL((k_axpy_real<float, false>), dim3(g), dim3(kNT), 0, s, (float *)d, (const float *)l, n, f);
    break;
  }
  case R8: {
    const double f = real_factor(re, im);
    if (acc) hipLaunchKernelGGL((k_axpy_real<double, true>), dim3(g), dim3(kNT), 0, s, (double *)d, (const double *)l, n, f);
    else hipLaunchKernelGGL((k_axpy_real<double, false>), dim3(g), dim3(kNT), 0, s, (double *)d, (const double *)l, n, f);
    break;
  }
  case C4: {
    auto *dd = (Cx<float> *)d; auto *ll = (const Cx<float> *)l;
    const float fr = (float)re, fi = (float)im;
    if (acc && conj) hipLaunchKernelGGL((k_axpy_cplx<float, true, true>), dim3(g), dim3(kNT), 0, s, dd, ll, n, fr, fi);
    else if (acc) hipLaunchKernelGGL((k_axpy_cplx<float, true, false>), dim3(g), dim3(kNT), 0, s, dd, ll, n, fr, fi);
    else if (conj) hipLaunchKernelGGL((k_axpy_cplx<float, false, true>), dim3(g), dim3(kNT), 0, s, dd, ll, n, fr, fi);
    else hipLaunchKernelGGL((k_axpy_cplx<float, false, false>), dim3(g), dim3(kNT), 0, s, dd, ll, n, fr, fi);
    break;
  }
  case C8: {
    auto *dd = (Cx<double> *)d; auto *ll = (const Cx<double> *)l;
    if (acc && conj) hipLaunchKernelGGL((k_axpy_cplx<double, true, true>), dim3(g), dim3(kNT), 0, s, dd, ll, n, re, im);
    else if (acc) hipLaunchKernelGGL((k_axpy_cplx<double, true, false>), dim3(g), dim3(kNT), 0, s, dd, ll, n, re, im);
    else if (conj) hipLaunchKernelGGL((k_axpy_cplx<double, false, true>), dim3(g), dim3(kNT), 0, s, dd, ll, n, re, im);
    else hipLaunchKernelGGL((k_axpy_cplx<double, false, false>), dim3(g), dim3(kNT), 0, s, dd, ll, n, re, im);
    break;
  }
  default: return TALSH_INVALID_ARGS;
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

int launch_axpy_dev(void *d, const void *l, const void *sp, int kind, size_t vol, double are, double aim, bool lconj,
                    bool sconj, bool acc, hipStream_t s)
{
  if (!vol) return 0;
  const int g = grid_for(vol);
  const uint64_t n = vol;
  switch (kind) {
  case R4: hipLaunchKernelGGL((k_axpy_dev<float, false>), dim3(g), dim3(kNT), 0, s, (float *)d, (const float *)l, (const float *)sp, n, are, aim, lconj, sconj, acc); break;
  case R8: hipLaunchKernelGGL((k_axpy_dev<double, false>), dim3(g), dim3(kNT), 0, s, (double *)d, (const double *)l, (const double *)sp, n, are, aim, lconj, sconj, acc); break;
  case C4: hipLaunchKernelGGL((k_axpy_dev<float, true>), dim3(g), dim3(kNT), 0, s, (float *)d, (const float *)l, (const float *)sp, n, are, aim, lconj, sconj, acc); break;
  case C8: hipLaunchKernelGGL((k_axpy_dev<double, true>), dim3(g), dim3(kNT), 0, s, (double *)d, (const double *)l, (const double *)sp, n, are, aim, lconj, sconj, acc); break;
  default: return TALSH_INVALID_ARGS;
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

int launch_dot(void *d, const void *l, const void *r, int kind, size_t n, double are, double aim, bool lconj,
               bool rconj, bool acc, void *scratch, size_t scratch_bytes, hipStream_t s)
{
  const int np = std::min(grid_for(n), 1024);
  if (scratch_bytes < (size_t)np * sizeof(double2)) return TALSH_INVALID_ARGS;
  double2 *part = (double2 *)scratch;
  const uint64_t nn = n;
  switch (kind) {
  case R4:
    hipLaunchKernelGGL((k_dot_partial<float, false>), dim3(np), dim3(kNT), 0, s, (const float *)l, (const float *)r, nn, false, false, part);
    hipLaunchKernelGGL((k_dot_final<float, false>), dim3(1), dim3(kNT), 0, s, part, np, (float *)d, are, 0.0, acc);
    break;
  case R8:
    hipLaunchKernelGGL((k_dot_partial<double, false>), dim3(np), dim3(kNT), 0, s, (const double *)l, (const double *)r, nn, false, false, part);
    hipLaunchKernelGGL((k_dot_final<double, false>), dim3(1), dim3(kNT), 0, s, part, np, (double *)d, are, 0.0, acc);
    break;
  case C4:
    hipLaunchKernelGGL((k_dot_partial<float, true>), dim3(np), dim3(kNT), 0, s, (const float *)l, (const float *)r, nn, lconj, rconj, part);
    hipLaunchKernelGGL((k_dot_final<float, true>), dim3(1), dim3(kNT), 0, s, part, np, (float *)d, are, aim, acc);
    break;
  case C8:
    hipLaunchKernelGGL((k_dot_partial<double, true>), dim3(np), dim3(kNT), 0, s, (const double *)l, (const double *)r, nn, lconj, rconj, part);
    hipLaunchKernelGGL((k_dot_final<double, true>), dim3(1), dim3(kNT), 0, s, part, np, (double *)d, are, aim, acc);
    break;
  default: return TALSH_INVALID_ARGS;
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

}  // namespace xt
