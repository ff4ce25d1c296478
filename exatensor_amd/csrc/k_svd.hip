// Singular value decomposition of a matricized tensor block on the GPU:
// one-sided (Hestenes) Jacobi in FP64 (complex: FP64 pairs), for
// talshTensorDecomposeSVD. The reference computes it on the host with LAPACK
// ?gesvdx (TALSH/tensor_algebra_cpu.F90:4360-4914); its GPU entry is a stub.
//
// B (m x n, column-major, m >= n) is orthogonalised column pair by column
// pair: for columns p, q with a = |b_p|^2, b = |b_q|^2, g = b_p^H b_q =
// |g| e^{i phi}, the rotation
//   b_p' = c b_p - s e^{-i phi} b_q,   b_q' = s b_p + c e^{-i phi} b_q,
//   zeta = (b - a) / (2 |g|),  t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
//   c = 1 / sqrt(1 + t^2),     s = c t
// makes them orthogonal (t is the smaller root of t^2 + 2 zeta t - 1 = 0);
// the same rotations accumulate in V (n x n, from the identity), so that at
// convergence B = A V has orthogonal columns: A = U diag(sigma) V^H with
// sigma_j = |b_j| and u_j = b_j / sigma_j. A sweep visits every pair once in
// n_e - 1 rounds of n_e / 2 disjoint pairs (round-robin tournament, n_e = n
// rounded up to even; the extra index is a bye); one workgroup per pair.
#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace xt {
namespace {

constexpr int kSvdThreads = 256;

template <bool CPLX>
struct Z;
template <>
struct Z<false> {
  typedef double T;
  __device__ static double nrm2(T x) { return x * x; }
  __device__ static T cdot(T a, T b) { return a * b; }  // conj(a) * b
};
template <>
struct Z<true> {
  typedef double2 T;
  __device__ static double nrm2(T x) { return x.x * x.x + x.y * x.y; }
  __device__ static T cdot(T a, T b) { return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
};

__device__ __forceinline__ void pair_of(int ne, int r, int k, int &p, int &q)
{
  const int n1 = ne - 1;
  if (k == 0) { p = n1; q = r; }
  else { p = (r + k) % n1; q = (r - k + n1) % n1; }
}

template <bool CPLX>
__global__ __launch_bounds__(kSvdThreads) void k_jacobi_round(typename Z<CPLX>::T *__restrict__ B,
                                                               typename Z<CPLX>::T *__restrict__ V, int64_t m, int n,
                                                               int ne, int round, double tol, double *__restrict__ off)
{
  typedef typename Z<CPLX>::T T;
  __shared__ double red[3][kSvdThreads];
  __shared__ double rot[4];  // c, s, cos(phi), sin(phi)
  int p, q;
  pair_of(ne, round, blockIdx.x, p, q);
  if (p >= n || q >= n) return;  // a bye
  if (p > q) { const int t = p; p = q; q = t; }
  T *bp = B + (int64_t)p * m, *bq = B + (int64_t)q * m;
  double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += kSvdThreads) {
    const T x = bp[i], y = bq[i];
    a += Z<CPLX>::nrm2(x);
    b += Z<CPLX>::nrm2(y);
    if constexpr (CPLX) {
      const T g = Z<CPLX>::cdot(x, y);
      gr += g.x;
      gi += g.y;
    } else {
      gr += x * y;
    }
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  red[2][threadIdx.x] = gr;
  __syncthreads();
  for (int w = kSvdThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + w];
    __syncthreads();
  }
  a = red[0][0];
  b = red[1][0];
  gr = red[2][0];
  __syncthreads();
  if constexpr (CPLX) {
    red[0][threadIdx.x] = gi;
    __syncthreads();
    for (int w = kSvdThreads / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[0][threadIdx.x] += red[0][threadIdx.x + w];
      __syncthreads();
    }
    gi = red[0][0];
  }
  if (threadIdx.x == 0) {
    const double ag = sqrt(gr * gr + gi * gi);
    const double scale = sqrt(a * b);
    double c = 1.0, s = 0.0, cp = 1.0, sp = 0.0;
    if (scale > 0.0 && ag > tol * scale) {
      const double zeta = (b - a) / (2.0 * ag);
      const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      c = 1.0 / sqrt(1.0 + t * t);
      s = c * t;
      cp = gr / ag;
      sp = gi / ag;
      off[blockIdx.x] = fmax(off[blockIdx.x], ag / scale);
    }
    rot[0] = c; rot[1] = s; rot[2] = cp; rot[3] = sp;
  }
  __syncthreads();
  const double c = rot[0], s = rot[1], cp = rot[2], sp = rot[3];
  if (s == 0.0) return;
  // y~ = e^{-i phi} y
  auto apply = [&](T *xp, T *yp, int64_t len) {
    for (int64_t i = threadIdx.x; i < len; i += kSvdThreads) {
      const T x = xp[i], y = yp[i];
      if constexpr (CPLX) {
        const double2 yt = make_double2(cp * y.x + sp * y.y, cp * y.y - sp * y.x);
        xp[i] = make_double2(c * x.x - s * yt.x, c * x.y - s * yt.y);
        yp[i] = make_double2(s * x.x + c * yt.x, s * x.y + c * yt.y);
      } else {
        const double yt = cp * y;
        xp[i] = c * x - s * yt;
        yp[i] = s * x + c * yt;
      }
    }
  };
  apply(bp, bq, m);
  apply(V + (int64_t)p * n, V + (int64_t)q * n, n);
}

template <bool CPLX>
__global__ __launch_bounds__(kSvdThreads) void k_colnorm(const typename Z<CPLX>::T *__restrict__ B, int64_t m,
                                                          double *__restrict__ sigma)
{
  __shared__ double red[kSvdThreads];
  const typename Z<CPLX>::T *b = B + (int64_t)blockIdx.x * m;
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += kSvdThreads) a += Z<CPLX>::nrm2(b[i]);
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = kSvdThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sigma[blockIdx.x] = sqrt(red[0]);
}

template <typename T>
__device__ __forceinline__ void put(T *p, double re, double im);
template <>
__device__ __forceinline__ void put<float>(float *p, double re, double) { *p = (float)re; }
template <>
__device__ __forceinline__ void put<double>(double *p, double re, double) { *p = re; }
template <>
__device__ __forceinline__ void put<float2>(float2 *p, double re, double im) { *p = make_float2((float)re, (float)im); }
template <>
__device__ __forceinline__ void put<double2>(double2 *p, double re, double im) { *p = make_double2(re, im); }

template <typename T>
__device__ __forceinline__ void get(const T *p, double &re, double &im);
template <>
__device__ __forceinline__ void get<float>(const float *p, double &re, double &im) { re = *p; im = 0.0; }
template <>
__device__ __forceinline__ void get<double>(const double *p, double &re, double &im) { re = *p; im = 0.0; }
template <>
__device__ __forceinline__ void get<float2>(const float2 *p, double &re, double &im) { re = p->x; im = p->y; }
template <>
__device__ __forceinline__ void get<double2>(const double2 *p, double &re, double &im) { re = p->x; im = p->y; }

// B (rows x cols, FP64) from the tensor-kind matrix A (lu x ru, column-major):
// B = A when not herm, B = A^H (ru x lu) when herm.
template <typename T, bool CPLX>
__global__ void k_load(const T *__restrict__ A, int64_t lu, int64_t ru, bool herm, typename Z<CPLX>::T *__restrict__ B)
{
  const int64_t n = lu * ru, st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += st) {
    const int64_t i = e % lu, j = e / lu;  // A(i, j)
    double re, im;
    get<T>(A + e, re, im);
    const int64_t o = herm ? (j + i * ru) : e;
    if constexpr (CPLX) B[o] = make_double2(re, herm ? -im : im);
    else B[o] = re;
  }
}

// Factors in the tensor kind: Lmat (lu x nv) = U(:, order) * fl(sigma),
// Rmat (nv x ru) = V(:, order)^H * fr(sigma), S(nv) = sigma(order) (zero past
// the available values). U / V come from the Jacobi factors: without the
// transposition U = B / sigma (lu x n) and V = Vj (ru x n); with it (B was
// A^H) U = Vj (lu x n) and V = B / sigma (ru x n).
struct EmitArgs {
  int64_t lu, ru, nv, n;
  bool herm;
  int absorb;  // 0 none, 1 left, 2 right, 3 square root into both
};

template <typename T, bool CPLX>
__global__ void k_emit(const typename Z<CPLX>::T *__restrict__ B, const typename Z<CPLX>::T *__restrict__ Vj,
                       const double *__restrict__ sigma, const int *__restrict__ order, const EmitArgs a,
                       T *__restrict__ Lm, T *__restrict__ Rm, T *__restrict__ S)
{
  const int64_t nl = a.lu * a.nv, nr = a.nv * a.ru, st = (int64_t)gridDim.x * blockDim.x;
  auto fac = [&](double sg, bool left) {
    if (a.absorb == 3) return sqrt(sg);
    if ((a.absorb == 1 && left) || (a.absorb == 2 && !left)) return sg;
    return 1.0;
  };
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nl + nr + a.nv; e += st) {
    if (e < nl) {  // L(i, k) = u_{order k}(i)
      const int64_t i = e % a.lu, k = e / a.lu;
      double re = 0.0, im = 0.0;
      if (k < a.n) {
        const int j = order[k];
        const double sg = sigma[j];
        const typename Z<CPLX>::T u = a.herm ? Vj[i + (int64_t)j * a.lu] : B[i + (int64_t)j * a.lu];
        const double inv = a.herm ? 1.0 : (sg > 0.0 ? 1.0 / sg : 0.0);
        const double f = fac(sg, true) * inv;
        if constexpr (CPLX) { re = u.x * f; im = u.y * f; } else { re = u * f; }
      }
      put<T>(Lm + e, re, im);
    } else if (e < nl + nr) {  // R(k, j) = conj(v_{order k}(j))
      const int64_t f0 = e - nl, k = f0 % a.nv, jj = f0 / a.nv;
      double re = 0.0, im = 0.0;
      if (k < a.n) {
        const int j = order[k];
        const double sg = sigma[j];
        const typename Z<CPLX>::T v = a.herm ? B[jj + (int64_t)j * a.ru] : Vj[jj + (int64_t)j * a.ru];
        const double inv = a.herm ? (sg > 0.0 ? 1.0 / sg : 0.0) : 1.0;
        const double f = fac(sg, false) * inv;
        if constexpr (CPLX) { re = v.x * f; im = -v.y * f; } else { re = v * f; }
      }
      put<T>(Rm + f0, re, im);
    } else {
      const int64_t k = e - nl - nr;
      put<T>(S + k, k < a.n ? sigma[order[k]] : 0.0, 0.0);
    }
  }
}

__global__ void k_eye(double *__restrict__ V, int n, int words)
{
  const int64_t tot = (int64_t)n * n * words, st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += st) {
    const int64_t el = e / words, w = e % words;
    V[e] = (w == 0 && el % n == el / n) ? 1.0 : 0.0;
  }
}

__global__ void k_zero(double *__restrict__ p, int64_t n)
{
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += st) p[e] = 0.0;
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 4096); }

template <typename T, bool CPLX>
int svd_t(const SvdArgs &g, hipStream_t s)
{
  typedef typename Z<CPLX>::T W;
  constexpr int words = CPLX ? 2 : 1;
  const bool herm = g.lu < g.ru;
  const int64_t m = herm ? g.ru : g.lu;
  const int n = (int)(herm ? g.lu : g.ru);
  const int ne = n + (n & 1);
  W *B = static_cast<W *>(g.work);
  W *V = B + m * n;
  double *off = reinterpret_cast<double *>(V + (int64_t)n * n);
  double *sigma = off + ne / 2;
  int *order = reinterpret_cast<int *>(sigma + n);
  hipLaunchKernelGGL((k_load<T, CPLX>), dim3(grid_for(g.lu * g.ru)), dim3(256), 0, s, static_cast<const T *>(g.a),
                     g.lu, g.ru, herm, B);
  hipLaunchKernelGGL(k_eye, dim3(grid_for((int64_t)n * n * words)), dim3(256), 0, s, reinterpret_cast<double *>(V), n,
                     words);
  const double tol = 1e-15;
  std::vector<double> offh(ne / 2);
  int sweep = 0;
  for (; sweep < g.max_sweeps && n > 1; ++sweep) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, s, off, (int64_t)(ne / 2));
    for (int r = 0; r < ne - 1; ++r)
      hipLaunchKernelGGL((k_jacobi_round<CPLX>), dim3(ne / 2), dim3(kSvdThreads), 0, s, B, V, m, n, ne, r, tol, off);
    if (!XT_HIP_OK(hipMemcpyAsync(offh.data(), off, sizeof(double) * (ne / 2), hipMemcpyDeviceToHost, s)) ||
        !XT_HIP_OK(hipStreamSynchronize(s)))
      return TALSH_FAILURE;
    double mx = 0.0;
    for (double v : offh) mx = std::max(mx, v);
    if (mx <= 1e-14) break;
  }
  if (n > 1 && sweep == g.max_sweeps) return NUM_INSTABILITY;
  hipLaunchKernelGGL((k_colnorm<CPLX>), dim3(n), dim3(kSvdThreads), 0, s, B, m, sigma);
  std::vector<double> sh(n);
  if (!XT_HIP_OK(hipMemcpyAsync(sh.data(), sigma, sizeof(double) * n, hipMemcpyDeviceToHost, s)) ||
      !XT_HIP_OK(hipStreamSynchronize(s)))
    return TALSH_FAILURE;
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return sh[x] > sh[y]; });
  if (g.sigma_out)
    for (int i = 0; i < n; ++i) g.sigma_out[i] = sh[ord[i]];
  if (!XT_HIP_OK(hipMemcpyAsync(order, ord.data(), sizeof(int) * n, hipMemcpyHostToDevice, s)) ||
      !XT_HIP_OK(hipStreamSynchronize(s)))
    return TALSH_FAILURE;
  EmitArgs a{g.lu, g.ru, g.nv, (int64_t)n, herm, g.absorb};
  hipLaunchKernelGGL((k_emit<T, CPLX>), dim3(grid_for(g.lu * g.nv + g.nv * g.ru + g.nv)), dim3(256), 0, s, B, V,
                     sigma, order, a, static_cast<T *>(g.lmat), static_cast<T *>(g.rmat), static_cast<T *>(g.svals));
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

}  // namespace

size_t svd_workspace_bytes(int kind, int64_t lu, int64_t ru)
{
  const int64_t m = std::max(lu, ru), n = std::min(lu, ru), ne = n + (n & 1);
  const int64_t w = kind_complex(kind) ? 16 : 8;
  return (size_t)(w * (m * n + n * n) + 8 * (ne / 2 + n) + 4 * n + 64);
}

int launch_svd(const SvdArgs &g, hipStream_t s)
{
  if (g.lu <= 0 || g.ru <= 0 || g.nv <= 0 || g.nv > std::min(g.lu, g.ru)) return TALSH_INVALID_ARGS;
  if (std::min(g.lu, g.ru) > (int64_t)1 << 30) return TALSH_LIMIT_EXCEEDED;
  switch (g.kind) {
  case R4: return svd_t<float, false>(g, s);
  case R8: return svd_t<double, false>(g, s);
  case C4: return svd_t<float2, true>(g, s);
  case C8: return svd_t<double2, true>(g, s);
  default: return TALSH_INVALID_ARGS;
  }
}

}  // namespace xt
