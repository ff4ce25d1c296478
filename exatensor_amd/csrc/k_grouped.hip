// Persistent grouped contraction kernel for batches of uneven block
// contractions (block-sparse tensors), gfx950.
//
// A batch holds thousands of block contractions D_i += alpha * L_i * R_i of
// one pattern. Each is a GEMM whose row (m), column (n) and contracted (k)
// indices are composites of up to four tensor dimensions; operands are read
// in place from their DLF bodies through per-block index maps (fused gather:
// no per-block transpose launches). Work items are 64x64 output tiles of a D
// block together with the chain of (L, R) pairs that accumulate into it, so
// every D element is produced by exactly one workgroup in a fixed order (no
// atomics on D, deterministic results). Workgroups pull items from a device
// work queue sorted by decreasing cost (longest first) until it is empty.
//
// Replaces the per-block talsh_tensor_contract calls issued by ExaTENSOR's
// block-level caller (INTERVIRT/tavp_worker.F90:9440-9480).
#include "grouped.hpp"

namespace xt {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct GAcc;
template <>
struct GAcc<double> { typedef d4 type; };
template <>
struct GAcc<float> { typedef f4 type; };
template <typename T>
struct GF4;
template <>
struct GF4<double> { typedef double4 type; };
template <>
struct GF4<float> { typedef float4 type; };

__device__ __forceinline__ d4 gmfma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 gmfma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

__device__ __forceinline__ uint32_t gdiv(uint32_t n, const GDiv &f)
{
  return (f.d == 1) ? n : ((__umulhi(n, f.m) + n) >> f.s);
}

// Offsets of composite index v in the two tensors that carry it.
__device__ __forceinline__ void map_offsets(const IdxMap &mp, uint32_t v, int64_t &o0, int64_t &o1)
{
  o0 = 0;
  o1 = 0;
  for (int q = 0; q < mp.n; ++q) {
    const uint32_t qq = gdiv(v, mp.div[q]);
    const uint32_t c = v - qq * mp.div[q].d;
    v = qq;
    o0 += (int64_t)c * mp.s0[q];
    o1 += (int64_t)c * mp.s1[q];
  }
}

__global__ void k_zero_counter(int *c) { *c = 0; }

template <typename T>
__global__ __launch_bounds__(256, 2) void k_grouped(const BlockOp *__restrict__ ops, const WorkItem *__restrict__ items,
                                                     int nitems, int *__restrict__ counter, double alpha, int accumulate)
{
  constexpr int BM = 64, BN = 64, BK = 16, KP = BK + (sizeof(T) == 4 ? 4 : 2);
  typedef typename GAcc<T>::type AccT;
  typedef typename GF4<T>::type F4;
  __shared__ __align__(16) T As[2][BM][KP];
  __shared__ __align__(16) T Bs[2][BN][KP];
  __shared__ int next_item;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int sr = tid & 63, sk = (tid >> 6) * 4;  // staging: row sr, k = sk..sk+3
  const int fr = lane & 15, fg = (lane >> 4) * 4, fk = lane >> 4;
  while (true) {
    if (tid == 0) next_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = next_item;
    __syncthreads();
    if (it >= nitems) break;
    const WorkItem w = items[it];
    const int64_t m0 = (int64_t)w.mt * BM, n0 = (int64_t)w.nt * BN;
    AccT acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = AccT{0, 0, 0, 0};
    for (int oi = w.op_begin; oi < w.op_end; ++oi) {
      const BlockOp &op = ops[oi];
      const T *L = static_cast<const T *>(op.l);
      const T *R = static_cast<const T *>(op.r);
      // this thread's staging row in A (m) and B (n): offsets into L and R
      const int64_t am = m0 + sr, bn = n0 + sr;
      int64_t offLm = 0, offRn = 0, dummy;
      const bool mok = am < op.m, nok = bn < op.n;
      if (mok) map_offsets(op.M, (uint32_t)am, offLm, dummy);
      if (nok) map_offsets(op.N, (uint32_t)bn, offRn, dummy);
      const int64_t nk = (op.k + BK - 1) / BK;
      T ra[4], rb[4];
      auto load = [&](int64_t k0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t kk = k0 + sk + j;
          int64_t offLk = 0, offRk = 0;
          const bool kok = kk < op.k;
          if (kok) map_offsets(op.K, (uint32_t)kk, offLk, offRk);
          ra[j] = (mok && kok) ? L[offLm + offLk] : T(0);
          rb[j] = (nok && kok) ? R[offRn + offRk] : T(0);
        }
      };
      auto store = [&](int buf) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          As[buf][sr][sk + j] = ra[j];
          Bs[buf][sr][sk + j] = rb[j];
        }
      };
      load(0);
      store(0);
      __syncthreads();
      for (int64_t kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) load((kt + 1) * BK);
        F4 fa[2], fb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) fa[i] = *reinterpret_cast<const F4 *>(&As[cur][wm + i * 16 + fr][fg]);
#pragma unroll
        for (int j = 0; j < 2; ++j) fb[j] = *reinterpret_cast<const F4 *>(&Bs[cur][wn + j * 16 + fr][fg]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc[i][j] = gmfma(fa[i].x, fb[j].x, acc[i][j]);
            acc[i][j] = gmfma(fa[i].y, fb[j].y, acc[i][j]);
            acc[i][j] = gmfma(fa[i].z, fb[j].z, acc[i][j]);
            acc[i][j] = gmfma(fa[i].w, fb[j].w, acc[i][j]);
          }
        if (kt + 1 < nk) store(cur ^ 1);
        __syncthreads();
      }
    }
    // write the tile of D (strides of the chain's D block from its first op)
    const BlockOp &op0 = ops[w.op_begin];
    T *D = static_cast<T *>(w.d);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t col = n0 + wn + j * 16 + fr;
        if (col >= op0.n) continue;
        int64_t dn, d0;
        map_offsets(op0.N, (uint32_t)col, d0, dn);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = m0 + wm + i * 16 + (sizeof(T) == 8 ? fk + 4 * r : fk * 4 + r);
          if (row >= op0.m) continue;
          int64_t dm, l0;
          map_offsets(op0.M, (uint32_t)row, l0, dm);
          T v = (T)alpha * acc[i][j][r];
          if (accumulate) v += D[dm + dn];
          D[dm + dn] = v;
        }
      }
  }
}

}  // namespace

int launch_grouped(int kind, const BlockOp *ops, const WorkItem *items, int nitems, int *counter, double alpha,
                   bool accumulate, int grid, hipStream_t s)
{
  if (nitems <= 0) return 0;
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, counter);
  if (kind == R8)
    hipLaunchKernelGGL(k_grouped<double>, dim3(grid), dim3(256), 0, s, ops, items, nitems, counter, alpha,
                       accumulate ? 1 : 0);
  else if (kind == R4)
    hipLaunchKernelGGL(k_grouped<float>, dim3(grid), dim3(256), 0, s, ops, items, nitems, counter, alpha,
                       accumulate ? 1 : 0);
  else
    return TALSH_INVALID_ARGS;
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

}  // namespace xt
