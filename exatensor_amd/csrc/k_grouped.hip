// Batched block-sparse contraction kernels for gfx950: a pack gather and a
// persistent grouped GEMM over a device work queue.
//
// A batch holds thousands of block contractions D_i += alpha * L_i * R_i of
// one pattern. Each is a GEMM whose row (m), column (n) and contracted (k)
// indices are composites of up to four tensor dimensions.
//
//  1. k_pack gathers every distinct L and R block once into a packed matrix
//     with k fastest and k padded to a multiple of 16 with zeros, so that the
//     GEMM reads 16-byte vectors along k with no bounds checks on k.
//  2. k_grouped_gemm is persistent: each workgroup pulls work items (a
//     128x128 tile of one D block plus the chain of (L, R) pairs that
//     accumulate into it) from per-XCD device queues until all are empty.
//     The inner loop is the dense engine GEMM's: double-buffered k-fast LDS
//     tiles feeding 16x16x4 MFMAs (FP64 for R8, FP32 for R4). The epilogue
//     scatters the tile into the D block through its index maps. Every D
//     element is written by one workgroup in chain order: results are
//     deterministic and need no atomics.
//
// Complex blocks (C4/C8) are packed realified (see PackJob) and run through
// the same real kernels.
//
// Replaces the per-block talsh_tensor_contract calls issued by ExaTENSOR's
// block-level caller (INTERVIRT/tavp_worker.F90:9440-9480).
#include <type_traits>

#include "bf16x3.hpp"
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

// Offset of composite index v in the tensor the map describes.
__device__ __forceinline__ int64_t map_off(const IdxMap &mp, uint32_t v)
{
  int64_t o = 0;
  for (int q = 0; q < mp.n; ++q) {
    const uint32_t qq = gdiv(v, mp.div[q]);
    o += (int64_t)(v - qq * mp.div[q].d) * mp.s[q];
    v = qq;
  }
  return o;
}

// 32-bit variant for D blocks (volume below 2^31, checked when the batch is built).
__device__ __forceinline__ int32_t map_off32(const IdxMap &mp, uint32_t v)
{
  int32_t o = 0;
  for (int q = 0; q < mp.n; ++q) {
    const uint32_t qq = gdiv(v, mp.div[q]);
    o += (int32_t)(v - qq * mp.div[q].d) * (int32_t)mp.s[q];
    v = qq;
  }
  return o;
}

__global__ void k_zero_counters(WorkQueues *q)
{
  for (int i = 0; i < kQueues; ++i) q->counter[i] = 0;
}

// The next work item for this workgroup (called by one thread): its XCD's
// queue first, then the others in turn; -1 when every queue is drained.
__device__ __forceinline__ int take_item(WorkQueues *q)
{
  const int home = (int)(blockIdx.x % kQueues);
  for (int t = 0; t < kQueues; ++t) {
    const int qq = (home + t) % kQueues;
    const int len = q->end[qq] - q->begin[qq];
    if (__hip_atomic_load(&q->counter[qq], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len) continue;
    const int v = atomicAdd(&q->counter[qq], 1);
    if (v < len) return q->begin[qq] + v;
  }
  return -1;
}

// D blocks are reached through generic pointers stored in the work items;
// accessed through a global-address-space pointer they compile to global
// (not flat) loads and stores, which count on the vector-memory counter only.
template <typename T>
using GlobalPtr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ GlobalPtr<T> as_global(void *p)
{
  return (GlobalPtr<T>)(static_cast<T *>(p));
}


// Three-way split operands are packed tile-blocked (B3Layout<16, 32>): row
// band b (16 rows) and k-block q (32 values) form the 1 KB block at
// b * 16 * kp + q * 512 of each plane; a plane holds round_up(rows, 16) * kp
// values (the padding rows are never stored: they feed only product rows or
// columns that are not written back).
typedef B3Layout<16, 32> PackLay;
__device__ __forceinline__ size_t b3_plane(int64_t nrows, int64_t kp) { return (size_t)((nrows + 15) / 16 * 16 * kp); }
__device__ __forceinline__ size_t b3_packed_offset(uint32_t r, uint32_t kk, int64_t kp)
{
  const uint32_t rr = r & 15u, q = kk >> 5, c = (kk >> 3) & 3u;
  return (size_t)(r >> 4) * 16 * kp + (size_t)q * 512 + rr * 32 + PackLay::pos((int)rr, (int)c) * 8 + (kk & 7u);
}

// Packs chunks of kPackChunk destination elements; a chunk belongs to the job
// with the largest chunk0 not above it (binary search over the sorted jobs).
template <typename T, bool SPLIT>
__global__ __launch_bounds__(256) void k_pack(const PackJob *__restrict__ jobs, int njobs, int64_t nchunks)
{
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].chunk0 <= c) lo = mid;
      else hi = mid - 1;
    }
    const PackJob &J = jobs[lo];
    const T *src = static_cast<const T *>(J.src);
    T *dst = static_cast<T *>(J.dst);
    const uint32_t total = (uint32_t)(J.nrows * J.kp), kp = (uint32_t)J.kp, k = (uint32_t)J.k;
    const uint32_t base = (uint32_t)((c - J.chunk0) * kPackChunk);
    const T isign = J.conj ? T(-1) : T(1);
    for (int e = threadIdx.x; e < kPackChunk; e += 256) {
      const uint32_t idx = base + (uint32_t)e;
      if (idx >= total) break;
      const uint32_t r = idx / kp, kk = idx - r * kp;
      T v = T(0);
      if (!J.cplx) {
        if (kk < k) v = src[map_off(J.rows, r) + map_off(J.cols, kk)];
      } else if (kk < 2 * k) {
        const uint32_t half = kk >= k ? 1u : 0u, kc = kk - half * k;
        const uint32_t rc = J.cplx == 1 ? r >> 1 : r, comp = J.cplx == 1 ? (r & 1u) : 0u;
        const int64_t o = 2 * (map_off(J.rows, rc) + map_off(J.cols, kc));
        const T re = src[o], im = isign * src[o + 1];
        // (comp, half): (0,0) re, (0,1) -im (L) or im (R), (1,0) im, (1,1) re
        v = comp == half ? re : (J.cplx == 1 && half) ? -im : im;
      }
      if constexpr (SPLIT) {
        unsigned short h0, h1, h2;
        split3(v, h0, h1, h2);
        unsigned short *d16 = static_cast<unsigned short *>(J.dst);
        const size_t o = b3_packed_offset(r, kk, kp), plane = b3_plane(J.nrows, kp);
        d16[o] = h0;
        d16[plane + o] = h1;
        d16[2 * plane + o] = h2;
      } else {
        dst[idx] = v;
      }
    }
  }
}

// Packs real blocks in tiles of 64 packed rows x 64 packed k-values staged
// through LDS. The tile's 64 row offsets and 64 column offsets in the source
// block are decoded once per tile, so an element costs one load, one LDS
// write, one LDS read and one store; the source is read with lanes along its
// contiguous side (rows or k, chosen per job on the host) and every packed
// row segment is written whole, 64 consecutive values per wave.
template <typename T, bool SPLIT>
__global__ __launch_bounds__(256) void k_pack_tiled(const PackJob *__restrict__ jobs, int njobs, int64_t ntiles)
{
  constexpr int TS = kPackTile;
  __shared__ T tile[TS][TS + 1];
  __shared__ int64_t roff[TS], coff[TS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].tile0 <= t) lo = mid;
      else hi = mid - 1;
    }
    const PackJob &J = jobs[lo];
    const uint32_t local = (uint32_t)(t - J.tile0), tk_n = (uint32_t)J.tiles_k;
    const uint32_t tr = local / tk_n, tk = local - tr * tk_n;
    const int64_t r0 = (int64_t)tr * TS, k0 = (int64_t)tk * TS;
    if (tid < TS) {
      const int64_t kk = k0 + tid;
      coff[tid] = kk < J.k ? map_off(J.cols, (uint32_t)kk) : -1;  // k <= kk < kp: zero padding
    } else if (tid < 2 * TS) {
      const int64_t r = r0 + (tid - TS);
      roff[tid - TS] = r < J.nrows ? map_off(J.rows, (uint32_t)r) : -1;
    }
    __syncthreads();
    const T *src = static_cast<const T *>(J.src);
    const bool lr = J.lane_rows != 0;
#pragma unroll 4
    for (int q = 0; q < TS / 4; ++q) {
      const int b = wave + 4 * q;
      const int i = lr ? lane : b, j = lr ? b : lane;
      const int64_t ro = roff[i], co = coff[j];
      tile[i][j] = (ro >= 0 && co >= 0) ? src[ro + co] : T(0);
    }
    __syncthreads();
    const int64_t kp = J.kp;
    if constexpr (SPLIT) {
      // the tile is 4 row bands x 2 k-blocks of the tile-blocked planes; each
      // wave splits two whole 1 KB blocks, lane l filling slot l of each
      const int64_t nbands = (J.nrows + 15) / 16, nkb = kp / 32;
      const size_t plane = b3_plane(J.nrows, kp);
      unsigned short *d16 = static_cast<unsigned short *>(J.dst);
      const int rr = lane >> 2, c = PackLay::pos(rr, lane & 3);
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int blk = wave * 2 + u, band = blk >> 1, kbl = blk & 1;
        const int64_t gband = r0 / 16 + band, gkb = k0 / 32 + kbl;
        if (gband >= nbands || gkb >= nkb) continue;
        us8 h0, h1, h2;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          unsigned short e0, e1, e2;
          split3((float)tile[band * 16 + rr][kbl * 32 + c * 8 + e], e0, e1, e2);
          h0[e] = e0; h1[e] = e1; h2[e] = e2;
        }
        const size_t off = (size_t)gband * 16 * kp + (size_t)gkb * 512 + lane * 8;
        *reinterpret_cast<us8 *>(d16 + off) = h0;
        *reinterpret_cast<us8 *>(d16 + plane + off) = h1;
        *reinterpret_cast<us8 *>(d16 + 2 * plane + off) = h2;
      }
    } else {
#pragma unroll 4
      for (int q = 0; q < TS / 4; ++q) {
        const int i = wave + 4 * q;
        const int64_t r = r0 + i, kk = k0 + lane;
        if (r < J.nrows && kk < kp) static_cast<T *>(J.dst)[r * kp + kk] = tile[i][lane];
      }
    }
    __syncthreads();
  }
}

template <typename T, int OCC>
__global__ __launch_bounds__(256, OCC) void k_grouped_gemm(const BlockOp *__restrict__ ops,
                                                           const WorkItem *__restrict__ items,
                                                           WorkQueues *__restrict__ queues, double alpha, int accumulate)
{
  constexpr int BM = kGroupedTile, BN = kGroupedTile, BK = 16, KP = BK + (sizeof(T) == 4 ? 4 : 2);
  constexpr int FM = BM / 32, FN = BN / 32;
  constexpr int LA = BM * BK / 256;           // elements staged per thread and operand
  constexpr int VW = 16 / (int)sizeof(T);     // elements per 16-byte access
  typedef typename GAcc<T>::type AccT;
  typedef typename GF4<T>::type F4;
  __shared__ __align__(16) T As[2][BM][KP];
  __shared__ __align__(16) T Bs[2][BN][KP];
  __shared__ int next_item;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
  const int s_r = tid / (BK / LA), s_k = (tid % (BK / LA)) * LA;  // staging row, first k
  const int fr = lane & 15, fg = (lane >> 4) * 4, fk = lane >> 4;
  while (true) {
    if (tid == 0) next_item = take_item(queues);
    __syncthreads();
    // uniform across the wave, so that the item's and its block products'
    // descriptors are read with scalar loads
    const int it = __builtin_amdgcn_readfirstlane(next_item);
    __syncthreads();  // everyone has read next_item before it is replaced
    if (it < 0) break;
    const WorkItem w = items[it];
    const int64_t m0 = (int64_t)w.mt * BM, n0 = (int64_t)w.nt * BN;
    AccT acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = AccT{0, 0, 0, 0};
    for (int oi = w.op_begin; oi < w.op_end; ++oi) {
      const BlockOp &op = ops[oi];
      const int64_t kp = op.kp;
      const bool aok = m0 + s_r < op.m, bok = n0 + s_r < op.n;
      const GlobalPtr<const T> ap = as_global<const T>(const_cast<void *>(op.a)) + (aok ? (m0 + s_r) * kp + s_k : 0);
      const GlobalPtr<const T> bp = as_global<const T>(const_cast<void *>(op.b)) + (bok ? (n0 + s_r) * kp + s_k : 0);
      T ra[LA], rb[LA];
      auto load = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < LA; i += VW) {
          typedef unsigned int u4 __attribute__((ext_vector_type(4)));
          *reinterpret_cast<u4 *>(&ra[i]) = aok ? *reinterpret_cast<GlobalPtr<const u4>>(ap + k0 + i) : u4{0, 0, 0, 0};
          *reinterpret_cast<u4 *>(&rb[i]) = bok ? *reinterpret_cast<GlobalPtr<const u4>>(bp + k0 + i) : u4{0, 0, 0, 0};
        }
      };
      auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < LA; i += VW) {
          *reinterpret_cast<float4 *>(&As[buf][s_r][s_k + i]) = *reinterpret_cast<const float4 *>(&ra[i]);
          *reinterpret_cast<float4 *>(&Bs[buf][s_r][s_k + i]) = *reinterpret_cast<const float4 *>(&rb[i]);
        }
      };
      const int64_t nk = kp / BK;
      load(0);
      store(0);
      __syncthreads();
      for (int64_t kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) load((kt + 1) * BK);
        F4 fa[FM], fb[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[i] = *reinterpret_cast<const F4 *>(&As[cur][wm + i * 16 + fr][fg]);
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = *reinterpret_cast<const F4 *>(&Bs[cur][wn + j * 16 + fr][fg]);
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
              const T a0 = st == 0 ? fa[i].x : st == 1 ? fa[i].y : st == 2 ? fa[i].z : fa[i].w;
              const T b0 = st == 0 ? fb[j].x : st == 1 ? fb[j].y : st == 2 ? fb[j].z : fb[j].w;
              acc[i][j] = gmfma(a0, b0, acc[i][j]);
            }
        if (kt + 1 < nk) store(cur ^ 1);
        __syncthreads();
      }
    }
    // scatter the tile into the D block (index maps of the chain's first op;
    // a complex block's product row 2r + c is component c of element r)
    const BlockOp &o0 = ops[w.op_begin];
    const GlobalPtr<T> D = as_global<T>(w.d);
    int32_t coff[FN];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int64_t col = n0 + wn + j * 16 + fr;
      coff[j] = col < o0.n ? (o0.cplx ? 2 : 1) * map_off32(o0.Nd, (uint32_t)col) : -1;
    }
    const T a = (T)alpha;
    // per 16-row block of the wave: the offsets (-1: outside the D block),
    // then every D load of the block at once (clamped addresses, no
    // lane-masked branches), one wait, then the stores. A load-add-store per
    // element waits for each load in turn behind the stores already issued.
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      int32_t off[4][FN];
      T dv[4][FN];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // FP64 16x16x4 keeps rows lane/16 + 4r in a lane, FP32 rows 4(lane/16) + r
        const int64_t row = m0 + wm + i * 16 + (sizeof(T) == 8 ? fk + 4 * r : fk * 4 + r);
        const int32_t roff = row >= o0.m ? -1
                           : o0.cplx ? 2 * map_off32(o0.Md, (uint32_t)(row >> 1)) + (int32_t)(row & 1)
                                     : map_off32(o0.Md, (uint32_t)row);
#pragma unroll
        for (int j = 0; j < FN; ++j) off[r][j] = (roff < 0 || coff[j] < 0) ? -1 : roff + coff[j];
      }
      if (accumulate) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int j = 0; j < FN; ++j) dv[r][j] = D[max(off[r][j], 0)];
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int j = 0; j < FN; ++j) dv[r][j] = T(0);
      }
      wait_all_vmem();  // on every lane, before the lane-masked stores
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          if (off[r][j] >= 0) D[off[r][j]] = a * acc[i][j][r] + dv[r][j];
    }
  }
}

// Three-way BF16 split form of k_grouped_gemm (R4 and C4 under fast math).
// The packed operands are tile-blocked split planes (see b3_packed_offset),
// so each k-stage of 32 is 48 whole 1 KB blocks (3 planes x 8 row bands of A
// and of B) that the 8 waves move with wave-wide DMA straight into LDS, six
// each, into a ring of three stage buffers (144 KB: two stages in flight
// while one is multiplied); the stage boundary is an explicit vmcnt wait and
// a bare barrier. A work item's chain is one pipeline: the DMA of the next
// k-stage may belong to the next block product of the chain. Each stage runs
// six 16x16x32 bf16 MFMA products per output block; the FP32 accumulator is
// folded into an FP64 sum every 8 stages. Waves are 2 x 4 over the 128 x 128
// tile (64 x 32 each, two per SIMD). Row bands beyond an operand's last are
// read from its last band (those product rows and columns are not stored).
constexpr int kB3GroupedThreads = 512;
__global__ __launch_bounds__(kB3GroupedThreads) void k_grouped_b3(const BlockOp *__restrict__ ops,
                                                                  const WorkItem *__restrict__ items,
                                                                  WorkQueues *__restrict__ queues, double alpha,
                                                                  int accumulate)
{
  constexpr int BM = kGroupedTile, BN = kGroupedTile, KB = 32, RB = 16;
  constexpr int WM = BM / 2, WN = BN / 4;      // wave sub-tile
  constexpr int FM = WM / 16, FN = WN / 16;    // 16x16 MFMA blocks per wave
  constexpr int NBA = BM / RB, NBB = BN / RB;  // row bands per stage and plane
  constexpr int NBLK = 3 * (NBA + NBB);        // 1 KB blocks per stage
  constexpr int PER = NBLK / 8;                // DMA blocks per wave per stage
  static_assert(PER * 8 == NBLK, "stage blocks spread evenly over the waves");
  constexpr int kStage = NBLK * 512;           // bf16 values per LDS stage
  constexpr int kFold = 8;
  __shared__ __align__(16) unsigned short lds[3 * kStage];
  __shared__ int next_item;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // the masks bound the fragment offsets inside one stage for the compiler,
  // which otherwise waits for the DMA into the other stages
  const int wm = (wave & 1) * WM, wn = ((wave >> 1) & 3) * WN;
  const int fr = lane & 15, fpos = PackLay::pos(fr, (lane >> 4) & 3) * 8;
  while (true) {
    if (tid == 0) next_item = take_item(queues);
    __syncthreads();
    const int it = next_item;
    __syncthreads();
    if (it < 0) break;
    const WorkItem w = items[it];
    const int64_t m0 = (int64_t)w.mt * BM, n0 = (int64_t)w.nt * BN;
    // k-stages of the whole chain
    int64_t nstages = 0;
    for (int oi = w.op_begin; oi < w.op_end; ++oi) nstages += ops[oi].kp / KB;
    // DMA cursor: block product and k-stage of the next stage to issue
    int iop = w.op_begin;
    int64_t ikb = 0;
    auto issue = [&](unsigned short *__restrict__ buf) {
      const BlockOp &op = ops[iop];
      const int64_t kp = op.kp, bands_a = (op.m + RB - 1) / RB, bands_b = (op.n + RB - 1) / RB;
      const int64_t pa = bands_a * RB * kp, pb = bands_b * RB * kp;  // plane strides
      const unsigned short *A = static_cast<const unsigned short *>(op.a);
      const unsigned short *B = static_cast<const unsigned short *>(op.b);
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int idx = wave + 8 * u;
        const unsigned short *g;
        if (idx < 3 * NBA) {
          const int64_t band = min(m0 / RB + idx % NBA, bands_a - 1);
          g = A + (idx / NBA) * pa + band * RB * kp;
        } else {
          const int jb = idx - 3 * NBA;
          const int64_t band = min(n0 / RB + jb % NBB, bands_b - 1);
          g = B + (jb / NBB) * pb + band * RB * kp;
        }
        lds_dma16(g + ikb * (RB * KB) + lane * 8, buf + idx * 512);
      }
      if (++ikb == kp / KB) {
        ++iop;
        ikb = 0;
      }
    };
    f4 acc[FM][FN];
    double sum[FM][FN][4];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        acc[i][j] = f4{0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; ++r) sum[i][j][r] = 0.0;
      }
    // stage st lives in buffer st % 3; the loop is unrolled over the three
    // buffers so that every LDS address is a constant offset into one array,
    // which lets the compiler see that the fragment reads of one buffer do
    // not depend on the DMA still in flight into the others (otherwise it
    // waits for all outstanding DMA before the first read of every stage)
    if (nstages > 0) issue(lds);
    if (nstages > 1) issue(lds + kStage);
    auto stage = [&](auto J, int64_t st) {
      constexpr int j = decltype(J)::value;
      const unsigned short *cur = lds + j * kStage;
      if (st + 1 < nstages) wait_vmcnt<PER>();  // stage st landed (st + 1 may still fly)
      else wait_vmcnt<0>();
      stage_barrier();  // everyone's stage st landed; stage st - 1's buffer is free
      if (st + 2 < nstages) issue(lds + ((j + 2) % 3) * kStage);
      const unsigned short *rb = cur + 3 * BM * KB;
      us8 fb[3][FN];
#pragma unroll
      for (int pl = 0; pl < 3; ++pl)
#pragma unroll
        for (int jj = 0; jj < FN; ++jj)
          fb[pl][jj] = *reinterpret_cast<const us8 *>(rb + (pl * BN + wn + jj * 16 + fr) * KB + fpos);
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        us8 fa[3];
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) fa[pl] = *reinterpret_cast<const us8 *>(cur + (pl * BM + wm + i * 16 + fr) * KB + fpos);
        // products outer, blocks inner; smallest terms first
#pragma unroll
        for (int t = 0; t < 6; ++t) {
          const int qa = t == 0 ? 0 : t == 1 ? 1 : t == 2 ? 2 : t == 3 ? 0 : t == 4 ? 1 : 0;
          const int qb = t == 0 ? 2 : t == 1 ? 1 : t == 2 ? 0 : t == 3 ? 1 : t == 4 ? 0 : 0;
#pragma unroll
          for (int jj = 0; jj < FN; ++jj) acc[i][jj] = mfma_bf(fa[qa], fb[qb][jj], acc[i][jj]);
        }
      }
      if ((st + 1) % kFold == 0) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int jj = 0; jj < FN; ++jj) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sum[i][jj][r] += (double)acc[i][jj][r];
            acc[i][jj] = f4{0, 0, 0, 0};
          }
      }
    };
    for (int64_t st = 0; st < nstages; st += 3) {
      stage(std::integral_constant<int, 0>{}, st);
      if (st + 1 < nstages) stage(std::integral_constant<int, 1>{}, st + 1);
      if (st + 2 < nstages) stage(std::integral_constant<int, 2>{}, st + 2);
    }
    // the LDS stages are reused by the next item only after every wave has
    // passed the item's first barrier, which follows these reads
    const BlockOp &o0 = ops[w.op_begin];
    const GlobalPtr<float> D = as_global<float>(w.d);
    int32_t coff[FN];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int64_t col = n0 + wn + j * 16 + fr;
      coff[j] = col < o0.n ? (o0.cplx ? 2 : 1) * map_off32(o0.Nd, (uint32_t)col) : -1;
    }
    // per 16-row block of the wave: offsets, then the block's D loads, one
    // wait, then the stores (see k_grouped_gemm)
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      int32_t off[4][FN];
      float dv[4][FN];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm + i * 16 + ((lane >> 4) & 3) * 4 + r;  // FP32 16x16 layout
        const int32_t roff = row >= o0.m ? -1
                           : o0.cplx ? 2 * map_off32(o0.Md, (uint32_t)(row >> 1)) + (int32_t)(row & 1)
                                     : map_off32(o0.Md, (uint32_t)row);
#pragma unroll
        for (int j = 0; j < FN; ++j) off[r][j] = (roff < 0 || coff[j] < 0) ? -1 : roff + coff[j];
      }
      if (accumulate) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int j = 0; j < FN; ++j) dv[r][j] = D[max(off[r][j], 0)];
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int j = 0; j < FN; ++j) dv[r][j] = 0.f;
      }
      wait_all_vmem();  // on every lane, before the lane-masked stores
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          if (off[r][j] >= 0)
            D[off[r][j]] = (float)(alpha * (sum[i][j][r] + (double)acc[i][j][r]) + (double)dv[r][j]);
    }
  }
}

}  // namespace

int launch_pack(int kind, bool split3, const PackJob *jobs, int njobs, int64_t nchunks, int grid, hipStream_t s)
{
  if (njobs <= 0 || nchunks <= 0) return 0;
  const bool dp = kind == R8 || kind == C8, sp = kind == R4 || kind == C4;
  if (dp && !split3) hipLaunchKernelGGL((k_pack<double, false>), dim3(grid), dim3(256), 0, s, jobs, njobs, nchunks);
  else if (sp && !split3) hipLaunchKernelGGL((k_pack<float, false>), dim3(grid), dim3(256), 0, s, jobs, njobs, nchunks);
  else if (sp) hipLaunchKernelGGL((k_pack<float, true>), dim3(grid), dim3(256), 0, s, jobs, njobs, nchunks);
  else return TALSH_INVALID_ARGS;
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

int launch_pack_tiled(int kind, bool split3, const PackJob *jobs, int njobs, int64_t ntiles, int grid,
                      hipStream_t s)
{
  if (njobs <= 0 || ntiles <= 0) return 0;
  if (kind == R8 && !split3)
    hipLaunchKernelGGL((k_pack_tiled<double, false>), dim3(grid), dim3(256), 0, s, jobs, njobs, ntiles);
  else if (kind == R4 && !split3)
    hipLaunchKernelGGL((k_pack_tiled<float, false>), dim3(grid), dim3(256), 0, s, jobs, njobs, ntiles);
  else if (kind == R4)
    hipLaunchKernelGGL((k_pack_tiled<float, true>), dim3(grid), dim3(256), 0, s, jobs, njobs, ntiles);
  else
    return TALSH_INVALID_ARGS;
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

int launch_grouped(int kind, bool split3, const BlockOp *ops, const WorkItem *items, WorkQueues *queues,
                   double alpha, bool accumulate, int grid, hipStream_t s)
{
  if (grid <= 0) return 0;
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(1), 0, s, queues);
  if (split3) {
    if (kind != R4 && kind != C4) return TALSH_INVALID_ARGS;
    hipLaunchKernelGGL(k_grouped_b3, dim3(grid), dim3(kB3GroupedThreads), 0, s, ops, items, queues, alpha,
                       accumulate ? 1 : 0);
  } else if (kind == R8 || kind == C8)
    hipLaunchKernelGGL((k_grouped_gemm<double, 2>), dim3(grid), dim3(256), 0, s, ops, items, queues, alpha,
                       accumulate ? 1 : 0);
  else if (kind == R4 || kind == C4)
    hipLaunchKernelGGL((k_grouped_gemm<float, 2>), dim3(grid), dim3(256), 0, s, ops, items, queues, alpha,
                       accumulate ? 1 : 0);
  else
    return TALSH_INVALID_ARGS;
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

}  // namespace xt
