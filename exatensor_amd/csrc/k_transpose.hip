// DLF tensor permutation for gfx950 (the engine's transpose/copy kernel).
//
// Replaces the reference's gpu_tensor_block_copy_dlf__ /
// gpu_tensor_block_copy_scatter_dlf__ (TALSH/tensor_algebra_gpu_nvidia.cu:1125-1457,
// :2172-2260). Semantics: out[idx'] = in[idx], input dimension i becoming output
// dimension o2n[i]; pure data movement, bit-exact.
//
// Design:
//  * host planner: drop unit extents, fuse input dimensions that stay adjacent
//    and in order in the output, then pick an LDS tile = the union of an
//    input-minor dimension group and an output-minor dimension group, grown
//    until both contiguous runs reach >= 256 bytes (the last dimension of each
//    group may be split, preferably so that the run is a whole number of
//    128-byte lines; ragged splits are masked), then grown to >= 8192 elements;
//  * LDS padding is chosen on the host by simulating the bank conflicts of the
//    store phase for each candidate padding period;
//  * persistent kernel (two resident 512-thread workgroups per CU): each thread decodes
//    the intra-tile offsets of its vectors once (input offset, output offset,
//    LDS read base) and keeps them in registers, so the per-tile work is one
//    base-offset decode per workgroup plus one coalesced load and one
//    coalesced store per vector.
#include <algorithm>
#include <mutex>
#include <string>
#include <unordered_map>
#include <cstring>
#include <numeric>

#include "kernels.hpp"

namespace xt {
namespace {

constexpr int kNT = 256;           // threads per workgroup (copy and scatter kernels)
constexpr int kTNT = 512;          // threads per workgroup of the tiled kernel
constexpr int kMaxTileDims = 8;    // dimensions inside one tile
constexpr int kMaxGridDims = 16;   // fused rank handled by the tiled path
constexpr int kMaxEPT = 8;         // tile vectors per thread (more cost occupancy)
constexpr int kTiledPerCU = 2;     // persistent workgroups per CU (16 waves)
constexpr int kTileMaxBytes = 65536;  // LDS tile (before padding)
constexpr long kTileMinVol = 8192;    // tiles are grown to at least this many elements
constexpr int kLine = 128;            // cache line bytes

struct DevDiv { uint32_t d, m, s; };

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const DevDiv &f)
{
  // d == 0 encodes "no division" (result 0); d == 1 is the identity
  if (f.d == 0) return 0;
  return (f.d == 1) ? n : ((__umulhi(n, f.m) + n) >> f.s);
}

struct TileParams {
  // tile dimensions in input (load) order
  int nt;
  DevDiv t_in[kMaxTileDims];       // tile extents, load order
  uint64_t is_in[kMaxTileDims];    // input strides, load order
  // tile dimensions in output (store) order
  DevDiv t_out[kMaxTileDims];      // tile extents, store order
  uint64_t os_out[kMaxTileDims];   // output strides, store order
  uint32_t ls_out[kMaxTileDims];   // LDS (load-order linear) strides, store order
  // ragged (partially tiled) dimensions: position in load / store order
  int nrag;
  int rag_in[2], rag_out[2];
  uint32_t rag_t[2];
  // grid over all fused dims
  int ng;
  DevDiv gdiv[kMaxGridDims];       // number of tiles along each dim
  uint64_t g_is[kMaxGridDims];     // input stride * tile extent
  uint64_t g_os[kMaxGridDims];     // output stride * tile extent
  uint32_t g_t[kMaxGridDims];      // tile extent along the dim
  uint32_t g_ext[kMaxGridDims];    // full extent
  int g_rag[kMaxGridDims];         // ragged slot (0/1) or -1
  uint32_t vol_tile;               // elements per tile
  uint32_t ntiles;
  DevDiv pad;                      // LDS slot = e + e / pad.d (pad.d = 0: no padding)
  DevDiv tv_in0, tv_out0;          // first tile extent on each side divided by the vector width
};

// Elements move as integer words of their size (pure data movement, so the
// result is bit-exact by construction); conjugation flips the sign bit of the
// imaginary half of a complex word.
template <typename W, bool CONJ>
__device__ __forceinline__ W conj_if(W v) { return v; }
template <>
__device__ __forceinline__ uint2 conj_if<uint2, true>(uint2 v) { v.y ^= 0x80000000u; return v; }
template <>
__device__ __forceinline__ uint4 conj_if<uint4, true>(uint4 v) { v.w ^= 0x80000000u; return v; }

template <typename W, int VEC>
struct __align__(sizeof(W) * VEC) Pack { W w[VEC]; };

// Hides a value's origin from the optimizer, so that addresses derived from it
// are recomputed where they are used instead of being hoisted out of the tile
// loop and held in registers for every element.
__device__ __forceinline__ void opaque(uint32_t &x) { asm volatile("" : "+v"(x)); }

// Streaming (nontemporal) global accesses: every element is read once and
// written once, so the data should not displace useful lines from L2/MALL.
typedef unsigned int nt1 __attribute__((ext_vector_type(1)));
typedef unsigned int nt2 __attribute__((ext_vector_type(2)));
typedef unsigned int nt4 __attribute__((ext_vector_type(4)));
template <int BYTES>
struct NtWord;
template <>
struct NtWord<4> { typedef unsigned int type; };
template <>
struct NtWord<8> { typedef nt2 type; };
template <>
struct NtWord<16> { typedef nt4 type; };

template <typename P>
__device__ __forceinline__ P ld_stream(const P *p)
{
  typedef typename NtWord<sizeof(P)>::type U;
  const U u = __builtin_nontemporal_load(reinterpret_cast<const U *>(p));
  return __builtin_bit_cast(P, u);
}
template <typename P>
__device__ __forceinline__ void st_stream(P *p, const P &v)
{
  typedef typename NtWord<sizeof(P)>::type U;
  __builtin_nontemporal_store(__builtin_bit_cast(U, v), reinterpret_cast<U *>(p));
}

// One workgroup walks tiles (persistent grid). Each thread owns EPT vectors of
// VEC consecutive elements along the input-minor dimension for the load and
// along the output-minor dimension for the store, so global accesses are
// 16-byte wide whenever both minor extents allow it.
//
// Register budget: a thread keeps per vector only its input offset, output
// offset and LDS read base (plus the ragged-edge coordinates); the LDS slots
// are recomputed per tile from those (a few integer operations) instead of
// being held for every element, which keeps the common instantiations at
// 30-60 VGPRs, i.e. at full occupancy.
template <typename W, int VEC, int EPT, bool CONJ, typename OffT, bool RAG>
__global__ __launch_bounds__(kTNT) void k_permute_tiled(const W *__restrict__ in, W *__restrict__ out,
                                                        const TileParams p)
{
  typedef Pack<W, VEC> P;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  W *lds = reinterpret_cast<W *>(lds_raw);
  const uint32_t tid = threadIdx.x;
  const uint32_t nvec = p.vol_tile / VEC;
  OffT in_off[EPT], out_off[EPT];
  uint32_t st_ls[EPT];
  uint32_t rin[RAG ? EPT : 1], rout[RAG ? EPT : 1];
  // one-time decode of this thread's tile vectors (kept in registers)
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const uint32_t v = tid + (uint32_t)k * kTNT;
    uint32_t r = v, cin[2] = {0, 0}, cout[2] = {0, 0};
    OffT io = 0, oo = 0;
    uint32_t ls = 0;
    for (int j = 0; j < p.nt; ++j) {
      const DevDiv &dv = j == 0 ? p.tv_in0 : p.t_in[j];
      const uint32_t q = fdiv(r, dv);
      uint32_t c = r - q * dv.d;
      if (j == 0) c *= VEC;
      r = q;
      io += (OffT)c * (OffT)p.is_in[j];
      if (RAG) {
        if (p.rag_in[0] == j) cin[0] = c;
        if (p.nrag > 1 && p.rag_in[1] == j) cin[1] = c;
      }
    }
    r = v;
    for (int j = 0; j < p.nt; ++j) {
      const DevDiv &dv = j == 0 ? p.tv_out0 : p.t_out[j];
      const uint32_t q = fdiv(r, dv);
      uint32_t c = r - q * dv.d;
      if (j == 0) c *= VEC;
      r = q;
      oo += (OffT)c * (OffT)p.os_out[j];
      ls += c * p.ls_out[j];
      if (RAG) {
        if (p.rag_out[0] == j) cout[0] = c;
        if (p.nrag > 1 && p.rag_out[1] == j) cout[1] = c;
      }
    }
    in_off[k] = io;
    out_off[k] = oo;
    st_ls[k] = ls;
    if (RAG) {
      const bool valid = v < nvec;
      rin[k] = valid ? (cin[0] | (cin[1] << 16)) : 0xFFFFFFFFu;
      rout[k] = valid ? (cout[0] | (cout[1] << 16)) : 0xFFFFFFFFu;
    }
  }
  const uint32_t step = p.ls_out[0];
  struct TileBase { OffT bin, bout; uint32_t rem0, rem1; bool full; };
  // tile base decode (uniform across the workgroup)
  auto decode = [&](uint32_t tile, TileBase &s) {
    s.bin = 0;
    s.bout = 0;
    uint32_t rem[2] = {0xFFFFu, 0xFFFFu};
    uint32_t r = tile;
    for (int g = 0; g < p.ng; ++g) {
      const uint32_t q = fdiv(r, p.gdiv[g]);
      const uint32_t c = r - q * p.gdiv[g].d;
      r = q;
      s.bin += (OffT)c * (OffT)p.g_is[g];
      s.bout += (OffT)c * (OffT)p.g_os[g];
      if (RAG && p.g_rag[g] >= 0) {
        const uint32_t left = p.g_ext[g] - c * p.g_t[g];
        rem[p.g_rag[g]] = left < p.g_t[g] ? left : p.g_t[g];
      }
    }
    s.rem0 = rem[0];
    s.rem1 = rem[1];
    s.full = !RAG || ((rem[0] == p.rag_t[0]) && (p.nrag < 2 || rem[1] == p.rag_t[1]));
  };
  P val[EPT];
  auto load = [&](const TileBase &s) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint32_t v = tid + (uint32_t)k * kTNT;
      bool ok = v < nvec;
      if (RAG && !s.full) ok = ok && (rin[k] & 0xFFFFu) < s.rem0 && (rin[k] >> 16) < s.rem1;
      if (ok) val[k] = ld_stream(reinterpret_cast<const P *>(in + (s.bin + in_off[k])));
    }
  };
  // the padding period is a multiple of VEC (checked by the planner), so the
  // VEC slots of one loaded vector are consecutive: base slot + j
  auto to_lds = [&](const TileBase &s) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      uint32_t v = tid + (uint32_t)k * kTNT;
      bool ok = v < nvec;
      if (RAG && !s.full) ok = ok && (rin[k] & 0xFFFFu) < s.rem0 && (rin[k] >> 16) < s.rem1;
      if (ok) {
        opaque(v);
        const uint32_t e0 = v * VEC;
        const uint32_t b = e0 + fdiv(e0, p.pad);
#pragma unroll
        for (int j = 0; j < VEC; ++j) lds[b + j] = val[k].w[j];
      }
    }
  };
  auto from_lds = [&](const TileBase &s) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const uint32_t f = tid + (uint32_t)k * kTNT;
      bool ok = f < nvec;
      if (RAG && !s.full) ok = ok && (rout[k] & 0xFFFFu) < s.rem0 && (rout[k] >> 16) < s.rem1;
      if (ok) {
        uint32_t base = st_ls[k];
        opaque(base);
        P o;
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
          const uint32_t e = base + j * step;
          o.w[j] = conj_if<W, CONJ>(lds[e + fdiv(e, p.pad)]);
        }
        st_stream(reinterpret_cast<P *>(out + (s.bout + out_off[k])), o);
      }
    }
  };
  uint32_t tile = blockIdx.x;
  if (tile >= p.ntiles) return;
  TileBase cur;
  decode(tile, cur);
  load(cur);
  // software pipeline: the next tile's loads are in flight while the current
  // tile drains from LDS to memory
  while (true) {
    to_lds(cur);
    __syncthreads();
    const uint32_t next = tile + gridDim.x;
    TileBase nxt;
    if (next < p.ntiles) {
      decode(next, nxt);
      load(nxt);
    }
    from_lds(cur);
    if (next >= p.ntiles) break;
    __syncthreads();
    tile = next;
    cur = nxt;
  }
}

// Generic per-element fallback (any rank): decode the input index, scatter.
struct ScatterParams {
  int rank;
  uint64_t vol;
  uint64_t ext[kMaxRank];
  uint64_t ostr[kMaxRank];
};

template <typename T, bool CONJ>
__global__ __launch_bounds__(kNT) void k_permute_scatter(const T *__restrict__ in, T *__restrict__ out,
                                                          const ScatterParams p)
{
  const uint64_t stride = (uint64_t)gridDim.x * kNT;
  for (uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x; i < p.vol; i += stride) {
    uint64_t r = i, o = 0;
    for (int d = 0; d < p.rank; ++d) {
      const uint64_t q = r / p.ext[d];
      o += (r - q * p.ext[d]) * p.ostr[d];
      r = q;
    }
    out[o] = conj_if<T, CONJ>(in[i]);
  }
}

template <typename T, bool CONJ>
__global__ __launch_bounds__(kNT) void k_copy(const T *__restrict__ in, T *__restrict__ out, uint64_t n)
{
  const uint64_t stride = (uint64_t)gridDim.x * kNT;
  for (uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += stride) out[i] = conj_if<T, CONJ>(in[i]);
}

// Identity permutation of a 16-byte aligned body: one 16-byte streaming
// access per thread in one-wave workgroups, one workgroup per KiB. The
// dispatcher hands the workgroups out in index order, so the lines in flight
// at any moment form a compact window that sweeps the body; on this part
// that moves more bytes per second than any persistent grid-stride form
// (measured: 6.4-6.7 TB/s against 5.5-5.8 TB/s).
constexpr int kCopyNT = 64;
template <typename W, bool CONJ>
__global__ __launch_bounds__(kCopyNT) void k_copy16(const nt4 *__restrict__ in, nt4 *__restrict__ out, uint64_t n16)
{
  const uint64_t i = (uint64_t)blockIdx.x * kCopyNT + threadIdx.x;
  if (i >= n16) return;
  nt4 v = __builtin_nontemporal_load(in + i);
  if (CONJ) {  // complex words: flip the sign bit of each imaginary part
    if (sizeof(W) == 8) { v.y ^= 0x80000000u; v.w ^= 0x80000000u; }
    else v.w ^= 0x80000000u;
  }
  __builtin_nontemporal_store(v, out + i);
}

// ---------------------------------------------------------------------------
// host planning

struct Fused {
  int rank = 0;
  std::vector<int> ext;   // fused extents, input order
  std::vector<int> o2n;   // output position of each fused dim
};

Fused fuse_dims(int rank, const int *dims, const int *o2n)
{
  // keep non-unit dims, re-rank their output positions
  std::vector<int> keep;
  for (int i = 0; i < rank; ++i)
    if (dims[i] > 1) keep.push_back(i);
  std::vector<int> order(keep.size());
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return o2n[keep[a]] < o2n[keep[b]]; });
  std::vector<int> pos(keep.size());
  for (size_t q = 0; q < order.size(); ++q) pos[order[q]] = (int)q;
  // fuse runs i, i+1 with pos[i+1] == pos[i] + 1 while the fused extent stays
  // an int (a run left unfused is still a valid description of the copy)
  Fused f;
  std::vector<int> grp_pos;
  for (size_t i = 0; i < keep.size(); ++i) {
    if (i > 0 && pos[i] == pos[i - 1] + 1 && (int64_t)f.ext.back() * dims[keep[i]] <= INT32_MAX) {
      f.ext.back() *= dims[keep[i]];
    } else {
      f.ext.push_back(dims[keep[i]]);
      grp_pos.push_back(pos[i]);
    }
  }
  f.rank = (int)f.ext.size();
  std::vector<int> ord(f.rank);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return grp_pos[a] < grp_pos[b]; });
  f.o2n.assign(f.rank, 0);
  for (int q = 0; q < f.rank; ++q) f.o2n[ord[q]] = q;
  return f;
}

// Smallest divisor of ext in [need, 2*need], else need itself. With
// unit_bytes > 0 (the bytes of the contiguous run the split extends), the
// smallest divisor in [need, 4*need] that makes the run a whole number of
// cache lines is preferred: runs that end mid-line leave lines split between
// tiles.
int pick_split(int ext, int need, long unit_bytes = 0)
{
  if (unit_bytes > 0)
    for (int t = need; t <= std::min(ext, 4 * need); ++t)
      if (ext % t == 0 && (unit_bytes * t) % kLine == 0) return t;
  for (int t = need; t <= std::min(ext, 2 * need); ++t)
    if (ext % t == 0) return t;
  return std::min(ext, need);
}

struct TilePlan {
  std::vector<int> t;   // tile extent per fused dim (0: not a tile dim)
  long vol = 0;
  int in_run = 0, out_run = 0;
};

TilePlan choose_tile(const Fused &f, int target_in, int target_out, int align_es)
{
  const int r = f.rank;
  std::vector<int> n2o(r);
  for (int i = 0; i < r; ++i) n2o[f.o2n[i]] = i;
  TilePlan tp;
  tp.t.assign(r, 0);
  long run = 1;
  for (int i = 0; i < r; ++i) {
    const long need = (target_in + run - 1) / run;
    if (need >= f.ext[i]) {
      tp.t[i] = f.ext[i];
      run *= f.ext[i];
      if (run >= target_in) break;
    } else {
      tp.t[i] = pick_split(f.ext[i], (int)need, run * align_es);
      run *= tp.t[i];
      break;
    }
  }
  tp.in_run = (int)std::min<long>(run, 1 << 30);
  run = 1;
  for (int q = 0; q < r; ++q) {
    const int i = n2o[q];
    const long need = (target_out + run - 1) / run;
    if (tp.t[i] > 0 && tp.t[i] < f.ext[i] && tp.t[i] < need)
      tp.t[i] = pick_split(f.ext[i], (int)std::min<long>(need, f.ext[i]), run * align_es);
    if (tp.t[i] == 0) tp.t[i] = need >= f.ext[i] ? f.ext[i] : pick_split(f.ext[i], (int)need, run * align_es);
    run *= tp.t[i];
    if (tp.t[i] < f.ext[i] || run >= target_out) break;
  }
  tp.out_run = (int)std::min<long>(run, 1 << 30);
  // the input run may have grown through the output walk
  run = 1;
  for (int i = 0; i < r; ++i) {
    if (tp.t[i] == 0) break;
    run *= tp.t[i];
    if (tp.t[i] < f.ext[i]) break;
  }
  tp.in_run = (int)std::min<long>(run, 1 << 30);
  tp.vol = 1;
  for (int i = 0; i < r; ++i)
    if (tp.t[i]) tp.vol *= tp.t[i];
  return tp;
}

// Grows a tile that meets its run targets but is too small to keep a
// workgroup busy (e.g. when the input- and output-minor dimensions coincide):
// alternately extends the input-order and output-order prefixes (doubling the
// first partial dimension, or opening the next one) until vol >= min_vol.
void grow_tile(const Fused &f, TilePlan &tp, long min_vol, long max_vol)
{
  const int r = f.rank;
  std::vector<int> n2o(r);
  for (int i = 0; i < r; ++i) n2o[f.o2n[i]] = i;
  auto volume = [&]() {
    long v = 1;
    for (int i = 0; i < r; ++i)
      if (tp.t[i]) v *= tp.t[i];
    return v;
  };
  int side = 0, stuck = 0;
  while (volume() < min_vol && stuck < 2) {
    int dim = -1;
    for (int q = 0; q < r; ++q) {
      const int i = side == 0 ? q : n2o[q];
      if (tp.t[i] < f.ext[i]) { dim = i; break; }
    }
    side ^= 1;
    if (dim < 0) { ++stuck; continue; }
    const int old = tp.t[dim];
    const int want = old == 0 ? 2 : old * 2;
    const int nt = pick_split(f.ext[dim], std::min(want, f.ext[dim]));
    const long nv = volume() / std::max(old, 1) * nt;
    if (nv > max_vol) { ++stuck; continue; }
    stuck = 0;
    tp.t[dim] = nt;
  }
  tp.vol = volume();
  long run = 1;
  for (int i = 0; i < r; ++i) {
    if (tp.t[i] == 0) break;
    run *= tp.t[i];
    if (tp.t[i] < f.ext[i]) break;
  }
  tp.in_run = (int)std::min<long>(run, 1 << 30);
  run = 1;
  for (int q = 0; q < r; ++q) {
    const int i = n2o[q];
    if (tp.t[i] == 0) break;
    run *= tp.t[i];
    if (tp.t[i] < f.ext[i]) break;
  }
  tp.out_run = (int)std::min<long>(run, 1 << 30);
}

// LDS bank-conflict degree of the store phase for a padding period.
int conflict_cost(const std::vector<uint32_t> &slots, int words_per_elem, uint32_t period)
{
  int cost = 0;
  for (int banks : {32, 64}) {
    for (size_t base = 0; base + 1 <= slots.size(); base += 64) {
      std::vector<std::vector<uint32_t>> hit(banks);
      for (size_t l = base; l < std::min(slots.size(), base + 64); ++l) {
        const uint32_t s = slots[l];
        const uint32_t slot = period ? s + s / period : s;
        for (int w = 0; w < words_per_elem; ++w) {
          const uint32_t word = slot * words_per_elem + w;
          auto &v = hit[word % banks];
          if (std::find(v.begin(), v.end(), word) == v.end()) v.push_back(word);
        }
      }
      int mx = 0;
      for (auto &v : hit) mx = std::max<int>(mx, (int)v.size());
      cost += mx;
    }
  }
  return cost;
}

DevDiv to_dev(uint32_t d)
{
  if (d == 0) return DevDiv{0, 0, 0};  // "no division" marker
  FastDiv f(d);
  return DevDiv{f.d, f.m, f.s};
}

struct Planned {
  int path = 0;  // 0 copy, 1 tiled, 2 scatter
  Fused f;
  TileParams tp{};
  int ept = 0;
  size_t lds_bytes = 0;
  int grid = 0;
  int vec = 1;        // elements per global access
  bool wide = false;  // 64-bit element offsets
  bool rag = false;   // masked edge tiles
  int cus = 0;        // compute units the grid was sized for
  TilePlan tile;
};

int num_cus_current()
{
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) cus = prop.multiProcessorCount;
  }
  return cus;
}

// Vector width the kernel can use for a tile: both minor dimensions (the first
// tile dimension in input and in output order) and their tile extents must be
// multiples of it, so that every vector is aligned and whole.
int tile_vec(const Fused &f, const TilePlan &tp, int es)
{
  int i0 = -1, o0 = -1;
  for (int i = 0; i < f.rank && i0 < 0; ++i)
    if (tp.t[i] > 0) i0 = i;
  std::vector<int> n2o(f.rank);
  for (int i = 0; i < f.rank; ++i) n2o[f.o2n[i]] = i;
  for (int q = 0; q < f.rank && o0 < 0; ++q)
    if (tp.t[n2o[q]] > 0) o0 = n2o[q];
  if (i0 < 0 || o0 < 0) return 1;
  for (int cand = 16 / es; cand >= 2; cand /= 2)
    if (f.ext[i0] % cand == 0 && tp.t[i0] % cand == 0 && f.ext[o0] % cand == 0 && tp.t[o0] % cand == 0) return cand;
  return 1;
}

int plan(int es, int rank, const int *dims, const int *o2n, Planned &pl, int cus, bool choose_pad = true)
{
  pl.f = fuse_dims(rank, dims, o2n);
  const Fused &f = pl.f;
  bool identity = true;
  for (int i = 0; i < f.rank; ++i) identity = identity && f.o2n[i] == i;
  if (f.rank <= 1 || identity) { pl.path = 0; return 0; }
  if (f.rank > kMaxGridDims) { pl.path = 2; return 0; }
  uint64_t vol = 1;
  for (int e : f.ext) vol *= (uint64_t)e;
  // a small tensor gets tiles small enough for about four per CU (a tile
  // takes one workgroup and the grid would otherwise leave CUs idle)
  const long min_vol = std::min<long>(kTileMinVol, std::max<long>(1024, (long)(vol / (4 * (uint64_t)std::max(cus, 1)))));
  // two attempts: first with splits that keep runs whole cache lines, then
  // without, if the aligned tile would need more vectors per thread than the
  // kernel is instantiated for
  TilePlan tp;
  bool ok = false;
  for (int attempt = 0; attempt < 2 && !ok; ++attempt) {
    const int align_es = attempt == 0 ? es : 0;
    // the second attempt bounds the tile by one element per vector, so that
    // it always fits the kernel's vector slots
    const long max_vol = std::min<long>((long)kTNT * kMaxEPT * (attempt == 0 ? 16 / es : 1), kTileMaxBytes / es);
    int target = std::max(8, 256 / es);
    while (true) {
      tp = choose_tile(f, target, target, align_es);
      int ntd = 0;
      for (int x : tp.t) ntd += x > 0;
      if ((tp.vol <= max_vol && ntd <= kMaxTileDims) || target <= 1) break;
      target = std::max(1, target / 2);
    }
    TilePlan g = tp;
    grow_tile(f, g, std::min(min_vol, max_vol), max_vol);
    int gd = 0;
    for (int x : g.t) gd += x > 0;
    if (gd <= kMaxTileDims) tp = g;
    ok = tp.vol <= max_vol && (tp.vol + (long)kTNT * tile_vec(f, tp, es) - 1) / ((long)kTNT * tile_vec(f, tp, es)) <= kMaxEPT;
  }
  int ntd = 0, nrag = 0;
  for (int i = 0; i < f.rank; ++i) {
    if (tp.t[i] > 0) ++ntd;
    if (tp.t[i] > 0 && f.ext[i] % tp.t[i] != 0) ++nrag;
  }
  uint64_t ntiles = 1;
  for (int i = 0; i < f.rank; ++i) ntiles *= (uint64_t)((f.ext[i] + std::max(tp.t[i], 1) - 1) / std::max(tp.t[i], 1));
  if (!ok || ntd > kMaxTileDims || nrag > 2 || ntiles >= (1ull << 31)) {
    pl.path = 2;
    return 0;
  }
  pl.path = 1;
  pl.tile = tp;
  TileParams &P = pl.tp;
  std::memset(&P, 0, sizeof(P));
  // strides
  std::vector<uint64_t> istr(f.rank), ostr_out(f.rank), ostr(f.rank);
  uint64_t s = 1;
  for (int i = 0; i < f.rank; ++i) { istr[i] = s; s *= (uint64_t)f.ext[i]; }
  std::vector<int> n2o(f.rank);
  for (int i = 0; i < f.rank; ++i) n2o[f.o2n[i]] = i;
  s = 1;
  for (int q = 0; q < f.rank; ++q) { ostr_out[q] = s; s *= (uint64_t)f.ext[n2o[q]]; }
  for (int i = 0; i < f.rank; ++i) ostr[i] = ostr_out[f.o2n[i]];
  // tile dims in load order
  std::vector<int> tin, tout;
  for (int i = 0; i < f.rank; ++i)
    if (tp.t[i] > 0) tin.push_back(i);
  for (int q = 0; q < f.rank; ++q)
    if (tp.t[n2o[q]] > 0) tout.push_back(n2o[q]);
  P.nt = (int)tin.size();
  std::vector<uint32_t> lstride(f.rank, 0);
  uint32_t ls = 1;
  for (int j = 0; j < P.nt; ++j) {
    P.t_in[j] = to_dev((uint32_t)tp.t[tin[j]]);
    P.is_in[j] = istr[tin[j]];
    lstride[tin[j]] = ls;
    ls *= (uint32_t)tp.t[tin[j]];
  }
  for (int j = 0; j < P.nt; ++j) {
    P.t_out[j] = to_dev((uint32_t)tp.t[tout[j]]);
    P.os_out[j] = ostr[tout[j]];
    P.ls_out[j] = lstride[tout[j]];
  }
  P.nrag = 0;
  for (int j = 0; j < P.nt; ++j) {
    const int i = tin[j];
    if (f.ext[i] % tp.t[i] != 0) {
      const int slot = P.nrag++;
      P.rag_in[slot] = j;
      P.rag_t[slot] = (uint32_t)tp.t[i];
      for (int jj = 0; jj < P.nt; ++jj)
        if (tout[jj] == i) P.rag_out[slot] = jj;
    }
  }
  // grid: tile index decoded fastest along the output's minor dimension, then
  // in output order, so that the tiles in flight at once (consecutive indices
  // across the persistent grid) write neighbouring regions of the output.
  // Measured against input order on the 174-transpose workload (+1.9%
  // geometric mean) and on 40 random shapes of ranks 2-6 (+5.9%).
  P.ng = f.rank;
  for (int q = 0; q < f.rank; ++q) {
    const int i = n2o[q];
    const uint32_t t = (uint32_t)std::max(tp.t[i], 1);
    const uint32_t n = (uint32_t)((f.ext[i] + t - 1) / t);
    P.gdiv[q] = to_dev(n);
    P.g_is[q] = istr[i] * t;
    P.g_os[q] = ostr[i] * t;
    P.g_t[q] = t;
    P.g_ext[q] = (uint32_t)f.ext[i];
    P.g_rag[q] = -1;
    for (int k = 0; k < P.nrag; ++k)
      if (tin[P.rag_in[k]] == i) P.g_rag[q] = k;
  }
  P.vol_tile = (uint32_t)tp.vol;
  P.ntiles = (uint32_t)ntiles;
  // LDS padding by simulated store-phase conflicts (first 256 store slots)
  std::vector<uint32_t> st;
  for (uint32_t e = 0; e < std::min<uint32_t>(P.vol_tile, kTNT); ++e) {
    uint32_t r = e, l = 0;
    for (int j = 0; j < P.nt; ++j) {
      const uint32_t d = P.t_out[j].d;
      l += (r % d) * P.ls_out[j];
      r /= d;
    }
    st.push_back(l);
  }
  const int words = std::max(1, es / 4);
  std::vector<uint32_t> cands = {0, 4, 8, 16, 32, 64, 128, 256};
  uint32_t prod = 1;
  for (int j = 0; j < P.nt && j < 3; ++j) { prod *= P.t_in[j].d; cands.push_back(prod); }
  uint32_t best = 0;
  int bestc = choose_pad ? conflict_cost(st, words, 0) : 0;
  for (uint32_t c : cands) {
    if (!choose_pad || c == 0 || c >= P.vol_tile) continue;
    const int cc = conflict_cost(st, words, c);
    if (cc < bestc) { bestc = cc; best = c; }
  }
  P.pad = to_dev(best);
  // vector width: both minor dimensions (input dim 0, output dim 0) and their
  // tile extents must be multiples of it so every vector is aligned and whole
  int vec = tile_vec(f, tp, es);
  // the kernel writes one vector to consecutive LDS slots, so a vector must
  // not straddle a padding period (every candidate period above is a
  // multiple of the width already; this keeps that invariant explicit)
  while (vec > 1 && best % (uint32_t)vec != 0) vec /= 2;
  pl.vec = vec;
  P.tv_in0 = to_dev((uint32_t)(tp.t[tin[0]] / vec));
  P.tv_out0 = to_dev((uint32_t)(tp.t[tout[0]] / vec));
  int ept = 1;
  while ((long)ept * kTNT * vec < tp.vol) ept *= 2;
  pl.ept = ept;
  pl.wide = vol >= (1ull << 32);
  pl.rag = P.nrag > 0;
  const size_t slots = P.vol_tile + (best ? P.vol_tile / best + 1 : 0);
  pl.lds_bytes = slots * (size_t)es;
  // resident workgroups per CU: measured on the transpose workload, 16 waves
  // per CU move more bytes per second than 32 (the DRAM sees fewer concurrent
  // streams); LDS may allow fewer
  const int per_cu = std::max(1, std::min(kTiledPerCU, (int)(160 * 1024 / std::max<size_t>(pl.lds_bytes, 1))));
  pl.cus = cus;
  pl.grid = (int)std::min<uint64_t>(ntiles, (uint64_t)cus * per_cu);
  return 0;
}

// Workgroups of one kernel instantiation that a CU holds at once as far as its
// registers allow (LDS is accounted for by the planner); the persistent grid
// is capped at this so that every workgroup is resident from the start.
template <typename W, int VEC, int EPT, bool CONJ, typename OffT, bool RAG>
int resident_per_cu()
{
  static const int n = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_permute_tiled<W, VEC, EPT, CONJ, OffT, RAG>, kTNT, 0) !=
            hipSuccess || b <= 0)
      b = 1;
    return b;
  }();
  return n;
}

template <typename W, int VEC, bool CONJ, typename OffT, bool RAG>
int launch_tiled_v(const Planned &pl, const void *in, void *out, hipStream_t s)
{
  const W *a = static_cast<const W *>(in);
  W *b = static_cast<W *>(out);
  const dim3 blk(kTNT);
  switch (pl.ept) {
#define XT_CASE(E)                                                                                             \
  case E: {                                                                                                    \
    const int cap = pl.cus * resident_per_cu<W, VEC, E, CONJ, OffT, RAG>();                                    \
    const dim3 g(std::min(pl.grid, cap));                                                                      \
    hipLaunchKernelGGL((k_permute_tiled<W, VEC, E, CONJ, OffT, RAG>), g, blk, pl.lds_bytes, s, a, b, pl.tp);   \
    break;                                                                                                     \
  }
    XT_CASE(1) XT_CASE(2) XT_CASE(4) XT_CASE(8)
#undef XT_CASE
  default: return TALSH_FAILURE;
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

template <typename W, int VEC, bool CONJ>
int launch_tiled_w(const Planned &pl, const void *in, void *out, hipStream_t s)
{
  if (pl.wide) return pl.rag ? launch_tiled_v<W, VEC, CONJ, uint64_t, true>(pl, in, out, s)
                             : launch_tiled_v<W, VEC, CONJ, uint64_t, false>(pl, in, out, s);
  return pl.rag ? launch_tiled_v<W, VEC, CONJ, uint32_t, true>(pl, in, out, s)
                : launch_tiled_v<W, VEC, CONJ, uint32_t, false>(pl, in, out, s);
}

template <typename W, bool CONJ>
int launch_tiled_t(const Planned &pl, const void *in, void *out, hipStream_t s)
{
  // only vector widths that keep one access <= 16 bytes are instantiated
  if constexpr (sizeof(W) == 4) {
    if (pl.vec == 4) return launch_tiled_w<W, 4, CONJ>(pl, in, out, s);
    if (pl.vec == 2) return launch_tiled_w<W, 2, CONJ>(pl, in, out, s);
  } else if constexpr (sizeof(W) == 8) {
    if (pl.vec == 2) return launch_tiled_w<W, 2, CONJ>(pl, in, out, s);
  }
  return launch_tiled_w<W, 1, CONJ>(pl, in, out, s);
}

template <typename W, bool CONJ>
int launch_copy_t(const void *in, void *out, uint64_t n, hipStream_t s, int cus)
{
  const uint64_t bytes = n * sizeof(W);
  if (bytes % 16 == 0 && ((uintptr_t)in % 16) == 0 && ((uintptr_t)out % 16) == 0) {
    const uint64_t n16 = bytes / 16;
    const uint64_t blocks = (n16 + kCopyNT - 1) / kCopyNT;
    // one launch covers up to 2^31 - 1 workgroups of 1 KiB (2 TiB): always
    // enough for a body that fits in device memory
    if (blocks >= (1ull << 31)) return TALSH_INVALID_ARGS;
    hipLaunchKernelGGL((k_copy16<W, CONJ>), dim3((unsigned)blocks), dim3(kCopyNT), 0, s,
                       static_cast<const nt4 *>(in), static_cast<nt4 *>(out), n16);
    return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
  }
  const uint64_t blocks = std::min<uint64_t>((n + kNT - 1) / kNT, (uint64_t)cus * 16);
  if (blocks == 0) return 0;
  hipLaunchKernelGGL((k_copy<W, CONJ>), dim3((unsigned)blocks), dim3(kNT), 0, s, static_cast<const W *>(in),
                     static_cast<W *>(out), n);
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

template <typename W, bool CONJ>
int launch_scatter_t(const Fused &f, const void *in, void *out, hipStream_t s, int cus)
{
  ScatterParams h;
  std::memset(&h, 0, sizeof(h));
  h.rank = f.rank;
  h.vol = 1;
  std::vector<int> n2o(f.rank);
  for (int i = 0; i < f.rank; ++i) n2o[f.o2n[i]] = i;
  std::vector<uint64_t> ostr_out(f.rank);
  uint64_t st = 1;
  for (int q = 0; q < f.rank; ++q) { ostr_out[q] = st; st *= (uint64_t)f.ext[n2o[q]]; }
  for (int i = 0; i < f.rank; ++i) {
    h.ext[i] = (uint64_t)f.ext[i];
    h.ostr[i] = ostr_out[f.o2n[i]];
    h.vol *= h.ext[i];
  }
  const uint64_t blocks = std::min<uint64_t>((h.vol + kNT - 1) / kNT, (uint64_t)cus * 16);
  hipLaunchKernelGGL((k_permute_scatter<W, CONJ>), dim3((unsigned)std::max<uint64_t>(blocks, 1)), dim3(kNT), 0, s,
                     static_cast<const W *>(in), static_cast<W *>(out), h);
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

template <typename W, bool CONJ>
int dispatch(const Planned &pl, const void *in, void *out, uint64_t vol, hipStream_t s, int cus)
{
  if (pl.path == 0) return launch_copy_t<W, CONJ>(in, out, vol, s, cus);
  if (pl.path == 1) return launch_tiled_t<W, CONJ>(pl, in, out, s);
  return launch_scatter_t<W, CONJ>(pl.f, in, out, s, cus);
}

}  // namespace

int plan_permute_info(int kind, int rank, const int *dims_in, const int *o2n, PermuteInfo *info, bool with_padding)
{
  const int es = kind_size(kind);
  if (es == 0 || rank < 0 || rank > kMaxRank) return TALSH_INVALID_ARGS;
  Planned pl;
  plan(es, rank, dims_in, o2n, pl, 256, with_padding);
  info->fused_rank = pl.f.rank;
  info->path = pl.path;
  info->tile_volume = pl.path == 1 ? (int)pl.tile.vol : 0;
  info->in_run = pl.path == 1 ? pl.tile.in_run : 0;
  info->out_run = pl.path == 1 ? pl.tile.out_run : 0;
  info->grid = pl.grid;
  info->vec = pl.path == 1 ? pl.vec : 1;
  return 0;
}

int launch_permute(const void *in, void *out, int kind, int rank, const int *dims_in, const int *o2n, bool conj,
                   hipStream_t s)
{
  const int es = kind_size(kind);
  if (es == 0 || rank < 0 || rank > kMaxRank) return TALSH_INVALID_ARGS;
  std::vector<int> seen(rank, 0);
  uint64_t vol = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims_in[i] <= 0 || o2n[i] < 0 || o2n[i] >= rank || seen[o2n[i]]) return TALSH_INVALID_ARGS;
    seen[o2n[i]] = 1;
    vol *= (uint64_t)dims_in[i];
  }
  if (vol == 0) return 0;
  if (in == out) return TALSH_INVALID_ARGS;  // out-of-place only
  static int cus = num_cus_current();
  Planned pl;
  {
    // plans are memoized per (element size, extents, permutation): planning
    // (fusion, tile choice, bank-conflict simulation) costs far more host time
    // than a launch
    static std::mutex mu;
    static std::unordered_map<std::string, Planned> cache;
    std::string key;
    key.reserve(8 * rank + 8);
    key.append(std::to_string(es)).push_back(':');
    for (int i = 0; i < rank; ++i) key.append(std::to_string(dims_in[i])).append(",").append(std::to_string(o2n[i])).push_back(';');
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      pl = it->second;
    } else {
      plan(es, rank, dims_in, o2n, pl, cus);
      if (cache.size() >= 4096) cache.clear();
      cache.emplace(key, pl);
    }
  }
  if (pl.path == 1) {
    // vector accesses need both bodies aligned to the vector size
    if (pl.vec > 1 && (((uintptr_t)in | (uintptr_t)out) % (uintptr_t)(pl.vec * es)) != 0) {
      pl.tp.tv_in0 = to_dev(pl.tp.t_in[0].d);
      pl.tp.tv_out0 = to_dev(pl.tp.t_out[0].d);
      pl.vec = 1;
      pl.ept = 1;
      while ((long)pl.ept * kTNT < (long)pl.tp.vol_tile) pl.ept *= 2;
      if (pl.ept > kMaxEPT) pl.path = 2;  // one element per access: too many per thread for the tiled kernel
    }
    // the persistent grid and per-thread vector slots must cover the tile
    if (pl.path == 1 && ((long)pl.ept * kTNT * pl.vec < (long)pl.tp.vol_tile || pl.grid <= 0 || pl.ept > kMaxEPT))
      return TALSH_FAILURE;
  }
  const bool cj = conj && kind_complex(kind);
  hipEvent_t pr[2];
  prof_begin(s, kKTranspose, pr);  // timing brackets the launch only
  int e = TALSH_INVALID_ARGS;
  switch (kind) {
  case R4: e = dispatch<uint32_t, false>(pl, in, out, vol, s, cus); break;
  case R8: e = dispatch<uint2, false>(pl, in, out, vol, s, cus); break;
  case C4: e = cj ? dispatch<uint2, true>(pl, in, out, vol, s, cus) : dispatch<uint2, false>(pl, in, out, vol, s, cus); break;
  case C8: e = cj ? dispatch<uint4, true>(pl, in, out, vol, s, cus) : dispatch<uint4, false>(pl, in, out, vol, s, cus); break;
  default: break;
  }
  prof_end(s, kKTranspose, pr);
  return e;
}

}  // namespace xt
