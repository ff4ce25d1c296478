// Matrix-core GEMMs for gfx950: the compute step of a tensor contraction
// (D(ll,lr) += alpha * L(lc,ll)^T * R(lc,lr) after matricization).
//
// Replaces the reference's cuBLAS calls and gpu_matrix_multiply_tn__
// (TALSH/tensor_algebra_gpu_nvidia.cu:6852-7056, :2263-2312) with in-house
// kernels:
//   R8 / C8 : v_mfma_f64_16x16x4_f64, FP64 accumulation;
//   R4 / C4 : v_mfma_f32_16x16x4_f32, FP32 accumulation ("native FP32");
//   R4 fast : three-way BF16 split on v_mfma_f32_16x16x32_bf16 (six partial
//             products per pair, two-level FP32 accumulation per K chunk,
//             FP64 sum over the chunks) -- see k_gemm_bf16x3 below.
// Operands: A stored k x m with k fastest ("T") or m x k with m fastest ("N");
// B stored k x n with k fastest ("N") or n x k with n fastest ("T");
// C is m x n column-major. Complex operands are interleaved (re, im) and are
// split into real/imaginary LDS planes while staging.
#include <algorithm>
#include <type_traits>

#include "bf16x3.hpp"
#include "kernels.hpp"

namespace xt {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Acc;
template <>
struct Acc<double> { typedef d4 type; };
template <>
struct Acc<float> { typedef f4 type; };

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma(float a, float b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

struct GemmParams {
  const void *a;
  const void *b;
  void *c;
  int64_t m, n, k;
  int64_t lda, ldb, ldc;     // leading dims in elements (complex: in complex elements)
  double alpha_re, alpha_im;
  int beta_zero;
  int conj_a, conj_b;
  int tiles_m, tiles_n;
  int xcd_remap;             // 1: blockIdx -> tile remap that keeps neighbours on one XCD
  int vec_a, vec_b;          // 16-byte aligned rows of k-fast operands (vector loads allowed)
  // split K (real kinds): workgroup row blockIdx.y takes k in [y * kchunk,
  // (y + 1) * kchunk) and writes its raw m x n partial to part + y * m * n;
  // part == nullptr: one pass over all of K, epilogue into C
  void *part;
  int64_t kchunk;
  // fused index gather (GATHER kernels): A element (r, kk) at a_row[r] +
  // a_kof[kk], B element (c, kk) at b_row[c] + b_kof[kk], C element (r, c) at
  // c_row[r] + c_col[c] (element offsets into the tensor bodies); vec_a /
  // vec_b: aligned groups of 16 bytes along the staging direction are
  // contiguous
  const int64_t *a_row, *a_kof, *b_row, *b_kof, *c_row, *c_col;
  int64_t a_ks0, b_ks0;      // stride of the first contracted dimension in A / B (gather)
  int64_t a_rs0, b_rs0;      // stride of the first row dimension of A / B (gather)
  const void *zpad;          // 64 zero bytes, 16-byte aligned (gather)
  int part_rowmajor;         // split-K partials m x n row-major (else column-major)
};

// Maps a workgroup to an output tile: first gather the workgroups that the
// dispatcher places on one XCD (round-robin over 8) into a contiguous range,
// then walk that range in column groups of G M-tiles so neighbouring
// workgroups share B panels (and A panels within a group) in their XCD's L2.
template <int G = 8>
__device__ __forceinline__ void tile_of(const GemmParams &p, int &tm, int &tn)
{
  const int nb = gridDim.x;
  int t = blockIdx.x;
  if (p.xcd_remap && (nb % 8) == 0) t = (t % 8) * (nb / 8) + t / 8;
  const int per_group = G * p.tiles_n;
  const int group = t / per_group;
  const int first_m = group * G;
  const int gm = min(p.tiles_m - first_m, G);
  const int r = t - group * per_group;
  tm = first_m + r % gm;
  tn = r / gm;
}

// ---------------------------------------------------------------------------
// FP64 / FP32 16x16x4 family. Workgroup: 256 threads = 2x2 waves; each wave
// owns a (BM/2) x (BN/2) sub-tile of 16x16 MFMA blocks.
//
// Operand tiles sit in LDS with k fastest ([row][k], padded rows). Inside a
// 16-deep k slice, lane group g = lane/16 feeds k = 4g+s at MFMA step s, so a
// lane's four steps use four consecutive k and its fragment for the whole
// slice is one contiguous LDS read (summation order over k is immaterial).
template <typename T>
struct Frag4;
template <>
struct Frag4<float> { typedef float4 type; };
template <>
struct Frag4<double> { typedef double4 type; };

// GATHER: 0 matricized operands; 1 fused gather where every thread's staged
// elements stay inside one run of the staging direction's first dimension
// (offsets by stride arithmetic from one table entry) and real operands load
// in vectors; 2 fused gather with a table lookup per staged element or
// vector.
template <typename T, bool CPLX, bool AK, bool BKF, int BM, int BN, int BK, int OCC, int GATHER = 0>
__global__ __launch_bounds__(256, OCC) void k_gemm_mfma(const GemmParams p)
{
  constexpr int NP = CPLX ? 2 : 1;                 // LDS planes (re, im)
  constexpr int KP = BK + (sizeof(T) == 4 ? 4 : 2); // padded k extent of an LDS row
  constexpr int FM = BM / 32, FN = BN / 32;        // MFMA blocks per wave along m / n
  constexpr int LA = BM * BK / 256;                // A elements staged per thread
  constexpr int LB = BN * BK / 256;
  typedef typename Acc<T>::type AccT;
  typedef typename Frag4<T>::type F4;
  __shared__ __align__(16) T As[2][NP][BM][KP];
  __shared__ __align__(16) T Bs[2][NP][BN][KP];

  int tm, tn;
  tile_of(p, tm, tn);
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
  const T *A = static_cast<const T *>(p.a);
  const T *B = static_cast<const T *>(p.b);
  const bool interior_mn = (m0 + BM <= p.m) && (n0 + BN <= p.n);
  const bool split = !CPLX && p.part != nullptr;
  const int64_t kbeg = split ? (int64_t)blockIdx.y * p.kchunk : 0;
  const int64_t kend = split ? min(p.k, kbeg + p.kchunk) : p.k;

  // staging coordinates. k-fast operands: a thread covers LA consecutive k of
  // one row. Row-fast operands: RV consecutive rows (one vector of two real
  // elements; one complex element) at each of KV = L / RV consecutive k, so a
  // wave's loads of one k are contiguous along rows and each row's KV values
  // go to LDS in one store. Registers: element (row + e, k + j) of a row-fast
  // operand is r[.][j * RV + e].
  constexpr int RV = CPLX ? 1 : 2;
  constexpr int KVA = LA / RV, KVB = LB / RV;
  int a_r, a_k, b_r, b_k;
  if (AK) { a_r = tid / (BK / LA); a_k = (tid % (BK / LA)) * LA; }
  else { a_r = (tid % (BM / RV)) * RV; a_k = (tid / (BM / RV)) * KVA; }
  if (BKF) { b_r = tid / (BK / LB); b_k = (tid % (BK / LB)) * LB; }
  else { b_r = (tid % (BN / RV)) * RV; b_k = (tid / (BN / RV)) * KVB; }

  // fused gather: offset of the thread's first staged row (loop invariant)
  int64_t ga_row = 0, gb_row = 0;
  if (GATHER) {
    ga_row = p.a_row[m0 + a_r < p.m ? m0 + a_r : 0];
    gb_row = p.b_row[n0 + b_r < p.n ? n0 + b_r : 0];
  }
  // One operand element (or vector of N elements) of a gather load; lanes out
  // of range (ok == false) read a zero block instead. The load itself is
  // unconditional and its value is used as loaded: a lane-masked load, or a
  // select on the loaded value, makes the compiler wait for the loads in
  // flight right there instead of at the LDS store.
  constexpr int VW = 16 / (int)sizeof(T);
  const T *zpad = static_cast<const T *>(p.zpad);
  auto gather_vec = [&](const T *X, int64_t off, bool ok, T *dst, auto n) {
    constexpr int N = decltype(n)::value;
    typedef T TN __attribute__((ext_vector_type(N)));
    const TN v = *reinterpret_cast<const TN *>(ok ? X + off : zpad);
#pragma unroll
    for (int e = 0; e < N; ++e) dst[e] = v[e];
  };
  auto gather_one = [&](const T *X, int64_t off, bool ok, int conj, T &re, T &im) {
    if (CPLX) {
      const T *q = ok ? X + 2 * off : zpad;
      re = q[0];
      const T b = q[1];
      im = conj ? -b : b;
    } else {
      re = *(ok ? X + off : zpad);
    }
  };
  // Gather staging of one operand tile, kf: along k (row offset fixed), else
  // along rows. kbase is the table entry of the thread's first k (loaded one
  // tile ahead), row_base the offset of its first row. GATHER == 1: a
  // thread's consecutive k (rows) stay inside one run of the first contracted
  // (row) dimension, stride ks0 (rs0) (the launcher checks the extents);
  // GATHER == 2: each k (row) looks up its own table entry. vec: 16-byte (k
  // direction) or RV-element (row direction) groups are contiguous.
  auto gather_load = [&](const T *X, const int64_t *rowtab, int64_t row_base, int64_t rs0, const int64_t *kof,
                         int64_t kbase, int64_t ks0, int64_t rows, int64_t row0,
                         int k_off, int64_t k0, bool vec, int conj, auto kf, auto dst) {
    constexpr int L = (int)(sizeof(dst[0]) / sizeof(T));
    constexpr bool RUNS = GATHER == 1;
    // GATHER == 1 kernels of real kinds always take the vector path (the
    // launcher checks it applies): a run-time choice between vector and
    // scalar loads made the compiler wait for one operand's loads before
    // issuing the other's
    if (RUNS && !CPLX) vec = true;
    const int64_t kk0 = k0 + k_off;
    if (decltype(kf)::value) {
      const bool rok = row0 < rows;
      if (RUNS) {
        if (!CPLX && vec) {
#pragma unroll
          for (int i = 0; i < L; i += VW)
            gather_vec(X, row_base + kbase + i, rok && kk0 + i < kend, &dst[0][i], std::integral_constant<int, VW>());
        } else {
#pragma unroll
          for (int i = 0; i < L; ++i)
            gather_one(X, row_base + kbase + i * ks0, rok && kk0 + i < kend, conj, dst[0][i], dst[1 % NP][i]);
        }
      } else if (!CPLX && vec) {
#pragma unroll
        for (int i = 0; i < L; i += VW) {
          const bool ok = rok && kk0 + i < kend;
          gather_vec(X, row_base + kof[ok ? kk0 + i : 0], ok, &dst[0][i], std::integral_constant<int, VW>());
        }
      } else {
#pragma unroll
        for (int i = 0; i < L; ++i) {
          const bool ok = rok && kk0 + i < kend;
          gather_one(X, row_base + kof[ok ? kk0 + i : 0], ok, conj, dst[0][i], dst[1 % NP][i]);
        }
      }
    } else {
      constexpr int KV = L / RV;
#pragma unroll
      for (int j = 0; j < KV; ++j) {
        const bool kok = kk0 + j < kend;
        const int64_t ko = RUNS ? kbase + j * ks0 : (j == 0 ? kbase : kof[kok ? kk0 + j : 0]);
        if (!CPLX && vec) {
          // rows come in contiguous groups of RV (the row extent is a multiple)
          const int64_t ro = RUNS ? row_base : rowtab[row0 < rows ? row0 : 0];
          gather_vec(X, ro + ko, kok && row0 < rows, &dst[0][j * RV], std::integral_constant<int, RV>());
        } else {
#pragma unroll
          for (int e = 0; e < RV; ++e) {
            const bool ok = kok && row0 + e < rows;
            const int64_t ro = RUNS ? row_base + e * rs0 : rowtab[ok ? row0 + e : 0];
            gather_one(X, ro + ko, ok, conj, dst[0][j * RV + e], dst[1 % NP][j * RV + e]);
          }
        }
      }
    }
  };
  // table entry of a thread's first k in the tile at k0 (clamped into range)
  auto kof_at = [&](const int64_t *kof, int64_t k0, int k_off) -> int64_t {
    const int64_t kk = k0 + k_off;
    return kof[kk < kend ? kk : kbeg];
  };
  int64_t ka_next = 0, kb_next = 0;  // gather: entries of the next tile to load

  // Matricized staging of one operand tile: X stored rows x k with leading
  // dimension ld (k fastest when kf, rows fastest otherwise); vec: 16-byte
  // aligned rows (kf) or columns.
  // force: the caller has checked that the vector form applies.
  auto mat_load = [&](const T *X, int64_t ld, int64_t rows, int64_t row0, int k_off, int64_t k0, bool vec,
                      bool interior, int conj, auto kf, auto dst, auto force) {
    constexpr int L = (int)(sizeof(dst[0]) / sizeof(T));
    constexpr bool F = decltype(force)::value;
    const int64_t kk0 = k0 + k_off;
    if (decltype(kf)::value) {
      if (F || (!CPLX && vec && interior && k0 + BK <= kend)) {
        const T *src = X + row0 * ld + kk0;
#pragma unroll
        for (int i = 0; i < L; i += VW) {
          typedef T TV __attribute__((ext_vector_type(16 / sizeof(T))));
          const TV v = *reinterpret_cast<const TV *>(src + i);
#pragma unroll
          for (int e = 0; e < VW; ++e) dst[0][i + e] = v[e];
        }
        return;
      }
#pragma unroll
      for (int i = 0; i < L; ++i) {
        const int64_t kk = kk0 + i;
        const bool ok = row0 < rows && kk < kend;
        const int64_t off = row0 * ld + kk;
        if (CPLX) {
          dst[0][i] = ok ? X[2 * off] : T(0);
          const T im = ok ? X[2 * off + 1] : T(0);
          dst[1 % NP][i] = conj ? -im : im;
        } else {
          dst[0][i] = ok ? X[off] : T(0);
        }
      }
    } else {
      constexpr int KV = L / RV;
      if (F || (!CPLX && vec && interior && k0 + BK <= kend)) {
        typedef T TR __attribute__((ext_vector_type(RV)));
#pragma unroll
        for (int j = 0; j < KV; ++j) {
          const TR v = *reinterpret_cast<const TR *>(X + (kk0 + j) * ld + row0);
#pragma unroll
          for (int e = 0; e < RV; ++e) dst[0][j * RV + e] = v[e];
        }
        return;
      }
#pragma unroll
      for (int j = 0; j < KV; ++j)
#pragma unroll
        for (int e = 0; e < RV; ++e) {
          const int64_t kk = kk0 + j, r = row0 + e;
          const bool ok = r < rows && kk < kend;
          const int64_t off = kk * ld + r;
          if (CPLX) {
            dst[0][j * RV + e] = ok ? X[2 * off] : T(0);
            const T im = ok ? X[2 * off + 1] : T(0);
            dst[1 % NP][j * RV + e] = conj ? -im : im;
          } else {
            dst[0][j * RV + e] = ok ? X[off] : T(0);
          }
        }
    }
  };

  T ra[NP][LA], rb[NP][LB];
  auto load_a = [&](int64_t k0) {
    if (GATHER)
      gather_load(A, p.a_row, ga_row, p.a_rs0, p.a_kof, ka_next, p.a_ks0, p.m, m0 + a_r, a_k, k0,
                  p.vec_a != 0, p.conj_a, std::integral_constant<bool, AK>(), ra);
    else
      mat_load(A, p.lda, p.m, m0 + a_r, a_k, k0, p.vec_a != 0, interior_mn, p.conj_a,
               std::integral_constant<bool, AK>(), ra, std::false_type());
  };
  auto load_b = [&](int64_t k0) {
    if (GATHER)
      gather_load(B, p.b_row, gb_row, p.b_rs0, p.b_kof, kb_next, p.b_ks0, p.n, n0 + b_r, b_k, k0,
                  p.vec_b != 0, p.conj_b, std::integral_constant<bool, BKF>(), rb);
    else
      mat_load(B, p.ldb, p.n, n0 + b_r, b_k, k0, p.vec_b != 0, interior_mn, p.conj_b,
               std::integral_constant<bool, BKF>(), rb, std::false_type());
  };
  // Both operands of a tile: when both take the vector form, their loads are
  // issued back to back in one branch (separate per-operand choices made the
  // compiler wait for A's loads at the join before issuing B's).
  auto load_ab = [&](int64_t k0) {
    if (!GATHER && !CPLX && p.vec_a && p.vec_b && interior_mn && k0 + BK <= kend) {
      mat_load(A, p.lda, p.m, m0 + a_r, a_k, k0, true, true, 0, std::integral_constant<bool, AK>(), ra,
               std::true_type());
      mat_load(B, p.ldb, p.n, n0 + b_r, b_k, k0, true, true, 0, std::integral_constant<bool, BKF>(), rb,
               std::true_type());
      return;
    }
    load_a(k0);
    load_b(k0);
  };
  // LDS store of one operand's staged tile: k-fast rows in 16-byte pieces;
  // row-fast: each of the RV rows gets its KV consecutive k in one store.
  auto lds_store = [&](auto S, const auto &r, int row, int k, auto kf) {
    constexpr int L = (int)(sizeof(r[0]) / sizeof(T));
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      if (decltype(kf)::value) {
#pragma unroll
        for (int i = 0; i < L; i += VW) {
          typedef T TV __attribute__((ext_vector_type(16 / sizeof(T))));
          TV v;
#pragma unroll
          for (int e = 0; e < VW; ++e) v[e] = r[q][i + e];
          *reinterpret_cast<TV *>(&S[q][row][k + i]) = v;
        }
      } else {
        constexpr int KV = L / RV;
#pragma unroll
        for (int e = 0; e < RV; ++e) {
          if constexpr (KV == 1) {
            S[q][row + e][k] = r[q][e];
          } else {
            typedef T TK __attribute__((ext_vector_type(KV)));
            TK v;
#pragma unroll
            for (int j = 0; j < KV; ++j) v[j] = r[q][j * RV + e];
            *reinterpret_cast<TK *>(&S[q][row + e][k]) = v;
          }
        }
      }
    }
  };
  auto store = [&](int buf) {
    lds_store(As[buf], ra, a_r, a_k, std::integral_constant<bool, AK>());
    lds_store(Bs[buf], rb, b_r, b_k, std::integral_constant<bool, BKF>());
  };

  AccT acc[NP][FM][FN];
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[q][i][j] = AccT{0, 0, 0, 0};

  const int64_t nk = (kend - kbeg + BK - 1) / BK;
  if (GATHER) {
    ka_next = kof_at(p.a_kof, kbeg, a_k);
    kb_next = kof_at(p.b_kof, kbeg, b_k);
  }
  load_ab(kbeg);
  if (GATHER) {  // table entries one tile ahead of the operand loads
    ka_next = kof_at(p.a_kof, kbeg + BK, a_k);
    kb_next = kof_at(p.b_kof, kbeg + BK, b_k);
  }
  store(0);
  __syncthreads();
  const int fr = lane & 15, fg = (lane >> 4) * 4;
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) {
      load_ab(kbeg + (kt + 1) * BK);
      if (GATHER) {
        ka_next = kof_at(p.a_kof, kbeg + (kt + 2) * BK, a_k);
        kb_next = kof_at(p.b_kof, kbeg + (kt + 2) * BK, b_k);
      }
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      F4 fa[NP][FM], fb[NP][FN];
#pragma unroll
      for (int q = 0; q < NP; ++q) {
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[q][i] = *reinterpret_cast<const F4 *>(&As[cur][q][wm + i * 16 + fr][ks + fg]);
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[q][j] = *reinterpret_cast<const F4 *>(&Bs[cur][q][wn + j * 16 + fr][ks + fg]);
      }
#pragma unroll
      for (int st = 0; st < 4; ++st) {
#pragma unroll
        for (int i = 0; i < FM; ++i) {
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            const T a0 = st == 0 ? fa[0][i].x : st == 1 ? fa[0][i].y : st == 2 ? fa[0][i].z : fa[0][i].w;
            const T b0 = st == 0 ? fb[0][j].x : st == 1 ? fb[0][j].y : st == 2 ? fb[0][j].z : fb[0][j].w;
            if (CPLX) {
              const F4 &ai4 = fa[1 % NP][i], &bi4 = fb[1 % NP][j];
              const T a1 = st == 0 ? ai4.x : st == 1 ? ai4.y : st == 2 ? ai4.z : ai4.w;
              const T b1 = st == 0 ? bi4.x : st == 1 ? bi4.y : st == 2 ? bi4.z : bi4.w;
              // (ar + i ai)(br + i bi) = (ar br - ai bi) + i (ar bi + ai br)
              acc[0][i][j] = mfma(a0, b0, acc[0][i][j]);
              acc[0][i][j] = mfma(-a1, b1, acc[0][i][j]);
              acc[1 % NP][i][j] = mfma(a0, b1, acc[1 % NP][i][j]);
              acc[1 % NP][i][j] = mfma(a1, b0, acc[1 % NP][i][j]);
            } else {
              acc[0][i][j] = mfma(a0, b0, acc[0][i][j]);
            }
          }
        }
      }
    }
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }
  // epilogue: C = alpha * acc + beta * C (beta 0 or 1), or the raw partial
  T *C = static_cast<T *>(p.c);
  T *Pz = split ? static_cast<T *>(p.part) + (int64_t)blockIdx.y * p.m * p.n : nullptr;
  const T ar = (T)p.alpha_re, ai = (T)p.alpha_im;
  const int fk = (lane >> 4) & 3;
  int64_t gcol[FN];
  if (GATHER && !split) {
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int64_t col = n0 + wn + j * 16 + fr;
      gcol[j] = p.c_col[col < p.n ? col : 0];
    }
  }
  // Per 16-row block of the wave: the offsets (-1: outside C), then every C
  // load of the block at once (clamped addresses, no lane-masked branches),
  // one wait, then the stores; a load-add-store per element would wait for
  // each load in turn behind the stores already issued.
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    int64_t off[FN][4];
    T cv[FN][4][NP];
    // accumulator layout: FP32 16x16x4 keeps rows 4*(lane/16)+r in a lane,
    // FP64 16x16x4 keeps rows lane/16 + 4*r
    int64_t grow[4];
    if (GATHER && !split) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm + i * 16 + (sizeof(T) == 8 ? fk + 4 * r : fk * 4 + r);
        grow[r] = p.c_row[row < p.m ? row : 0];
      }
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int64_t col = n0 + wn + j * 16 + fr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm + i * 16 + (sizeof(T) == 8 ? fk + 4 * r : fk * 4 + r);
        off[j][r] = (row < p.m && col < p.n)
                        ? (split ? (p.part_rowmajor ? row * p.n + col : row + col * p.m)
                                 : GATHER ? grow[r] + gcol[j] : row + col * p.ldc)
                        : -1;
      }
    }
    if (!split && !p.beta_zero) {
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < NP; ++q) cv[j][r][q] = C[NP * (off[j][r] < 0 ? 0 : off[j][r]) + q];
      wait_all_vmem();  // on every lane, before the lane-masked stores
    } else {
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < NP; ++q) cv[j][r][q] = T(0);
    }
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (off[j][r] < 0) continue;
        if (CPLX) {
          const T xr = acc[0][i][j][r], xi = acc[1 % NP][i][j][r];
          C[2 * off[j][r]] = ar * xr - ai * xi + cv[j][r][0];
          C[2 * off[j][r] + 1] = ar * xi + ai * xr + cv[j][r][1 % NP];
        } else if (split) {
          Pz[off[j][r]] = acc[0][i][j][r];
        } else {
          C[off[j][r]] = ar * acc[0][i][j][r] + cv[j][r][0];
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Three-way BF16 split (helpers in bf16x3.hpp): six bf16 products per k-step
// on the bf16 MFMAs, FP32 accumulation folded every 256 k-values into a
// second FP32 sum (two-level summation), split K over gridDim.y chunks of at
// most kB3Chunk k-values whose partial tiles k_splitk_reduce sums in FP64.
//
// Operand layout ("tile-blocked"): the split pass stores each operand as three
// bf16 planes cut into blocks of RB rows x KB k-values, 1 KB each, laid out
// exactly as the matrix kernel's LDS stage holds them (rows of KB values,
// 16-byte chunks swizzled against LDS bank conflicts). One wave-wide
// global_load_lds_dwordx4 then moves one whole block, 1 KB contiguous in
// memory, straight into LDS. (Row-major planes moved 32 or 64 bytes per row
// per stage: partial 128-byte lines, and the memory path, not the MFMAs, set
// the kernel time.) Blocks of one row band follow each other along k; row
// bands are blk = RB * k_pad elements apart.
constexpr int kB3Chunk = 8192;
constexpr int kB3M = 256, kB3N = 128;   // output tile of one workgroup (8 waves)

// Split pass: FP32 operand -> three bf16 planes in the tile-blocked layout,
// zero-padded to rows_pad x k_pad, so the matrix kernel needs no bounds
// checks. Both forms write whole 1 KB blocks per wave (lane l fills the
// block's 16-byte slot l).
//
// k-fast source (rows of k contiguous): one wave per block, each lane reads
// the 8 k-values of its chunk straight from its row.
template <int RB, int KB>
__global__ __launch_bounds__(256) void k_split3(const float *__restrict__ x, int64_t rows, int64_t k, int64_t ld,
                                                 unsigned short *__restrict__ planes, int64_t rows_pad, int64_t k_pad,
                                                 int64_t blk, int64_t plane)
{
  typedef B3Layout<RB, KB> Lay;
  static_assert(RB * Lay::kChunks == 64, "one chunk per lane");
  const int lane = threadIdx.x & 63;
  const int64_t nkb = k_pad / KB;
  const int64_t nblocks = rows_pad / RB * nkb;
  const int rr = lane / Lay::kChunks;                    // slot l holds row l / 4 at position l % 4
  const int c = Lay::pos(rr, lane % Lay::kChunks);       // (the position map is an involution in c)
  const int64_t wst = (int64_t)gridDim.x * 4;
  // rows that start 16-byte aligned: two 16-byte loads per lane
  const bool vec = ((uintptr_t)x % 16 == 0) && (ld % 4 == 0);
  for (int64_t bi = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); bi < nblocks; bi += wst) {
    const int64_t band = bi / nkb, kb = bi % nkb;
    const int64_t r = band * RB + rr, k0 = kb * KB + c * 8;
    float v[8];
    if (vec && r < rows && k0 + 8 <= k) {
      const float4 *src = reinterpret_cast<const float4 *>(x + r * ld + k0);
      const float4 u0 = src[0], u1 = src[1];
      v[0] = u0.x; v[1] = u0.y; v[2] = u0.z; v[3] = u0.w; v[4] = u1.x; v[5] = u1.y; v[6] = u1.z; v[7] = u1.w;
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (r < rows && k0 + i < k) ? x[r * ld + k0 + i] : 0.f;
    }
    us8 h0, h1, h2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      unsigned short e0, e1, e2;
      split3(v[i], e0, e1, e2);
      h0[i] = e0; h1[i] = e1; h2[i] = e2;
    }
    const int64_t off = band * blk + kb * (RB * KB) + lane * 8;
    *reinterpret_cast<us8 *>(planes + off) = h0;
    *reinterpret_cast<us8 *>(planes + plane + off) = h1;
    *reinterpret_cast<us8 *>(planes + 2 * plane + off) = h2;
  }
}

// Row-fast source (x[kk * ld + r]): a workgroup takes 64 rows x KB k-values
// (four blocks), reads them as KB runs of 64 contiguous floats into LDS, and
// each wave then splits one block out of LDS.
template <int RB, int KB>
__global__ __launch_bounds__(256) void k_split3_rows(const float *__restrict__ x, int64_t rows, int64_t k, int64_t ld,
                                                      unsigned short *__restrict__ planes, int64_t rows_pad,
                                                      int64_t k_pad, int64_t blk, int64_t plane)
{
  typedef B3Layout<RB, KB> Lay;
  static_assert(RB == 16 && KB == 32, "four 16-row blocks per 64-row tile, 8 threads per k-run");
  __shared__ float t[KB][64 + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t nkb = k_pad / KB;
  const int64_t ntiles = rows_pad / 64 * nkb;
  const bool vec = ((uintptr_t)x % 16 == 0) && (ld % 4 == 0);
  const int lk = tid >> 3, lr = (tid & 7) * 8;           // load: k-run lk, rows lr .. lr+7
  const int rr = lane / Lay::kChunks;
  const int c = Lay::pos(rr, lane % Lay::kChunks);
  for (int64_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
    const int64_t rt = ti / nkb, kb = ti % nkb;
    const int64_t r0 = rt * 64 + lr, kk = kb * KB + lk;
    float v[8];
    if (vec && kk < k && r0 + 8 <= rows) {
      const float4 *src = reinterpret_cast<const float4 *>(x + kk * ld + r0);
      const float4 u0 = src[0], u1 = src[1];
      v[0] = u0.x; v[1] = u0.y; v[2] = u0.z; v[3] = u0.w; v[4] = u1.x; v[5] = u1.y; v[6] = u1.z; v[7] = u1.w;
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = (kk < k && r0 + i < rows) ? x[kk * ld + r0 + i] : 0.f;
    }
    __syncthreads();  // the previous tile's reads of t are done
#pragma unroll
    for (int i = 0; i < 8; ++i) t[lk][lr + i] = v[i];
    __syncthreads();
    us8 h0, h1, h2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      unsigned short e0, e1, e2;
      split3(t[c * 8 + i][wave * 16 + rr], e0, e1, e2);
      h0[i] = e0; h1[i] = e1; h2[i] = e2;
    }
    const int64_t off = (rt * 4 + wave) * blk + kb * (RB * KB) + lane * 8;
    *reinterpret_cast<us8 *>(planes + off) = h0;
    *reinterpret_cast<us8 *>(planes + plane + off) = h1;
    *reinterpret_cast<us8 *>(planes + 2 * plane + off) = h2;
  }
}

// Complex operands (C4) under the three-way split run as one real product of
// the realified operands, with the contracted extent doubled (k' = 2k):
//   A'' (2m x k'):  row 2r   = [ Re a(r,:), -Im a(r,:) ]
//                   row 2r+1 = [ Im a(r,:),  Re a(r,:) ]
//   B'' (n x k'):   row j    = [ Re b(j,:),  Im b(j,:) ]
// so that (A'' B''^T)(2r, j) = Re c(r,j) and (2r+1, j) = Im c(r,j): the real
// product's column-major 2m x n output is the complex m x n output,
// interleaved. Conjugation negates the imaginary parts on the way in. The
// pass writes the same tile-blocked planes as k_split3, one wave per 1 KB
// block, reading the complex source element by element (k-fast or row-fast).
template <int RB, int KB, bool IS_A>
__global__ __launch_bounds__(256) void k_split3_cx(const float2 *__restrict__ x, int64_t rows, int64_t k, int64_t ld,
                                                    int kfast, int conj, unsigned short *__restrict__ planes,
                                                    int64_t rows_pad, int64_t k_pad, int64_t blk, int64_t plane)
{
  typedef B3Layout<RB, KB> Lay;
  static_assert(RB * Lay::kChunks == 64, "one chunk per lane");
  const int lane = threadIdx.x & 63;
  const int64_t nkb = k_pad / KB;
  const int64_t nblocks = rows_pad / RB * nkb;
  const int rr = lane / Lay::kChunks;
  const int c = Lay::pos(rr, lane % Lay::kChunks);
  const int64_t wst = (int64_t)gridDim.x * 4;
  const int64_t rrows = IS_A ? 2 * rows : rows;          // realified rows
  const float isign = conj ? -1.f : 1.f;
  for (int64_t bi = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); bi < nblocks; bi += wst) {
    const int64_t band = bi / nkb, kb = bi % nkb;
    const int64_t rp = band * RB + rr, k0 = kb * KB + c * 8;
    const int64_t r = IS_A ? rp >> 1 : rp;
    const int comp = IS_A ? (int)(rp & 1) : 0;
    us8 h0, h1, h2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t kk2 = k0 + i;
      float v = 0.f;
      if (rp < rrows && kk2 < 2 * k) {
        const int half = kk2 >= k ? 1 : 0;
        const int64_t kk = kk2 - half * k;
        const float2 z = x[kfast ? r * ld + kk : kk * ld + r];
        const float zi = isign * z.y;
        // A: (comp, half) = (0,0) re, (0,1) -im, (1,0) im, (1,1) re; B: (0,0) re, (0,1) im
        v = (comp == half) ? z.x : (IS_A && half) ? -zi : zi;
      }
      unsigned short e0, e1, e2;
      split3(v, e0, e1, e2);
      h0[i] = e0; h1[i] = e1; h2[i] = e2;
    }
    const int64_t off = band * blk + kb * (RB * KB) + lane * 8;
    *reinterpret_cast<us8 *>(planes + off) = h0;
    *reinterpret_cast<us8 *>(planes + plane + off) = h1;
    *reinterpret_cast<us8 *>(planes + 2 * plane + off) = h2;
  }
}

struct B3Params {
  const unsigned short *a3;  // 3 tile-blocked planes of A, plane_a elements apart
  const unsigned short *b3;
  float *part;               // split-K partial C tiles, m x n column-major each
  int64_t m, n, k_pad, kchunk;
  int64_t blk;               // elements between row bands (both operands)
  int64_t plane_a, plane_b;
  int tiles_m, tiles_n;
};

// The workgroups walk columns of 4 M-tiles: the 32 workgroups an XCD runs at
// once then cover 4 x 8 tiles, the arrangement that reads the fewest operand
// rows per k-step for 32 tiles.
__device__ __forceinline__ void tile_of_b3(const B3Params &p, int &tm, int &tn)
{
  GemmParams q{};
  q.tiles_m = p.tiles_m;
  q.tiles_n = p.tiles_n;
  q.xcd_remap = 1;
  tile_of<4>(q, tm, tn);
}

// Issue the DMA of one k-stage: NL 1 KB blocks (A planes x row bands, then B),
// spread over the 8 waves; block idx lands at wr + idx * 512.
template <int RB, int KB, int NL>
__device__ __forceinline__ void b3_issue(unsigned short *__restrict__ wr, const unsigned short *__restrict__ ga,
                                         const unsigned short *__restrict__ gb, const B3Params &p, int64_t kb,
                                         int wave, int lane)
{
  constexpr int NA = 3 * (kB3M / RB);
  constexpr int PER = (NL + 7) / 8;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int idx = wave + 8 * u;
    if (NL % 8 != 0 && u == PER - 1 && idx >= NL) break;
    const unsigned short *g;
    if (idx < NA) g = ga + (idx / (kB3M / RB)) * p.plane_a + (idx % (kB3M / RB)) * p.blk;
    else g = gb + ((idx - NA) / (kB3N / RB)) * p.plane_b + ((idx - NA) % (kB3N / RB)) * p.blk;
    lds_dma16(g + kb * (RB * KB) + lane * 8, wr + idx * 512);
  }
}

// Form 16: k-stages of 32, 16x16x32 MFMAs, each wave 64x64 (4x4 blocks),
// LDS double-buffered (2 x 72 KB). (A form with k-stages of 16, 32x32x16
// MFMAs and four 36 KB stages in flight measured 4% slower.)
//
// A stage function both issues the DMA of a later stage into `wr` and
// multiplies the stage in `rd`. The pointers are restrict-qualified because
// they name different stage buffers: that lets the compiler's wait-count pass
// see that the LDS reads of `rd` do not depend on the DMA still in flight
// into other buffers (without it every LDS read waits for all outstanding
// loads and the stages serialise). The dependence the DMA does carry (stage
// kt written before it is read) is enforced by the caller's vmcnt wait and
// barrier.
template <int FORM>
struct B3Form;

template <>
struct B3Form<16> {
  static constexpr int RB = 16, KB = 32, S = 2;
  static constexpr int NL = 3 * (kB3M + kB3N) / RB;   // 72 blocks per stage, 9 per wave
  typedef f4 Acc;
  static constexpr int FB = 4;                        // MFMA blocks per wave side
  __device__ static void stage(const unsigned short *__restrict__ rd, unsigned short *__restrict__ wr, bool issue,
                               bool compute, const unsigned short *__restrict__ ga,
                               const unsigned short *__restrict__ gb, const B3Params &p, int64_t kb, int wave,
                               int lane, int wm, int wn, Acc (&acc)[FB][FB])
  {
    if (issue) b3_issue<RB, KB, NL>(wr, ga, gb, p, kb, wave, lane);
    if (!compute) return;
    typedef B3Layout<RB, KB> Lay;
    const int fr = lane & 15, fc = lane >> 4;
    const int fpos = Lay::pos(fr, fc) * 8;
    const unsigned short *rb = rd + 3 * kB3M * KB;
    us8 fb[3][FB];
#pragma unroll
    for (int pl = 0; pl < 3; ++pl)
#pragma unroll
      for (int j = 0; j < FB; ++j) fb[pl][j] = *reinterpret_cast<const us8 *>(rb + (pl * kB3N + wn + j * 16 + fr) * KB + fpos);
#pragma unroll
    for (int i = 0; i < FB; ++i) {
      us8 fa[3];
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) fa[pl] = *reinterpret_cast<const us8 *>(rd + (pl * kB3M + wm + i * 16 + fr) * KB + fpos);
      // products outer, blocks inner: consecutive MFMAs write different
      // accumulators; smallest terms first
#pragma unroll
      for (int t = 0; t < 6; ++t) {
        const int qa = t == 0 ? 0 : t == 1 ? 1 : t == 2 ? 2 : t == 3 ? 0 : t == 4 ? 1 : 0;
        const int qb = t == 0 ? 2 : t == 1 ? 1 : t == 2 ? 0 : t == 3 ? 1 : t == 4 ? 0 : 0;
#pragma unroll
        for (int j = 0; j < FB; ++j) acc[i][j] = mfma_bf(fa[qa], fb[qb][j], acc[i][j]);
      }
    }
  }
  // item r of lane l: row 4 (l / 16) + r, column l % 16
  __device__ static int row_of(int lane, int r) { return 4 * (lane >> 4) + r; }
  __device__ static int col_of(int lane) { return lane & 15; }
  static constexpr int BS = 16, NR = 4;
};

template <int FORM>
__global__ __launch_bounds__(512, 1) void k_gemm_bf16x3(const B3Params p)
{
  typedef B3Form<FORM> F;
  constexpr int S = F::S, KB = F::KB, FB = F::FB;
  constexpr int kStage = 3 * (kB3M + kB3N) * KB;      // bf16 values per LDS stage (A planes, then B planes)
  constexpr int kFold = 256 / KB;                     // stages per FP32 fold (256 k-values)
  __shared__ __align__(16) unsigned short lds[S * kStage];
  int tm, tn;
  tile_of_b3(p, tm, tn);
  const int64_t m0 = (int64_t)tm * kB3M, n0 = (int64_t)tn * kB3N;
  const int64_t kbeg = (int64_t)blockIdx.y * p.kchunk;
  const int64_t kend = min(kbeg + p.kchunk, p.k_pad);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 3) * 64, wn = (wave >> 2) * 64;
  const unsigned short *ga = p.a3 + (m0 / F::RB) * p.blk, *gb = p.b3 + (n0 / F::RB) * p.blk;
  typename F::Acc acc[FB][FB], sum[FB][FB];
#pragma unroll
  for (int i = 0; i < FB; ++i)
#pragma unroll
    for (int j = 0; j < FB; ++j) {
      acc[i][j] = typename F::Acc{};
      sum[i][j] = typename F::Acc{};
    }
  const int64_t nk = (kend - kbeg) / KB, kb0 = kbeg / KB;
  for (int s = 0; s < S - 1; ++s)
    if (s < nk) F::stage(lds + (S - 1) * kStage, lds + s * kStage, true, false, ga, gb, p, kb0 + s, wave, lane, wm, wn, acc);
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int cur = (int)(kt % S);
    // this wave's DMA of stage kt has landed (the only stage in flight here)
    static_assert(S == 2, "stage count");
    wait_vmcnt<0>();
    stage_barrier();  // everyone's stage kt is in LDS; stage kt-1's buffer is free
    F::stage(lds + cur * kStage, lds + ((cur + S - 1) % S) * kStage, kt + S - 1 < nk, true, ga, gb, p,
             kb0 + kt + S - 1, wave, lane, wm, wn, acc);
    if ((kt + 1) % kFold == 0 || kt + 1 == nk) {
#pragma unroll
      for (int i = 0; i < FB; ++i)
#pragma unroll
        for (int j = 0; j < FB; ++j) {
          sum[i][j] += acc[i][j];
          acc[i][j] = typename F::Acc{};
        }
    }
  }
  float *P = p.part + (int64_t)blockIdx.y * p.m * p.n;
#pragma unroll
  for (int i = 0; i < FB; ++i)
#pragma unroll
    for (int j = 0; j < FB; ++j) {
      const int64_t col = n0 + wn + j * F::BS + F::col_of(lane);
#pragma unroll
      for (int r = 0; r < F::NR; ++r) {
        const int64_t row = m0 + wm + i * F::BS + F::row_of(lane, r);
        if (row < p.m && col < p.n) P[row + col * p.m] = sum[i][j][r];
      }
    }
}

// C = alpha * sum_z part[z] (+ C): FP64 sum over the K chunks. C and the
// partials are all m x n column-major with leading dimension m, so this is a
// 1-D pass over mn values, four per thread when mn allows.
__global__ __launch_bounds__(256) void k_splitk_reduce(const float *__restrict__ part, int nz, int64_t mn,
                                                       float *__restrict__ c, double alpha, int beta_zero)
{
  const int64_t st = (int64_t)gridDim.x * 256;
  const int64_t n4 = (mn % 4 == 0 && ((uintptr_t)c & 15) == 0) ? mn / 4 : 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += st) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int z = 0; z < nz; ++z) {
      const float4 v = reinterpret_cast<const float4 *>(part + (int64_t)z * mn)[i];
      s0 += (double)v.x; s1 += (double)v.y; s2 += (double)v.z; s3 += (double)v.w;
    }
    float4 o;
    if (beta_zero) {
      o = make_float4((float)(alpha * s0), (float)(alpha * s1), (float)(alpha * s2), (float)(alpha * s3));
    } else {
      const float4 cv = reinterpret_cast<const float4 *>(c)[i];
      o = make_float4((float)(alpha * s0 + (double)cv.x), (float)(alpha * s1 + (double)cv.y),
                      (float)(alpha * s2 + (double)cv.z), (float)(alpha * s3 + (double)cv.w));
    }
    reinterpret_cast<float4 *>(c)[i] = o;
  }
  for (int64_t i = n4 * 4 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < mn; i += st) {
    double s = 0.0;
    for (int z = 0; z < nz; ++z) s += (double)part[(int64_t)z * mn + i];
    double v = alpha * s;
    if (!beta_zero) v += (double)c[i];
    c[i] = (float)v;
  }
}

// Complex form (C4 under the three-way split): the partials hold the m x n
// complex result interleaved; C = alpha * sum_z part[z] (+ C) with complex
// alpha, summed in FP64.
__global__ __launch_bounds__(256) void k_splitk_reduce_cx(const float2 *__restrict__ part, int nz, int64_t mn,
                                                          float2 *__restrict__ c, double are, double aim,
                                                          int beta_zero)
{
  const int64_t st = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < mn; i += st) {
    double sr = 0.0, si = 0.0;
    for (int z = 0; z < nz; ++z) {
      const float2 v = part[(int64_t)z * mn + i];
      sr += (double)v.x;
      si += (double)v.y;
    }
    double orr = are * sr - aim * si, oi = are * si + aim * sr;
    if (!beta_zero) {
      const float2 cv = c[i];
      orr += (double)cv.x;
      oi += (double)cv.y;
    }
    c[i] = make_float2((float)orr, (float)oi);
  }
}

// Realified complex operands for the native-precision complex product (C8 on
// FP64, C4 on FP32 matrix cores) of large problems: the same layout as the
// three-way split's complex form (k' = 2k, k fastest),
//   A'' (2m x k'):  row 2r   = [ Re a(r,:), -Im a(r,:) ],  row 2r+1 = [ Im a(r,:), Re a(r,:) ]
//   B'' (n x k'):   row j    = [ Re b(j,:),  Im b(j,:) ]
// so that the real product A'' B''^T, column-major 2m x n, is the complex m x n
// product interleaved. Conjugation negates imaginary parts; a complex alpha
// (are, aim) is applied to A on the way in (the real product then runs with
// alpha 1). One thread per source element.
template <typename T, bool IS_A>
__global__ __launch_bounds__(256) void k_realify(const T *__restrict__ x, int64_t rows, int64_t k, int64_t ld,
                                                 int kfast, int conj, double are, double aim,
                                                 T *__restrict__ out)
{
  const int64_t k2 = 2 * k;
  const int64_t total = rows * k;  // source elements
  const T isign = conj ? T(-1) : T(1);
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
    // i walks the source in its storage order (k fastest or rows fastest),
    // so that the reads are coalesced
    int64_t r, kk;
    if (kfast) { r = i / k; kk = i - r * k; }
    else { kk = i / rows; r = i - kk * rows; }
    const int64_t src = kfast ? r * ld + kk : kk * ld + r;
    T re = x[2 * src], im = isign * x[2 * src + 1];
    if (IS_A) {
      if (aim != 0.0 || are != 1.0) {
        const T nr = (T)are * re - (T)aim * im, ni = (T)are * im + (T)aim * re;
        re = nr;
        im = ni;
      }
      T *r0 = out + (2 * r) * k2, *r1 = r0 + k2;
      r0[kk] = re;
      r0[k + kk] = -im;
      r1[kk] = im;
      r1[k + kk] = re;
    } else {
      T *r0 = out + r * k2;
      r0[kk] = re;
      r0[k + kk] = im;
    }
  }
}

template <typename K>
int launch_grid(K kern, const GemmParams &p, hipStream_t s)
{
  const int nb = p.tiles_m * p.tiles_n;
  const int nz = p.part ? (int)((p.k + p.kchunk - 1) / p.kchunk) : 1;
  hipLaunchKernelGGL(kern, dim3(nb, nz), dim3(256), 0, s, p);
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

// C = alpha * sum_z part[z] (+ C) for split-K partials of a real GEMM, summed
// in FP64; C and the partials are m x n column-major with leading dimension m.
template <typename T>
__global__ __launch_bounds__(256) void k_splitk_sum(const T *__restrict__ part, int nz, int64_t mn,
                                                    T *__restrict__ c, double alpha, int beta_zero)
{
  const int64_t st = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < mn; i += st) {
    double v = 0.0;
    for (int z = 0; z < nz; ++z) v += (double)part[(int64_t)z * mn + i];
    v *= alpha;
    if (!beta_zero) v += (double)c[i];
    c[i] = (T)v;
  }
}

// Split-K sum of a gathered product: the partials are m x n, column-major or
// (rowmajor) row-major; C element (r, c) sits at c_row[r] + c_col[c]. The
// launcher makes the partials row-major when D's stride-1 index leads the
// columns, so both the product's epilogue (16 consecutive columns per row)
// and this sum write along it.
template <typename T>
__global__ __launch_bounds__(256) void k_splitk_sum_gather(const T *__restrict__ part, int nz, int64_t m, int64_t n,
                                                           int rowmajor, const int64_t *__restrict__ c_row,
                                                           const int64_t *__restrict__ c_col, T *__restrict__ c,
                                                           double alpha, int beta_zero)
{
  const int64_t st = (int64_t)gridDim.x * 256, mn = m * n;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < mn; i += st) {
    int64_t row, col;
    if (rowmajor) { row = i / n; col = i - row * n; }
    else { col = i / m; row = i - col * m; }
    const int64_t off = c_row[row] + c_col[col];
    double v = 0.0;
    for (int z = 0; z < nz; ++z) v += (double)part[(int64_t)z * mn + i];
    v *= alpha;
    if (!beta_zero) v += (double)c[off];
    c[off] = (T)v;
  }
}

// Offset tables of a gathered product, back to back: table t holds
// count[t] = prod(ext[t][.]) entries, entry i = sum_j digit_j(i) * str[t][j].
// After the tables, from entry zero0 on, kGatherZeroWords zero entries (the
// zero block of out-of-range gather loads).
constexpr int kGatherZeroWords = 8;
struct GatherTables {
  int64_t *out;
  int64_t start[7];  // first entry of table t; start[6] = total
  int64_t zero0;     // first zero entry (even, so the block is 16-byte aligned)
  int nd[6];
  int ext[6][kGatherMaxDims];
  int64_t str[6][kGatherMaxDims];
};

__global__ __launch_bounds__(256) void k_gather_tables(const GatherTables g)
{
  const int64_t st = (int64_t)gridDim.x * 256;
  if (blockIdx.x == 0 && threadIdx.x < kGatherZeroWords) g.out[g.zero0 + threadIdx.x] = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < g.start[6]; i += st) {
    int t = 0;
    while (t < 5 && i >= g.start[t + 1]) ++t;
    int64_t x = i - g.start[t], off = 0;
    for (int j = 0; j < g.nd[t]; ++j) {
      const int64_t e = g.ext[t][j];
      const int64_t q = x / e;
      off += (x - q * e) * g.str[t][j];
      x = q;
    }
    g.out[i] = off;
  }
}

template <typename T, bool CPLX, int BM, int BN, int BK, int OCC, int GATHER = 0>
int launch_family(GemmParams p, bool ak, bool bk, hipStream_t s)
{
  p.tiles_m = (int)((p.m + BM - 1) / BM);
  p.tiles_n = (int)((p.n + BN - 1) / BN);
  if (ak && bk) return launch_grid(k_gemm_mfma<T, CPLX, true, true, BM, BN, BK, OCC, GATHER>, p, s);
  if (ak) return launch_grid(k_gemm_mfma<T, CPLX, true, false, BM, BN, BK, OCC, GATHER>, p, s);
  if (bk) return launch_grid(k_gemm_mfma<T, CPLX, false, true, BM, BN, BK, OCC, GATHER>, p, s);
  return launch_grid(k_gemm_mfma<T, CPLX, false, false, BM, BN, BK, OCC, GATHER>, p, s);
}

constexpr int kB3Form = 16;               // kernel form of the three-way split

int64_t round_up(int64_t x, int64_t q) { return (x + q - 1) / q * q; }

struct B3Geom {
  int64_t kp, mp, np, blk, plane_a, plane_b, nz;
};

// Geometry of the split planes; a complex product runs realified (2m rows of
// A, contracted extent 2k).
B3Geom b3_geom(const GemmArgs &g)
{
  B3Geom G;
  const int64_t cx = g.kind == C4 ? 2 : 1;
  G.kp = round_up(std::max<int64_t>(cx * (int64_t)g.k, 1), 32);
  G.mp = round_up(cx * (int64_t)g.m, kB3M);
  G.np = round_up((int64_t)g.n, kB3N);
  const int rb = B3Form<kB3Form>::RB;
  G.blk = rb * G.kp;
  G.plane_a = G.mp / rb * G.blk;
  G.plane_b = G.np / rb * G.blk;
  G.nz = (G.kp + kB3Chunk - 1) / kB3Chunk;
  return G;
}

int launch_bf16x3(const GemmArgs &g, void *ws, hipStream_t s)
{
  typedef B3Form<kB3Form> F;
  if (!ws) return TALSH_INVALID_ARGS;
  const B3Geom G = b3_geom(g);
  unsigned short *a3 = static_cast<unsigned short *>(ws);
  unsigned short *b3 = a3 + 3 * G.plane_a;
  auto split = [&](const void *x, int64_t rows, bool kfast, unsigned short *out, int64_t rows_pad, int64_t plane) {
    const float *xf = static_cast<const float *>(x);
    if (kfast) {
      const int64_t nblocks = rows_pad / F::RB * (G.kp / F::KB);  // one wave each, 4 per workgroup
      const int blocks = (int)std::min<int64_t>((nblocks + 3) / 4, 65536);
      hipLaunchKernelGGL((k_split3<F::RB, F::KB>), dim3(blocks), dim3(256), 0, s, xf, rows, (int64_t)g.k,
                         (int64_t)g.k, out, rows_pad, G.kp, G.blk, plane);
    } else {
      const int64_t ntiles = rows_pad / 64 * (G.kp / F::KB);
      const int blocks = (int)std::min<int64_t>(ntiles, 65536);
      hipLaunchKernelGGL((k_split3_rows<F::RB, F::KB>), dim3(blocks), dim3(256), 0, s, xf, rows, (int64_t)g.k, rows,
                         out, rows_pad, G.kp, G.blk, plane);
    }
  };
  if (g.kind == R4) {
    split(g.a, (int64_t)g.m, g.a_kfast, a3, G.mp, G.plane_a);
    split(g.b, (int64_t)g.n, g.b_kfast, b3, G.np, G.plane_b);
  }
  const bool cx = g.kind == C4;
  if (cx) {
    const int64_t nblocks_a = G.mp / F::RB * (G.kp / F::KB), nblocks_b = G.np / F::RB * (G.kp / F::KB);
    const int ba = (int)std::min<int64_t>((nblocks_a + 3) / 4, 65536);
    const int bb = (int)std::min<int64_t>((nblocks_b + 3) / 4, 65536);
    const int64_t lda = g.a_kfast ? (int64_t)g.k : (int64_t)g.m, ldb = g.b_kfast ? (int64_t)g.k : (int64_t)g.n;
    hipLaunchKernelGGL((k_split3_cx<F::RB, F::KB, true>), dim3(ba), dim3(256), 0, s,
                       static_cast<const float2 *>(g.a), (int64_t)g.m, (int64_t)g.k, lda, g.a_kfast ? 1 : 0,
                       g.conj_a ? 1 : 0, a3, G.mp, G.kp, G.blk, G.plane_a);
    hipLaunchKernelGGL((k_split3_cx<F::RB, F::KB, false>), dim3(bb), dim3(256), 0, s,
                       static_cast<const float2 *>(g.b), (int64_t)g.n, (int64_t)g.k, ldb, g.b_kfast ? 1 : 0,
                       g.conj_b ? 1 : 0, b3, G.np, G.kp, G.blk, G.plane_b);
  }
  B3Params p{};
  p.a3 = a3; p.b3 = b3;
  p.part = reinterpret_cast<float *>(b3 + 3 * G.plane_b);
  p.m = (cx ? 2 : 1) * (int64_t)g.m; p.n = (int64_t)g.n; p.k_pad = G.kp;
  p.kchunk = std::min<int64_t>(G.kp, kB3Chunk);
  p.blk = G.blk; p.plane_a = G.plane_a; p.plane_b = G.plane_b;
  p.tiles_m = (int)(G.mp / kB3M);
  p.tiles_n = (int)(G.np / kB3N);
  hipLaunchKernelGGL(k_gemm_bf16x3<kB3Form>, dim3(p.tiles_m * p.tiles_n, (int)G.nz), dim3(512), 0, s, p);
  const int64_t mn = (int64_t)g.m * (int64_t)g.n;
  if (cx) {
    const int rb = (int)std::min<int64_t>((mn + 255) / 256, 8192);
    hipLaunchKernelGGL(k_splitk_reduce_cx, dim3(rb), dim3(256), 0, s, reinterpret_cast<const float2 *>(p.part),
                       (int)G.nz, mn, static_cast<float2 *>(g.c), g.alpha_re, g.alpha_im, g.beta_zero ? 1 : 0);
  } else {
    const int rb = (int)std::min<int64_t>((mn / 4 + 255) / 256 + 1, 8192);
    hipLaunchKernelGGL(k_splitk_reduce, dim3(rb), dim3(256), 0, s, p.part, (int)G.nz, mn, static_cast<float *>(g.c),
                       g.alpha_re, g.beta_zero ? 1 : 0);
  }
  return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
}

}  // namespace

int gemm_default_tile(int kind, R4Mode r4mode, size_t m, size_t n, size_t k);

int current_cus()
{
  static const int cus = []() {
    int dev = 0, c = 256;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
      c = prop.multiProcessorCount;
    return c;
  }();
  return cus;
}

// Split-K of a real FP64/FP32 GEMM on 64x64 tiles: when the output has fewer
// tiles than two resident workgroups per CU, K is cut into parts of >= 256
// k-values (multiples of the 16-deep k-step) so that the grid fills the chip;
// the partials are summed in FP64 by k_splitk_sum. Returns the number of
// parts (1: no split) and the chunk length.
int splitk_parts_cus(const GemmArgs &g, int tile, int cus, int64_t *kchunk)
{
  *kchunk = (int64_t)g.k;
  const bool real = g.kind == R8 || (g.kind == R4 && g.r4mode == R4Mode::kNativeF32);
  if (!real || tile != 2 || g.k < 512) return 1;
  const int64_t tiles = (int64_t)((g.m + 63) / 64) * (int64_t)((g.n + 63) / 64);
  const int64_t want = 2 * (int64_t)cus;
  if (tiles >= want) return 1;
  int64_t nz = std::min<int64_t>({8, want / tiles, (int64_t)g.k / 256});
  if (nz < 2) return 1;
  const int64_t chunk = ((int64_t)g.k + nz - 1) / nz;
  *kchunk = (chunk + 15) / 16 * 16;
  return (int)(((int64_t)g.k + *kchunk - 1) / *kchunk);
}

int splitk_parts(const GemmArgs &g, int tile, int64_t *kchunk) { return splitk_parts_cus(g, tile, current_cus(), kchunk); }

// Large complex products in native precision run realified through the real
// 128x128 kernels (which re-read operands from L2 half as often as the 64x64
// complex kernel), when the real product has enough 128x128 tiles to fill the
// chip twice and K is deep enough to amortise the realify pass.
bool complex_realified(const GemmArgs &g)
{
  if ((g.kind != C8 && g.kind != C4) || gemm_uses_split3(g)) return false;
  const size_t tiles = ((2 * g.m + 127) / 128) * ((g.n + 127) / 128);
  return tiles >= (size_t)(2 * current_cus()) && g.k >= 512;
}

size_t realified_bytes(const GemmArgs &g)
{
  const size_t es = g.kind == C8 ? 8 : 4, k2 = 2 * g.k;
  return ((2 * g.m * k2 * es + 255) / 256 * 256) + g.n * k2 * es + 16;
}

size_t gemm_workspace_bytes(const GemmArgs &g)
{
  if (!gemm_uses_split3(g)) {
    if (g.kind == C8 || g.kind == C4) return complex_realified(g) ? realified_bytes(g) : 0;
    const int tile = g.tile < 0 ? gemm_default_tile(g.kind, g.r4mode, g.m, g.n, g.k) : g.tile;
    int64_t kc;
    const int nz = splitk_parts(g, tile, &kc);
    return nz > 1 ? (size_t)nz * g.m * g.n * (size_t)kind_size(g.kind) + 16 : 0;
  }
  const B3Geom G = b3_geom(g);
  // split planes (bf16) + K-chunk partials (FP32, m x n each; complex: 2m x n)
  const int64_t cx = g.kind == C4 ? 2 : 1;
  return (size_t)(3 * (G.plane_a + G.plane_b)) * sizeof(unsigned short) +
         (size_t)(G.nz * cx * (int64_t)g.m * (int64_t)g.n) * 4 + 16;
}

int launch_gemm(const GemmArgs &g, hipStream_t s) { return launch_gemm_ws(g, nullptr, 0, s); }

// Scratch of talshAmdGemmNoAlloc. In host code this name is a host-side
// stand-in, not a device address: only the pointer that
// gemm_noalloc_scratch returns may be handed to kernels or copies.
static __device__ __align__(256) unsigned char scratch[kGemmNoAllocScratchBytes];

int gemm_noalloc_scratch(void **ptr)
{
  *ptr = nullptr;
  if (!XT_HIP_OK(hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)))) {
    *ptr = nullptr;
    return TALSH_FAILURE;
  }
  return 0;
}

bool gemm_uses_split3(const GemmArgs &g) { return (g.kind == R4 || g.kind == C4) && g.r4mode == R4Mode::kBF16x3; }

int gemm_tile_choices(int kind, R4Mode r4mode);

int gemm_default_tile(int kind, R4Mode r4mode, size_t m, size_t n, size_t k)
{
  if (gemm_tile_choices(kind, r4mode) == 1) return 0;
  const int cus = current_cus();
  // 128x128 tiles pay off only when there are enough of them to fill the
  // chip twice over and K is deep enough to amortise their prologue;
  // otherwise 64x64 tiles balance better (measured over random contractions)
  const size_t tiles128 = ((m + 127) / 128) * ((n + 127) / 128);
  return (tiles128 >= (size_t)(2 * cus) && k >= 2048) ? 0 : 2;
}

int gemm_splitk_count(int kind, R4Mode r4mode, size_t m, size_t n, size_t k, int tile, int cus)
{
  GemmArgs g{};
  g.kind = kind; g.r4mode = r4mode; g.m = m; g.n = n; g.k = k;
  int64_t kc;
  return splitk_parts_cus(g, tile, cus, &kc);
}

int gemm_tile_choices(int kind, R4Mode r4mode)
{
  if (kind == R8 || (kind == R4 && r4mode == R4Mode::kNativeF32)) return 3;
  return 1;
}

void gemm_tile_shape(int kind, R4Mode r4mode, int tile, int *bm, int *bn, int *occupancy)
{
  *bm = *bn = 128;
  *occupancy = 2;
  if (kind == C8 || kind == C4) { *bm = *bn = 64; return; }
  if (kind == R4 && r4mode == R4Mode::kBF16x3) { *occupancy = 1; return; }
  if (tile == 1) *occupancy = 1;
  if (tile == 2) *bm = *bn = 64;
}

int launch_gemm_ws(const GemmArgs &g, void *ws, size_t ws_bytes, hipStream_t s)
{
  if (g.m == 0 || g.n == 0) return 0;
  if (g.k == 0 && !g.beta_zero) return 0;  // empty contraction leaves C unchanged
  if (ws && ws_bytes < gemm_workspace_bytes(g)) {
    if (gemm_uses_split3(g)) return TALSH_INVALID_ARGS;
    ws = nullptr;  // too small for split-K partials: unsplit
  }
  GemmParams p{};
  p.a = g.a; p.b = g.b; p.c = g.c;
  p.m = (int64_t)g.m; p.n = (int64_t)g.n; p.k = (int64_t)g.k;
  p.lda = g.a_kfast ? p.k : p.m;
  p.ldb = g.b_kfast ? p.k : p.n;
  p.ldc = p.m;
  p.alpha_re = g.alpha_re; p.alpha_im = g.alpha_im;
  p.beta_zero = g.beta_zero ? 1 : 0;
  p.conj_a = g.conj_a; p.conj_b = g.conj_b;
  p.xcd_remap = 1;
  {
    const int es = kind_size(g.kind);
    p.vec_a = (((uintptr_t)g.a % 16) == 0) && ((p.lda * es) % 16 == 0);
    p.vec_b = (((uintptr_t)g.b % 16) == 0) && ((p.ldb * es) % 16 == 0);
  }
  if ((p.m + 63) / 64 * ((p.n + 63) / 64) >= (1ll << 31)) return TALSH_LIMIT_EXCEEDED;
  const int tile = g.tile < 0 ? gemm_default_tile(g.kind, g.r4mode, g.m, g.n, g.k) : g.tile;
  {
    int64_t kc;
    const int nz = ws ? splitk_parts(g, tile, &kc) : 1;
    if (nz > 1) {
      p.part = ws;
      p.kchunk = kc;
      const int64_t mn = (int64_t)g.m * (int64_t)g.n;
      const unsigned rb = (unsigned)std::min<int64_t>((mn + 255) / 256, 4096);
      int e;
      if (g.kind == R8) {
        e = launch_family<double, false, 64, 64, 16, 2>(p, g.a_kfast, g.b_kfast, s);
        if (!e) hipLaunchKernelGGL(k_splitk_sum<double>, dim3(rb), dim3(256), 0, s, static_cast<const double *>(ws), nz,
                                   mn, static_cast<double *>(g.c), g.alpha_re, p.beta_zero);
      } else {
        e = launch_family<float, false, 64, 64, 16, 2>(p, g.a_kfast, g.b_kfast, s);
        if (!e) hipLaunchKernelGGL(k_splitk_sum<float>, dim3(rb), dim3(256), 0, s, static_cast<const float *>(ws), nz,
                                   mn, static_cast<float *>(g.c), g.alpha_re, p.beta_zero);
      }
      if (e) return e;
      return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
    }
  }
  if (ws && complex_realified(g)) {
    const bool dp = g.kind == C8;
    const size_t es = dp ? 8 : 4;
    const int64_t k2 = 2 * (int64_t)g.k;
    char *a2 = static_cast<char *>(ws);
    char *b2 = a2 + (2 * g.m * (size_t)k2 * es + 255) / 256 * 256;
    const bool real_alpha = g.alpha_im == 0.0;
    const double are = real_alpha ? 1.0 : g.alpha_re, aim = real_alpha ? 0.0 : g.alpha_im;
    const int64_t lda = g.a_kfast ? (int64_t)g.k : (int64_t)g.m, ldb = g.b_kfast ? (int64_t)g.k : (int64_t)g.n;
    const int ga = (int)std::min<int64_t>(((int64_t)g.m * (int64_t)g.k + 255) / 256, 65536);
    const int gb = (int)std::min<int64_t>(((int64_t)g.n * (int64_t)g.k + 255) / 256, 65536);
    GemmParams q = p;
    q.m = 2 * (int64_t)g.m;
    q.n = (int64_t)g.n;
    q.k = k2;
    q.lda = q.ldb = k2;
    q.ldc = q.m;
    q.a = a2;
    q.b = b2;
    q.conj_a = q.conj_b = 0;
    q.alpha_re = real_alpha ? g.alpha_re : 1.0;
    q.alpha_im = 0.0;
    q.vec_a = q.vec_b = ((k2 * (int64_t)es) % 16) == 0;
    if (dp) {
      hipLaunchKernelGGL((k_realify<double, true>), dim3(ga), dim3(256), 0, s, static_cast<const double *>(g.a),
                         (int64_t)g.m, (int64_t)g.k, lda, g.a_kfast ? 1 : 0, g.conj_a ? 1 : 0, are, aim,
                         reinterpret_cast<double *>(a2));
      hipLaunchKernelGGL((k_realify<double, false>), dim3(gb), dim3(256), 0, s, static_cast<const double *>(g.b),
                         (int64_t)g.n, (int64_t)g.k, ldb, g.b_kfast ? 1 : 0, g.conj_b ? 1 : 0, 1.0, 0.0,
                         reinterpret_cast<double *>(b2));
      return launch_family<double, false, 128, 128, 16, 2>(q, true, true, s);
    }
    hipLaunchKernelGGL((k_realify<float, true>), dim3(ga), dim3(256), 0, s, static_cast<const float *>(g.a),
                       (int64_t)g.m, (int64_t)g.k, lda, g.a_kfast ? 1 : 0, g.conj_a ? 1 : 0, are, aim,
                       reinterpret_cast<float *>(a2));
    hipLaunchKernelGGL((k_realify<float, false>), dim3(gb), dim3(256), 0, s, static_cast<const float *>(g.b),
                       (int64_t)g.n, (int64_t)g.k, ldb, g.b_kfast ? 1 : 0, g.conj_b ? 1 : 0, 1.0, 0.0,
                       reinterpret_cast<float *>(b2));
    return launch_family<float, false, 128, 128, 16, 2>(q, true, true, s);
  }
  switch (g.kind) {
  case R8:
    switch (tile) {
    case 1: return launch_family<double, false, 128, 128, 16, 1>(p, g.a_kfast, g.b_kfast, s);
    case 2: return launch_family<double, false, 64, 64, 16, 2>(p, g.a_kfast, g.b_kfast, s);
    default: return launch_family<double, false, 128, 128, 16, 2>(p, g.a_kfast, g.b_kfast, s);
    }
  case C8: return launch_family<double, true, 64, 64, 16, 2>(p, g.a_kfast, g.b_kfast, s);
  case R4:
    if (g.r4mode == R4Mode::kBF16x3) return launch_bf16x3(g, ws, s);
    switch (tile) {
    case 1: return launch_family<float, false, 128, 128, 16, 1>(p, g.a_kfast, g.b_kfast, s);
    case 2: return launch_family<float, false, 64, 64, 16, 2>(p, g.a_kfast, g.b_kfast, s);
    default: return launch_family<float, false, 128, 128, 16, 2>(p, g.a_kfast, g.b_kfast, s);
    }
  case C4:
    if (g.r4mode == R4Mode::kBF16x3) return launch_bf16x3(g, ws, s);
    return launch_family<float, true, 64, 64, 16, 2>(p, g.a_kfast, g.b_kfast, s);
  default: return TALSH_INVALID_ARGS;
  }
}

namespace {

int64_t group_count(const GatherGroup &g)
{
  int64_t v = 1;
  for (int j = 0; j < g.nd; ++j) v *= g.ext[j];
  return v;
}

// 16-byte groups along a staging direction are contiguous and aligned when the
// group's first dimension is the tensor's stride-1 dimension with an extent
// that is a multiple of the vector width (every other stride of the tensor is
// then a multiple of that extent) and the body is 16-byte aligned.
bool gather_vec_ok(const GatherGroup &g, int es, const void *base)
{
  return g.nd > 0 && g.str[0] == 1 && (g.ext[0] * es) % 16 == 0 && ((uintptr_t)base % 16) == 0;
}

size_t gather_tables_bytes(const GemmArgs &g)
{
  return ((2 * (g.m + g.n + g.k) + 2 + kGatherZeroWords) * sizeof(int64_t) + 255) / 256 * 256;
}

}  // namespace

size_t gemm_gather_workspace_bytes(const GemmArgs &g)
{
  if (gemm_uses_split3(g)) return 0;
  const size_t tb = gather_tables_bytes(g);
  if (g.kind == C8 || g.kind == C4) return tb;
  const int tile = g.tile < 0 ? gemm_default_tile(g.kind, g.r4mode, g.m, g.n, g.k) : g.tile;
  int64_t kc;
  const int nz = splitk_parts(g, tile, &kc);
  return tb + (nz > 1 ? (size_t)nz * g.m * g.n * (size_t)kind_size(g.kind) + 16 : 0);
}

int launch_gemm_gather(const GemmArgs &g, const GatherMaps &mp, void *ws, size_t ws_bytes, hipStream_t s)
{
  if (gemm_uses_split3(g)) return TALSH_NOT_IMPLEMENTED;
  if (g.m == 0 || g.n == 0) return 0;
  if (g.k == 0 && !g.beta_zero) return 0;  // empty contraction leaves C unchanged
  const GatherGroup *grp[6] = {&mp.a_row, &mp.a_k, &mp.b_row, &mp.b_k, &mp.c_row, &mp.c_col};
  const int64_t want[6] = {(int64_t)g.m, (int64_t)g.k, (int64_t)g.n, (int64_t)g.k, (int64_t)g.m, (int64_t)g.n};
  GatherTables T{};
  T.start[0] = 0;
  for (int t = 0; t < 6; ++t) {
    if (grp[t]->nd < 0 || grp[t]->nd > kGatherMaxDims || group_count(*grp[t]) != want[t]) return TALSH_INVALID_ARGS;
    T.nd[t] = grp[t]->nd;
    for (int j = 0; j < grp[t]->nd; ++j) {
      if (grp[t]->ext[j] < 1 || grp[t]->str[j] < 0) return TALSH_INVALID_ARGS;
      T.ext[t][j] = grp[t]->ext[j];
      T.str[t][j] = grp[t]->str[j];
    }
    T.start[t + 1] = T.start[t] + want[t];
  }
  const size_t tb = gather_tables_bytes(g);
  if (!ws || ws_bytes < tb) return TALSH_INVALID_ARGS;
  if ((int64_t)((g.m + 63) / 64) * (int64_t)((g.n + 63) / 64) >= (1ll << 31)) return TALSH_LIMIT_EXCEEDED;
  int64_t *tab = static_cast<int64_t *>(ws);
  T.out = tab;
  T.zero0 = (T.start[6] + 1) / 2 * 2;
  hipLaunchKernelGGL(k_gather_tables, dim3((unsigned)std::min<int64_t>((T.start[6] + 255) / 256, 1024)), dim3(256), 0,
                     s, T);
  if (!XT_HIP_OK(hipGetLastError())) return TALSH_FAILURE;

  const int es = kind_size(g.kind);
  const bool cplx = g.kind == C8 || g.kind == C4;
  GemmParams p{};
  p.a = g.a; p.b = g.b; p.c = g.c;
  p.m = (int64_t)g.m; p.n = (int64_t)g.n; p.k = (int64_t)g.k;
  p.alpha_re = g.alpha_re; p.alpha_im = g.alpha_im;
  p.beta_zero = g.beta_zero ? 1 : 0;
  p.conj_a = g.conj_a; p.conj_b = g.conj_b;
  p.xcd_remap = 1;
  p.vec_a = !cplx && gather_vec_ok(g.a_kfast ? mp.a_k : mp.a_row, es, g.a);
  p.vec_b = !cplx && gather_vec_ok(g.b_kfast ? mp.b_k : mp.b_row, es, g.b);
  p.a_row = tab + T.start[0];
  p.a_kof = tab + T.start[1];
  p.b_row = tab + T.start[2];
  p.b_kof = tab + T.start[3];
  p.c_row = tab + T.start[4];
  p.c_col = tab + T.start[5];
  p.zpad = tab + T.zero0;
  p.a_ks0 = mp.a_k.nd ? mp.a_k.str[0] : 0;
  p.b_ks0 = mp.b_k.nd ? mp.b_k.str[0] : 0;
  p.a_rs0 = mp.a_row.nd ? mp.a_row.str[0] : 0;
  p.b_rs0 = mp.b_row.nd ? mp.b_row.str[0] : 0;
  // every thread's staged elements (L = 8 on 128x128 tiles, 4 on 64x64) stay
  // inside runs of the first dimensions: L consecutive k of a k-fast operand;
  // RV consecutive rows (2 real, 1 complex) at L / RV consecutive k of a
  // row-fast one
  auto runs = [&](int L) {
    const int rv = cplx ? 1 : 2;
    auto e0 = [](const GatherGroup &q) { return q.nd ? q.ext[0] : 1; };
    auto ok = [&](bool kf, const GatherGroup &row, const GatherGroup &k) {
      return kf ? e0(k) % L == 0 : (e0(row) % rv == 0 && e0(k) % (L / rv) == 0);
    };
    return ok(g.a_kfast, mp.a_row, mp.a_k) && ok(g.b_kfast, mp.b_row, mp.b_k) && (cplx || (p.vec_a && p.vec_b));
  };
  const int tile = g.tile < 0 ? gemm_default_tile(g.kind, g.r4mode, g.m, g.n, g.k) : g.tile;
  if (!cplx) {
    int64_t kc;
    const int nz = splitk_parts(g, tile, &kc);
    const int64_t mn = (int64_t)g.m * (int64_t)g.n;
    if (nz > 1 && ws_bytes >= tb + (size_t)nz * (size_t)mn * es) {
      p.part = static_cast<char *>(ws) + tb;
      p.kchunk = kc;
      p.part_rowmajor = mp.c_col.nd > 0 && mp.c_col.str[0] == 1;
      const unsigned rb = (unsigned)std::min<int64_t>((mn + 255) / 256, 4096);
      int e;
      if (g.kind == R8) {
        e = runs(4) ? launch_family<double, false, 64, 64, 16, 2, 1>(p, g.a_kfast, g.b_kfast, s)
                    : launch_family<double, false, 64, 64, 16, 2, 2>(p, g.a_kfast, g.b_kfast, s);
        if (!e)
          hipLaunchKernelGGL(k_splitk_sum_gather<double>, dim3(rb), dim3(256), 0, s,
                             static_cast<const double *>(p.part), nz, p.m, p.n, p.part_rowmajor, p.c_row, p.c_col,
                             static_cast<double *>(g.c), g.alpha_re, p.beta_zero);
      } else {
        e = runs(4) ? launch_family<float, false, 64, 64, 16, 2, 1>(p, g.a_kfast, g.b_kfast, s)
                    : launch_family<float, false, 64, 64, 16, 2, 2>(p, g.a_kfast, g.b_kfast, s);
        if (!e)
          hipLaunchKernelGGL(k_splitk_sum_gather<float>, dim3(rb), dim3(256), 0, s,
                             static_cast<const float *>(p.part), nz, p.m, p.n, p.part_rowmajor, p.c_row, p.c_col,
                             static_cast<float *>(g.c), g.alpha_re, p.beta_zero);
      }
      if (e) return e;
      return XT_HIP_OK(hipGetLastError()) ? 0 : TALSH_FAILURE;
    }
  }
#define XT_GATHER(T_, CPLX_, BM_, OCC_)                                                          \
  (runs(BM_ / 16) ? launch_family<T_, CPLX_, BM_, BM_, 16, OCC_, 1>(p, g.a_kfast, g.b_kfast, s) \
                  : launch_family<T_, CPLX_, BM_, BM_, 16, OCC_, 2>(p, g.a_kfast, g.b_kfast, s))
  switch (g.kind) {
  case R8:
    switch (tile) {
    case 1: return XT_GATHER(double, false, 128, 1);
    case 2: return XT_GATHER(double, false, 64, 2);
    default: return XT_GATHER(double, false, 128, 2);
    }
  case C8: return XT_GATHER(double, true, 64, 2);
  case R4:
    switch (tile) {
    case 1: return XT_GATHER(float, false, 128, 1);
    case 2: return XT_GATHER(float, false, 64, 2);
    default: return XT_GATHER(float, false, 128, 2);
    }
  case C4: return XT_GATHER(float, true, 64, 2);
  default: return TALSH_INVALID_ARGS;
  }
#undef XT_GATHER
}

}  // namespace xt
