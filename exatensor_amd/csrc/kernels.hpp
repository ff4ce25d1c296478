// Launchers of the engine's gfx950 kernels. All launches are asynchronous on
// the given stream; every launcher checks its operand shapes against what the
// kernel and grid assume before launching and returns 0 or a TAL-SH code.
#pragma once

#include "common.hpp"
#include "runtime.hpp"

namespace xt {

// ---- DLF permutation (k_transpose.hip) ------------------------------------
// out[idx'] = in[idx] where input dimension i becomes output dimension o2n[i]
// (0-based); conj negates imaginary parts of complex kinds. Bit-exact.
int launch_permute(const void *in, void *out, int kind, int rank, const int *dims_in, const int *o2n, bool conj,
                   hipStream_t s);

// Transpose plan summary for diagnostics and the plan selector.
struct PermuteInfo {
  int fused_rank;       // rank after dropping unit extents and fusing runs
  int tile_volume;      // elements per LDS tile (0: direct copy path)
  int in_run, out_run;  // contiguous run lengths (elements) on each side
  int grid;             // workgroups launched
  int path;             // 0 copy, 1 tiled, 2 scatter fallback
  int vec;              // elements per global access
};
// with_padding = false skips the LDS padding search (the fields above do not
// depend on it).
int plan_permute_info(int kind, int rank, const int *dims_in, const int *o2n, PermuteInfo *info,
                      bool with_padding = true);

// ---- element-wise (k_elementwise.hip) -------------------------------------
int launch_fill(void *d, int kind, size_t vol, double re, double im, hipStream_t s);
int launch_scale(void *d, int kind, size_t vol, double re, double im, hipStream_t s);
// d = (accumulate ? d : 0) + f * (conj ? conj(l) : l)
int launch_axpy(void *d, const void *l, int kind, size_t vol, double re, double im, bool conj, bool accumulate,
                hipStream_t s);
// d[0] = (accumulate ? d[0] : 0) + alpha * sum_i l[i]*r[i] (conj flags per operand)
int launch_dot(void *d, const void *l, const void *r, int kind, size_t n, double are, double aim, bool lconj,
               bool rconj, bool accumulate, void *scratch, size_t scratch_bytes, hipStream_t s);

// out[dbase + sum_i x_i dstr_i] (+)= in[sbase + sum_i x_i sstr_i] for every x
// in the box 0 <= x_i < box[i] (tensor slice / insert; element strides).
int launch_box(const void *in, void *out, int kind, int rank, const int *box, const int64_t *sstr,
               const int64_t *dstr, int64_t sbase, int64_t dbase, bool acc, hipStream_t s);

// ---- matrix products (k_gemm.hip) -----------------------------------------
// C(m x n, column-major, ldc = m) = beta*C + alpha * op(A) * op(B)
//   a_kfast: A stored k x m with k fastest (else m x k with m fastest)
//   b_kfast: B stored k x n with k fastest (else n x k with n fastest)
// conj_a / conj_b conjugate complex operands; beta is 0 or 1.
struct GemmArgs {
  int kind;
  R4Mode r4mode;
  size_t m, n, k;
  const void *a;
  const void *b;
  void *c;
  bool a_kfast, b_kfast, conj_a, conj_b, beta_zero;
  double alpha_re, alpha_im;
  int tile;  // tile-shape choice (-1: default)
};
int launch_gemm(const GemmArgs &g, hipStream_t s);
// R4 and C4 products under fast math run as the three-way BF16 split (C4
// realified: one real product over 2m x 2k and n x 2k operands).
bool gemm_uses_split3(const GemmArgs &g);
// Bytes of scratch needed by launch_gemm_ws: BF16x3 split planes and K-chunk
// partials, or the partials of a split-K real product; 0 otherwise.
size_t gemm_workspace_bytes(const GemmArgs &g);
// ws holds ws_bytes; a workspace smaller than gemm_workspace_bytes(g) is an
// error for the three-way split and runs a real product unsplit.
int launch_gemm_ws(const GemmArgs &g, void *ws, size_t ws_bytes, hipStream_t s);
// Fused index gather: a contraction's product read straight from the
// tensors, no transposes. Each matrix index (rows and k of A and of B, rows
// and columns of C) is a group of tensor dimensions; element i of a group
// sits at sum_j digit_j(i) * str[j] with the digits of i over ext[] (first
// fastest). GemmArgs::a, b and c are the tensor bodies; a_kfast / b_kfast
// name the staging direction of A and B (along k, or along rows), the one
// that holds the tensor's stride-1 dimension. Native precision only (not the
// three-way split).
constexpr int kGatherMaxDims = 16;
struct GatherGroup {
  int nd = 0;
  int ext[kGatherMaxDims] = {};
  int64_t str[kGatherMaxDims] = {};
};
struct GatherMaps {
  GatherGroup a_row, a_k, b_row, b_k, c_row, c_col;
};
// Workspace of launch_gemm_gather: the six offset tables (8 bytes per index)
// and, for split K, the partial products.
size_t gemm_gather_workspace_bytes(const GemmArgs &g);
int launch_gemm_gather(const GemmArgs &g, const GatherMaps &maps, void *ws, size_t ws_bytes, hipStream_t s);
// Fixed per-GPU scratch of the allocation-free product entry point: a
// file-scope device array in k_gemm.hip (64 MiB, held by every GPU the
// library runs on until the process ends). gemm_noalloc_scratch returns its
// address on the current device (hipGetSymbolAddress; call once per device,
// with that device current, and keep the pointer per device).
constexpr size_t kGemmNoAllocScratchBytes = size_t(64) << 20;
int gemm_noalloc_scratch(void **ptr);
// Number of tile-shape choices of launch_gemm for a kind, and the output tile
// (bm x bn) and occupancy target of a choice.
int gemm_tile_choices(int kind, R4Mode r4mode);
void gemm_tile_shape(int kind, R4Mode r4mode, int tile, int *bm, int *bn, int *occupancy);
// Number of K parts of a real product in native precision on tile choice
// `tile` on a GPU with cus compute units (1: no split K).
int gemm_splitk_count(int kind, R4Mode r4mode, size_t m, size_t n, size_t k, int tile, int cus);
// Tile choice of launch_gemm when GemmArgs::tile < 0.
int gemm_default_tile(int kind, R4Mode r4mode, size_t m, size_t n, size_t k);

// Truncated SVD of a column-major lu x ru matrix a (tensor kind) in FP64:
// a ~= Lmat (lu x nv) * diag(svals) * Rmat (nv x ru), the nv largest singular
// values in descending order, Rmat holding the conjugate-transposed right
// singular vectors; absorb: 0 none, 1 into Lmat, 2 into Rmat, 3 square roots
// into both. Blocking on stream s (it synchronizes once per Jacobi sweep).
// NUM_INSTABILITY if max_sweeps sweeps do not converge.
struct SvdArgs {
  int kind;
  int64_t lu, ru, nv;
  const void *a;
  void *lmat, *rmat, *svals;
  void *work;          // svd_workspace_bytes(kind, lu, ru)
  int absorb;
  int max_sweeps;
  double *sigma_out;   // optional host copy of the min(lu, ru) singular values
};
int launch_svd(const SvdArgs &g, hipStream_t s);
size_t svd_workspace_bytes(int kind, int64_t lu, int64_t ru);

}  // namespace xt
