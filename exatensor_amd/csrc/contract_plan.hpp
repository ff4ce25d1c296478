// Contraction planning: matricization and fused-gather candidates of a
// GEMM-shaped tensor contraction, their features, and the learned plan
// selector.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace xt {

// How the operands of D += L * R are matricized for the GEMM
// D(m x n) = Lmat(k x m)^T * Rmat(k x n) (or its transpose when swap).
struct ContractPlan {
  int kase = 0;                          // 0 scalar product, 1 scaled add, 2 full, 3 GEMM
  bool swap = false;                     // GEMM computes D^T = R^T L^T
  std::vector<int> l_o2n, r_o2n;         // transposes of L / R (empty: none)
  std::vector<int> d_in_o2n, d_out_o2n;  // D -> Dmat and back (empty: none)
  std::vector<int> lmat_dims, rmat_dims, dmat_dims;
  bool a_kfast = true, b_kfast = true;
  size_t m = 1, n = 1, k = 1;
  double bytes_moved = 0;
  // fused index gather: no transposes, the product reads L and R and updates
  // D in place through offset tables (maps in the launched orientation)
  bool fused = false;
  GatherMaps maps;
};

// Plan variants from kFusedVariantBase on are the fused-gather candidates
// (kFusedVariantBase + i: candidate i of plan_fused).
constexpr int kFusedVariantBase = 1000;

// Fused-gather candidates of a GEMM-shaped contraction: contracted indices in
// L's or in R's order, D or D^T as the product's result; none when an index
// group has more than kGatherMaxDims dimensions after unit extents are
// dropped and contiguous runs merged.
int plan_fused(const std::vector<int> &dig, int dr, int lr, int rr, const std::vector<int> &dd,
               const std::vector<int> &ld, const std::vector<int> &rd, std::vector<ContractPlan> &out);

std::vector<int> reorder(const std::vector<int> &from, const std::vector<int> &to);
bool is_identity(const std::vector<int> &p);

// Distinct matricization candidates of a GEMM-shaped contraction (digital
// pattern dig, extents dd/ld/rd), ranked by the byte-count heuristic; P is
// candidate `variant` (the heuristic's choice for variant < 0 or out of
// range; variant >= kFusedVariantBase picks a fused-gather candidate). Hyper-
// indices (in L, R and D) are rejected with TALSH_INVALID_ARGS.
int plan_contraction(const std::vector<int> &dig, int dr, int lr, int rr, const std::vector<int> &dd,
                     const std::vector<int> &ld, const std::vector<int> &rd, int es, bool acc, int variant,
                     ContractPlan &P, std::vector<ContractPlan> *all);

// ---- learned plan selector ---------------------------------------------------
constexpr int kPlanFeatures = 54;

// One (matricization, GEMM tile) candidate and its feature vector.
struct PlanCandidate {
  int variant = 0, tile = 0;
  float f[kPlanFeatures] = {};
};

// Every candidate of a GEMM-shaped contraction with its features.
// The matricization candidates, then the fused-gather candidates (native
// precision only: not for R4/C4 under the three-way split).
int plan_candidates(const std::vector<int> &dig, int dr, int lr, int rr, const std::vector<int> &dd,
                    const std::vector<int> &ld, const std::vector<int> &rd, int kind, R4Mode r4mode, bool acc,
                    int num_cus, std::vector<PlanCandidate> &out);

// Training domain of the selector (random contractions up to 2^35 flops and
// operands of up to 2^26 elements, with margin); larger problems use the
// heuristic.
constexpr double kSelectorMaxLog2Flops = 35.5;
constexpr double kSelectorMaxLog2Volume = 26.5;

// Memoization key of a plan choice: everything the candidates and their
// timings depend on (pattern, extents, data kind, precision mode,
// accumulation, compute units).
std::string plan_key(const std::vector<int> &dig, int lr, int rr, const std::vector<int> &dd,
                     const std::vector<int> &ld, const std::vector<int> &rd, int kind, R4Mode r4mode, bool acc,
                     int num_cus);

// If a selector model is installed: its choice of (variant, tile) for this
// contraction (memoized per shape) and true; false otherwise (no model, or
// the contraction lies outside the model's training domain).
bool selector_choose(const std::vector<int> &dig, int dr, int lr, int rr, const std::vector<int> &dd,
                     const std::vector<int> &ld, const std::vector<int> &rd, int kind, R4Mode r4mode, bool acc,
                     int num_cus, int *variant, int *tile);

}  // namespace xt
