// Device-side descriptors of a batched (grouped) block contraction.
#pragma once

#include "common.hpp"

namespace xt {

struct GDiv { uint32_t d, m, s; };

// A composite index over up to four tensor dimensions (fastest first) and its
// strides in one tensor.
struct IdxMap {
  int n;
  GDiv div[4];
  int64_t s[4];
};

// Gather of one operand block into a packed k-fast matrix:
//   dst[r * kp + kk] = src[rows(r) + cols(kk)]  (kk < k), 0 for k <= kk < kp.
// With the three-way BF16 split (R4 fast math) dst holds three bf16 planes of
// nrows * kp values each.
struct PackJob {
  const void *src;
  void *dst;
  IdxMap rows, cols;
  int64_t nrows, k, kp;
  int64_t chunk0;  // index of this job's first chunk of kPackChunk elements
};
constexpr int kPackChunk = 4096;

// One block product D(m x n) += A(m x k) B(n x k)^T on packed operands
// (rows of kp elements, k fastest, zero padded); Md / Nd give the offsets of
// the composite m / n indices in the D block.
struct BlockOp {
  const void *a;
  const void *b;
  int64_t m, n, kp;
  IdxMap Md, Nd;
};

// A tile of one D block and the chain ops[op_begin, op_end) that accumulate
// into it, in order.
struct WorkItem {
  int op_begin, op_end;
  int mt, nt;
  void *d;
};

constexpr int kGroupedTile = 128;

// split3: R4 operands packed as three bf16 planes and multiplied with the
// three-way split (six MFMA products, FP64 folds); kp must then be a multiple
// of 32.
int launch_pack(int kind, bool split3, const PackJob *jobs, int njobs, int64_t nchunks, int grid, hipStream_t s);
int launch_grouped(int kind, bool split3, const BlockOp *ops, const WorkItem *items, int nitems, int *counter,
                   double alpha, bool accumulate, int grid, hipStream_t s);

}  // namespace xt
