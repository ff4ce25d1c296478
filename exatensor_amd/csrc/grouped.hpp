// Device-side descriptors of a batched (grouped) block contraction.
#pragma once

#include "common.hpp"

namespace xt {

struct GDiv { uint32_t d, m, s; };

// A composite GEMM index over up to four tensor dimensions (fastest first),
// with its strides in the two tensors that carry it.
struct IdxMap {
  int n;
  GDiv div[4];
  int64_t s0[4], s1[4];
};

// One block contraction: M maps (L, D), N maps (R, D), K maps (L, R).
struct BlockOp {
  const void *l;
  const void *r;
  IdxMap M, N, K;
  int64_t m, n, k;
};

// A 64x64 output tile of one D block and the chain ops[op_begin, op_end) that
// accumulate into it, in order.
struct WorkItem {
  int op_begin, op_end;
  int mt, nt;
  void *d;
};

int launch_grouped(int kind, const BlockOp *ops, const WorkItem *items, int nitems, int *counter, double alpha,
                   bool accumulate, int grid, hipStream_t s);

}  // namespace xt
