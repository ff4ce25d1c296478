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
// With the three-way BF16 split (R4/C4 fast math) dst holds three bf16 planes
// of nrows * kp values each.
// Complex blocks are packed realified (the real product of the packed
// matrices is the complex product, interleaved in its rows):
//   cplx 1 (L operand): packed row 2r   = [ Re x(r,:), -Im x(r,:) ],
//                       packed row 2r+1 = [ Im x(r,:),  Re x(r,:) ];
//   cplx 2 (R operand): packed row r    = [ Re x(r,:),  Im x(r,:) ];
// nrows counts packed rows, k is the source's contracted extent (complex
// elements) and kp the packed row length (>= 2k); conj negates Im x.
struct PackJob {
  const void *src;
  void *dst;
  IdxMap rows, cols;
  int64_t nrows, k, kp;
  int64_t chunk0;  // index of this job's first chunk of kPackChunk elements
  int cplx, conj;
  // tiled pack (real blocks): first tile of this job, tiles along the packed
  // row (kp), and whether lanes read along rows (the rows' first dimension is
  // the source's contiguous one) or along k
  int64_t tile0;
  int tiles_k, lane_rows;
};
constexpr int kPackChunk = 4096;
constexpr int kPackTile = 64;  // tiled pack: 64 packed rows x 64 packed k-values

// One block product D(m x n) += A(m x k) B(n x k)^T on packed operands
// (rows of kp elements, k fastest, zero padded); Md / Nd give the offsets of
// the composite m / n indices in the D block.
// A complex block (cplx) runs as the real product of its realified packed
// operands: m counts packed rows (twice the block's rows) and row 2r + c of
// the product is component c of D element r.
struct BlockOp {
  const void *a;
  const void *b;
  int64_t m, n, kp;
  IdxMap Md, Nd;
  int cplx;
};

// A tile of one D block and the chain ops[op_begin, op_end) that accumulate
// into it, in order.
struct WorkItem {
  int op_begin, op_end;
  int mt, nt;
  void *d;
};

constexpr int kGroupedTile = 128;

// Work queues of a grouped launch, one per XCD (each XCD has its own L2): a
// workgroup takes items from the queue of its XCD (workgroup ids are dealt
// to the XCDs round robin) and, once that is empty, from the others. Each
// queue is a contiguous range of the work-item list holding whole D blocks,
// so the tiles that share operand panels meet in one L2.
constexpr int kQueues = 8;
struct WorkQueues {
  int counter[kQueues];  // items taken from each queue (reset before every launch)
  int begin[kQueues], end[kQueues];
};

// kind selects the element type (R8/C8 run on FP64, R4/C4 on FP32 matrix
// cores); split3: R4/C4 operands packed as three bf16 planes and multiplied
// with the three-way split (six MFMA products, FP64 folds); kp must then be a
// multiple of 32.
int launch_pack(int kind, bool split3, const PackJob *jobs, int njobs, int64_t nchunks, int grid, hipStream_t s);
// Real (R8/R4) blocks only: the same packing in LDS-staged tiles.
int launch_pack_tiled(int kind, bool split3, const PackJob *jobs, int njobs, int64_t ntiles, int grid,
                      hipStream_t s);
int launch_grouped(int kind, bool split3, const BlockOp *ops, const WorkItem *items, WorkQueues *queues,
                   double alpha, bool accumulate, int grid, hipStream_t s);

}  // namespace xt
