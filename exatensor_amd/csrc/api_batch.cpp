// Batched block-sparse contraction (engine extension, talsh_amd.h).
//
// A batch is built once from arrays of D/L/R block tensors that share one
// contraction pattern (every tensor resident on the GPU) and can then be
// executed many times, optionally as a captured hipGraph (counter reset +
// persistent grouped kernel). Blocks that share a D tensor form a chain that
// one work item accumulates in op order.
#include <algorithm>
#include <map>
#include <numeric>
#include <vector>

#include "grouped.hpp"
#include "kernels.hpp"
#include "runtime_impl.hpp"
#include "talsh_amd.h"

namespace xt {
namespace {

GDiv make_div(uint32_t d)
{
  FastDiv f(d);
  return GDiv{f.d, f.m, f.s};
}

struct Batch {
  Device *dev = nullptr;
  int kind = NO_TYPE;
  int nops = 0, nitems = 0;
  double alpha = 1.0;
  bool acc = true;
  int grid = 0;
  double flops = 0;
  talsh_dev_rsc_t mem{};     // ops + items + counter in the arena
  BlockOp *d_ops = nullptr;
  WorkItem *d_items = nullptr;
  int *d_counter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

void *gpu_body(talsh_tens_t *t, int flat, int kind)
{
  if (!t || !t->shape_p) return nullptr;
  for (int i = 0; i < t->ndev; ++i)
    if (t->dev_rsc[i].dev_id == flat && t->data_kind[i] == kind && t->avail[i] == YEP) return t->dev_rsc[i].gmem_p;
  return nullptr;
}

}  // namespace
}  // namespace xt

using namespace xt;

extern "C" {

int talshAmdBatchCreate(const char *cptrn, int nops, talsh_tens_t *const *dtens, talsh_tens_t *const *ltens,
                        talsh_tens_t *const *rtens, double alpha_re, int accumulative, int dev_num, void **batch)
{
  Engine &E = engine();
  if (!E.on) return TALSH_NOT_INITIALIZED;
  if (!cptrn || nops <= 0 || !dtens || !ltens || !rtens || !batch) return TALSH_INVALID_ARGS;
  if (dev_num < 0 || dev_num >= (int)E.gpus.size()) return TALSH_INVALID_ARGS;
  Device *d = E.gpus[dev_num];
  int dig[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, dig, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr <= 0 || lr <= 0 || rr <= 0) return TALSH_INVALID_ARGS;
  // index groups of the pattern: D positions from L (M), from R (N), contracted (K in L order)
  std::vector<int> mdims_l, ndims_r, kdims_l;  // positions in L / R
  for (int i = 0; i < lr; ++i) (dig[i] > 0 ? mdims_l : kdims_l).push_back(i);
  for (int j = 0; j < rr; ++j)
    if (dig[lr + j] > 0) ndims_r.push_back(j);
  if (mdims_l.size() > 4 || ndims_r.size() > 4 || kdims_l.size() > 4) return TALSH_NOT_IMPLEMENTED;
  for (int i : mdims_l)
    for (int j : ndims_r)
      if (dig[i] == dig[lr + j]) return TALSH_INVALID_ARGS;  // hyper-index
  const int kind = nops > 0 && dtens[0] && dtens[0]->ndev > 0 ? dtens[0]->data_kind[0] : NO_TYPE;
  if (kind != R8 && kind != R4) return TALSH_NOT_IMPLEMENTED;
  std::vector<BlockOp> ops(nops);
  std::vector<void *> dptr(nops);
  double flops = 0;
  for (int o = 0; o < nops; ++o) {
    talsh_tens_t *D = dtens[o], *L = ltens[o], *R = rtens[o];
    if (talshTensorRank(D) != dr || talshTensorRank(L) != lr || talshTensorRank(R) != rr) return TALSH_INVALID_ARGS;
    void *pd = gpu_body(D, d->flat_id, kind), *pl = gpu_body(L, d->flat_id, kind), *pr = gpu_body(R, d->flat_id, kind);
    if (!pd || !pl || !pr) return TALSH_INVALID_ARGS;  // operands must be resident on the GPU
    const int *dd = D->shape_p->dims, *ld = L->shape_p->dims, *rd = R->shape_p->dims;
    std::vector<int64_t> ds(dr), ls(lr), rs(rr);
    int64_t st = 1;
    for (int q = 0; q < dr; ++q) { ds[q] = st; st *= dd[q]; }
    st = 1;
    for (int q = 0; q < lr; ++q) { ls[q] = st; st *= ld[q]; }
    st = 1;
    for (int q = 0; q < rr; ++q) { rs[q] = st; st *= rd[q]; }
    BlockOp &b = ops[o];
    b.l = pl;
    b.r = pr;
    b.M.n = (int)mdims_l.size();
    b.m = 1;
    for (int q = 0; q < b.M.n; ++q) {
      const int i = mdims_l[q];
      if (dd[dig[i] - 1] != ld[i]) return TALSH_INVALID_ARGS;
      b.M.div[q] = make_div((uint32_t)ld[i]);
      b.M.s0[q] = ls[i];
      b.M.s1[q] = ds[dig[i] - 1];
      b.m *= ld[i];
    }
    b.N.n = (int)ndims_r.size();
    b.n = 1;
    for (int q = 0; q < b.N.n; ++q) {
      const int j = ndims_r[q];
      if (dd[dig[lr + j] - 1] != rd[j]) return TALSH_INVALID_ARGS;
      b.N.div[q] = make_div((uint32_t)rd[j]);
      b.N.s0[q] = rs[j];
      b.N.s1[q] = ds[dig[lr + j] - 1];
      b.n *= rd[j];
    }
    b.K.n = (int)kdims_l.size();
    b.k = 1;
    for (int q = 0; q < b.K.n; ++q) {
      const int i = kdims_l[q], j = -dig[i] - 1;
      if (rd[j] != ld[i]) return TALSH_INVALID_ARGS;
      b.K.div[q] = make_div((uint32_t)ld[i]);
      b.K.s0[q] = ls[i];
      b.K.s1[q] = rs[j];
      b.k *= ld[i];
    }
    if (b.m >= (1ll << 31) || b.n >= (1ll << 31) || b.k >= (1ll << 31)) return TALSH_LIMIT_EXCEEDED;
    dptr[o] = pd;
    flops += 2.0 * (double)b.m * (double)b.n * (double)b.k;
  }
  // chains: ops grouped by D body, op order kept inside a chain
  std::vector<int> order(nops);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return dptr[a] < dptr[b]; });
  std::vector<BlockOp> sorted(nops);
  for (int i = 0; i < nops; ++i) sorted[i] = ops[order[i]];
  std::vector<WorkItem> items;
  std::vector<double> cost;
  for (int i = 0; i < nops;) {
    int j = i;
    while (j < nops && dptr[order[j]] == dptr[order[i]]) ++j;
    const BlockOp &b = sorted[i];
    double kc = 0;
    for (int q = i; q < j; ++q) kc += (double)sorted[q].k;
    const int tm = (int)((b.m + 63) / 64), tn = (int)((b.n + 63) / 64);
    for (int a = 0; a < tm; ++a)
      for (int c = 0; c < tn; ++c) {
        items.push_back(WorkItem{i, j, a, c, dptr[order[i]]});
        cost.push_back(kc);
      }
    i = j;
  }
  // longest work first
  std::vector<int> io(items.size());
  std::iota(io.begin(), io.end(), 0);
  std::stable_sort(io.begin(), io.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  std::vector<WorkItem> sitems(items.size());
  for (size_t i = 0; i < items.size(); ++i) sitems[i] = items[io[i]];
  Batch *B = new Batch();
  B->dev = d;
  B->kind = kind;
  B->nops = nops;
  B->nitems = (int)sitems.size();
  B->alpha = alpha_re;
  B->acc = accumulative != NOPE;
  B->flops = flops;
  B->grid = std::min(B->nitems, d->num_cus * 4);  // 4 resident workgroups per CU (LDS-bound)
  const size_t bytes = sizeof(BlockOp) * nops + sizeof(WorkItem) * sitems.size() + 256;
  int e = image_alloc(d->flat_id, bytes, &B->mem);
  if (e) { delete B; return e; }
  char *base = static_cast<char *>(B->mem.gmem_p);
  B->d_ops = reinterpret_cast<BlockOp *>(base);
  B->d_items = reinterpret_cast<WorkItem *>(base + sizeof(BlockOp) * nops);
  B->d_counter = reinterpret_cast<int *>(base + bytes - 256);
  (void)hipSetDevice(d->ordinal);
  bool ok = XT_HIP_OK(hipMemcpy(B->d_ops, sorted.data(), sizeof(BlockOp) * nops, hipMemcpyHostToDevice));
  ok = ok && XT_HIP_OK(hipMemcpy(B->d_items, sitems.data(), sizeof(WorkItem) * sitems.size(), hipMemcpyHostToDevice));
  if (!ok) {
    image_release(B->mem, nullptr);
    delete B;
    return TALSH_FAILURE;
  }
  *batch = B;
  return 0;
}

int talshAmdBatchExecute(void *batch, int use_graph)
{
  Batch *B = static_cast<Batch *>(batch);
  if (!B) return TALSH_INVALID_ARGS;
  Device *d = B->dev;
  (void)hipSetDevice(d->ordinal);
  for (auto st : d->streams)  // operands may still be written by earlier tasks
    if (!XT_HIP_OK(hipStreamSynchronize(st))) return TALSH_FAILURE;
  hipStream_t s = d->streams[0];
  hipEvent_t pr[2];
  if (use_graph) {
    if (!B->exec) {
      if (!XT_HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal))) return TALSH_FAILURE;
      const int e = launch_grouped(B->kind, B->d_ops, B->d_items, B->nitems, B->d_counter, B->alpha, B->acc, B->grid, s);
      hipGraph_t g = nullptr;
      const bool ok = XT_HIP_OK(hipStreamEndCapture(s, &g));
      if (e || !ok) return TALSH_FAILURE;
      B->graph = g;
      if (!XT_HIP_OK(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0))) return TALSH_FAILURE;
    }
    prof_begin(s, kKGrouped, pr);
    const bool ok = XT_HIP_OK(hipGraphLaunch(B->exec, s));
    prof_end(s, kKGrouped, pr);
    if (!ok) return TALSH_FAILURE;
  } else {
    prof_begin(s, kKGrouped, pr);
    const int e = launch_grouped(B->kind, B->d_ops, B->d_items, B->nitems, B->d_counter, B->alpha, B->acc, B->grid, s);
    prof_end(s, kKGrouped, pr);
    if (e) return e;
  }
  if (!XT_HIP_OK(hipStreamSynchronize(s))) return TALSH_FAILURE;
  d->stats.flops += B->flops;
  return 0;
}

int talshAmdBatchInfo(void *batch, int *nops, int *nitems, double *flops)
{
  Batch *B = static_cast<Batch *>(batch);
  if (!B) return TALSH_INVALID_ARGS;
  if (nops) *nops = B->nops;
  if (nitems) *nitems = B->nitems;
  if (flops) *flops = B->flops;
  return 0;
}

int talshAmdBatchDestroy(void *batch)
{
  Batch *B = static_cast<Batch *>(batch);
  if (!B) return TALSH_INVALID_ARGS;
  (void)hipSetDevice(B->dev->ordinal);
  (void)hipStreamSynchronize(B->dev->streams[0]);
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  image_release(B->mem, nullptr);
  delete B;
  return 0;
}

}  // extern "C"
