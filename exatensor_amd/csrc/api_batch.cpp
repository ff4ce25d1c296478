// Batched block-sparse contraction (engine extension, talsh_amd.h).
//
// A batch is built once from arrays of D/L/R block tensors that share one
// contraction pattern (every tensor resident on the GPU) and can then be
// executed many times, optionally as a captured hipGraph (pack gather of the
// distinct L and R blocks, work-queue reset, persistent grouped GEMM). Blocks
// that share a D tensor form a chain that one work item accumulates in op
// order.
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
  int nops = 0, nitems = 0, njobs = 0;
  int64_t nchunks = 0, ntiles = 0;
  bool cplx = false;         // complex blocks: element-wise realifying pack
  double alpha = 1.0;
  bool acc = true;
  int grid = 0, pack_grid = 0;
  bool split3 = false;       // R4 under fast math: three-way BF16 split
  double flops = 0;
  talsh_dev_rsc_t mem{};     // ops + items + pack jobs + work queues in the arena
  talsh_dev_rsc_t packed{};  // packed L and R blocks
  BlockOp *d_ops = nullptr;
  WorkItem *d_items = nullptr;
  PackJob *d_jobs = nullptr;
  WorkQueues *d_queues = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int launch_batch(const Batch *B, hipStream_t s)
{
  int e = B->cplx ? launch_pack(B->kind, B->split3, B->d_jobs, B->njobs, B->nchunks, B->pack_grid, s)
                  : launch_pack_tiled(B->kind, B->split3, B->d_jobs, B->njobs, B->ntiles, B->pack_grid, s);
  if (e) return e;
  return launch_grouped(B->kind, B->split3, B->d_ops, B->d_items, B->d_queues, B->alpha, B->acc, B->grid, s);
}

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
  // order of the contracted indices inside the packed k index (any order is
  // the same product, both operands share it): an operand whose first
  // (stride-1) dimension is contracted reads its source contiguously only if
  // that index is the fastest one, so R's, else L's, goes first
  {
    int first = -1;
    if (dig[lr] < 0) first = -dig[lr] - 1;  // L position of R's first dimension
    else if (dig[0] < 0) first = 0;
    if (first >= 0) {
      auto it = std::find(kdims_l.begin(), kdims_l.end(), first);
      if (it != kdims_l.end()) std::rotate(kdims_l.begin(), it, it + 1);
    }
  }
  for (int i : mdims_l)
    for (int j : ndims_r)
      if (dig[i] == dig[lr + j]) return TALSH_INVALID_ARGS;  // hyper-index
  const int kind = nops > 0 && dtens[0] && dtens[0]->ndev > 0 ? dtens[0]->data_kind[0] : NO_TYPE;
  if (kind != R8 && kind != R4 && kind != C8 && kind != C4) return TALSH_INVALID_ARGS;
  const bool cplx = kind == C8 || kind == C4;
  // conjugation flags of the complex operands (D+ conjugates both)
  bool lconj = cplx && (cj & 2), rconj = cplx && (cj & 4);
  if (cplx && (cj & 1)) { lconj = !lconj; rconj = !rconj; }
  const int es = kind_size(kind) / (cplx ? 2 : 1);  // bytes of one real component
  const bool split3 = (kind == R4 || kind == C4) && d->r4mode == R4Mode::kBF16x3;
  const int kq = split3 ? 32 : 16;                 // k padding of packed operands
  const size_t packed_es = split3 ? 3 * 2 : es;    // bytes per packed element (three bf16 planes)
  std::vector<BlockOp> ops(nops);
  std::vector<void *> dptr(nops);
  std::vector<PackJob> jobs;
  // (operand body, role 0 = L / 1 = R) -> byte offset in the packed buffer
  std::map<std::pair<const void *, int>, size_t> packed_at;
  size_t packed_bytes = 0;
  auto add_job = [&](const void *src, int role, const IdxMap &rows, const IdxMap &cols, int64_t nrows, int64_t k) {
    auto it = packed_at.find({src, role});
    if (it != packed_at.end()) return it->second;
    PackJob j{};
    j.src = src;
    j.rows = rows;
    j.cols = cols;
    // complex blocks are packed realified: L as 2 rows x 2k per row, R as 2k per row
    j.cplx = cplx ? (role == 0 ? 1 : 2) : 0;
    j.conj = cplx && (role == 0 ? lconj : rconj);
    j.nrows = j.cplx == 1 ? 2 * nrows : nrows;
    j.k = k;
    const int64_t kr = cplx ? 2 * k : k;
    j.kp = (kr + kq - 1) / kq * kq;
    // tiled pack: lanes along the side whose first dimension has the smaller source stride
    j.lane_rows = cols.n == 0 || (rows.n > 0 && rows.s[0] <= cols.s[0]) ? 1 : 0;
    nrows = j.nrows;
    j.dst = reinterpret_cast<void *>(packed_bytes);  // offset for now, pointer after allocation
    jobs.push_back(j);
    packed_at[{src, role}] = packed_bytes;
    const size_t at = packed_bytes;
    // split planes are tile-blocked in bands of 16 rows
    const int64_t prow = split3 ? (nrows + 15) / 16 * 16 : nrows;
    packed_bytes += ((size_t)(prow * j.kp) * packed_es + 255) / 256 * 256;  // 256-byte aligned sub-buffers
    return at;
  };
  std::vector<size_t> a_at(nops), b_at(nops);
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
    if (st * (cplx ? 2 : 1) >= (1ll << 31)) return TALSH_LIMIT_EXCEEDED;  // 32-bit offsets inside a D block
    st = 1;
    for (int q = 0; q < lr; ++q) { ls[q] = st; st *= ld[q]; }
    st = 1;
    for (int q = 0; q < rr; ++q) { rs[q] = st; st *= rd[q]; }
    IdxMap Ml{}, Md{}, Nr{}, Nd{}, Kl{}, Kr{};
    int64_t m = 1, n = 1, k = 1;
    Ml.n = Md.n = (int)mdims_l.size();
    for (int q = 0; q < Ml.n; ++q) {
      const int i = mdims_l[q];
      if (dd[dig[i] - 1] != ld[i]) return TALSH_INVALID_ARGS;
      Ml.div[q] = Md.div[q] = make_div((uint32_t)ld[i]);
      Ml.s[q] = ls[i];
      Md.s[q] = ds[dig[i] - 1];
      m *= ld[i];
    }
    Nr.n = Nd.n = (int)ndims_r.size();
    for (int q = 0; q < Nr.n; ++q) {
      const int j = ndims_r[q];
      if (dd[dig[lr + j] - 1] != rd[j]) return TALSH_INVALID_ARGS;
      Nr.div[q] = Nd.div[q] = make_div((uint32_t)rd[j]);
      Nr.s[q] = rs[j];
      Nd.s[q] = ds[dig[lr + j] - 1];
      n *= rd[j];
    }
    Kl.n = Kr.n = (int)kdims_l.size();  // contracted indices in L's order for both operands
    for (int q = 0; q < Kl.n; ++q) {
      const int i = kdims_l[q], j = -dig[i] - 1;
      if (rd[j] != ld[i]) return TALSH_INVALID_ARGS;
      Kl.div[q] = Kr.div[q] = make_div((uint32_t)ld[i]);
      Kl.s[q] = ls[i];
      Kr.s[q] = rs[j];
      k *= ld[i];
    }
    const int64_t mr = cplx ? 2 * m : m, kr = cplx ? 2 * k : k;  // realified rows of L, contracted extent
    const int64_t kp = (kr + kq - 1) / kq * kq;
    if (mr * kp >= (1ll << 31) || n * kp >= (1ll << 31)) return TALSH_LIMIT_EXCEEDED;
    a_at[o] = add_job(pl, 0, Ml, Kl, m, k);
    b_at[o] = add_job(pr, 1, Nr, Kr, n, k);
    BlockOp &b = ops[o];
    b.m = mr;
    b.n = n;
    b.kp = kp;
    b.Md = Md;
    b.Nd = Nd;
    b.cplx = cplx ? 1 : 0;
    dptr[o] = pd;
    flops += (cplx ? 8.0 : 2.0) * (double)m * (double)n * (double)k;
  }
  int64_t nchunks = 0, ntiles = 0;
  for (PackJob &j : jobs) {
    j.chunk0 = nchunks;
    nchunks += (j.nrows * j.kp + kPackChunk - 1) / kPackChunk;
    j.tile0 = ntiles;
    j.tiles_k = (int)((j.kp + kPackTile - 1) / kPackTile);
    ntiles += (int64_t)j.tiles_k * ((j.nrows + kPackTile - 1) / kPackTile);
  }
  if (ntiles >= (1ll << 31)) return TALSH_LIMIT_EXCEEDED;
  // chains: ops grouped by D body, op order kept inside a chain
  std::vector<int> order(nops);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return dptr[a] < dptr[b]; });
  // D blocks (chains) as groups of 128 x 128 work items; a group's items
  // share operand panels (an item row shares its L panel, a column its R
  // panel), and groups whose chains start with the same L block share it
  struct Group { int i, j, tm, tn; double cost; size_t key; void *d; };
  std::vector<Group> groups;
  double total_cost = 0;
  for (int i = 0; i < nops;) {
    int j = i;
    while (j < nops && dptr[order[j]] == dptr[order[i]]) ++j;
    const BlockOp &b = ops[order[i]];
    double kc = 0;
    for (int q = i; q < j; ++q) kc += (double)ops[order[q]].kp;
    const int tm = (int)((b.m + kGroupedTile - 1) / kGroupedTile), tn = (int)((b.n + kGroupedTile - 1) / kGroupedTile);
    groups.push_back(Group{i, j, tm, tn, kc * tm * tn, a_at[order[i]], dptr[order[i]]});
    total_cost += kc * tm * tn;
    i = j;
  }
  std::stable_sort(groups.begin(), groups.end(), [](const Group &x, const Group &y) { return x.key < y.key; });
  // one queue per XCD: contiguous runs of groups of about equal cost
  WorkQueues wq{};
  std::vector<WorkItem> sitems;
  {
    double cum = 0;
    int qx = 0;
    wq.begin[0] = 0;
    for (const Group &g : groups) {
      while (qx < kQueues - 1 && cum >= total_cost * (qx + 1) / kQueues) {
        wq.end[qx] = (int)sitems.size();
        wq.begin[++qx] = (int)sitems.size();
      }
      for (int a = 0; a < g.tm; ++a)
        for (int c = 0; c < g.tn; ++c) sitems.push_back(WorkItem{g.i, g.j, a, c, g.d});
      cum += g.cost;
    }
    wq.end[qx] = (int)sitems.size();
    for (int x = qx + 1; x < kQueues; ++x) wq.begin[x] = wq.end[x] = (int)sitems.size();
  }
  Batch *B = new Batch();
  B->dev = d;
  B->kind = kind;
  B->nops = nops;
  B->nitems = (int)sitems.size();
  B->njobs = (int)jobs.size();
  B->nchunks = nchunks;
  B->ntiles = ntiles;
  B->cplx = cplx;
  B->alpha = alpha_re;
  B->acc = accumulative != NOPE;
  B->flops = flops;
  B->split3 = split3;
  // resident workgroups per CU: two (FP32/FP64 tiles), one (bf16 planes fill the LDS)
  B->grid = std::min(B->nitems, d->num_cus * (split3 ? 1 : 2));
  B->pack_grid = (int)std::min<int64_t>(cplx ? nchunks : ntiles, (int64_t)d->num_cus * 8);
  int e = image_alloc(d->flat_id, std::max<size_t>(packed_bytes, 256), &B->packed);
  if (e) { delete B; return e; }
  char *pk = static_cast<char *>(B->packed.gmem_p);
  std::vector<BlockOp> sorted(nops);
  for (int i = 0; i < nops; ++i) {
    sorted[i] = ops[order[i]];
    sorted[i].a = pk + a_at[order[i]];
    sorted[i].b = pk + b_at[order[i]];
  }
  for (PackJob &j : jobs) j.dst = pk + reinterpret_cast<size_t>(j.dst);
  const size_t ob = sizeof(BlockOp) * nops, ib = sizeof(WorkItem) * sitems.size(), jb = sizeof(PackJob) * jobs.size();
  const size_t bytes = ob + ib + jb + (sizeof(WorkQueues) + 255) / 256 * 256;
  e = image_alloc(d->flat_id, bytes, &B->mem);
  if (e) {
    image_release(B->packed, nullptr);
    delete B;
    return e;
  }
  char *base = static_cast<char *>(B->mem.gmem_p);
  B->d_ops = reinterpret_cast<BlockOp *>(base);
  B->d_items = reinterpret_cast<WorkItem *>(base + ob);
  B->d_jobs = reinterpret_cast<PackJob *>(base + ob + ib);
  B->d_queues = reinterpret_cast<WorkQueues *>(base + ob + ib + jb);
  (void)hipSetDevice(d->ordinal);
  bool ok = XT_HIP_OK(hipMemcpy(B->d_ops, sorted.data(), ob, hipMemcpyHostToDevice));
  ok = ok && XT_HIP_OK(hipMemcpy(B->d_items, sitems.data(), ib, hipMemcpyHostToDevice));
  ok = ok && XT_HIP_OK(hipMemcpy(B->d_jobs, jobs.data(), jb, hipMemcpyHostToDevice));
  ok = ok && XT_HIP_OK(hipMemcpy(B->d_queues, &wq, sizeof(wq), hipMemcpyHostToDevice));
  if (!ok) {
    image_release(B->mem, nullptr);
    image_release(B->packed, nullptr);
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
      const int e = launch_batch(B, s);
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
    const int e = launch_batch(B, s);
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
  image_release(B->packed, nullptr);
  delete B;
  return 0;
}

}  // extern "C"
