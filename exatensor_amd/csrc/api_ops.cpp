// talsh.h task objects and tensor operations (place/discard, init, scale,
// copy, add, contract) on MI355X GPUs.
//
// Dispatch and coherence semantics follow TALSH/talshc.cpp (tasks :1935-2597,
// place/discard :3318-3560, init/scale :3566-4002, copy :4530-4789, add
// :4795-5067, contract :5073-5385) and the copy-control meanings of
// TALSH/tensor_algebra.h:119-203:
//   D discard the argument everywhere after the task; M keep it only on the
//   executing device; T keep it only on its source device; K keep both.
// The GPU work of one task runs in order on one stream of the pool; tasks that
// touch the same device image are ordered through per-image events.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "contract_plan.hpp"
#include "selector.hpp"
#include "kernels.hpp"
#include "runtime_impl.hpp"
#include "talsh_amd.h"

namespace xt {

int launch_axpy_dev(void *d, const void *l, const void *s, int kind, size_t vol, double are, double aim, bool lconj,
                    bool sconj, bool acc, hipStream_t st);

namespace {

Task *task_of(const talsh_task_t *t) { return t ? static_cast<Task *>(t->task_p) : nullptr; }

void destroy_task_events(Task *k);

int new_task(talsh_task_t *tt, Device *d, int kind, unsigned coh, int nargs, hipStream_t prefer, bool timed)
{
  Task *k = new Task();
  k->timed = timed;
  k->dev_flat = d->flat_id;
  k->stream = d->streams[d->next_stream.fetch_add(1) % kStreamsPerDevice];
  for (hipStream_t st : d->streams)
    if (prefer && st == prefer) k->stream = prefer;
  bool ok = true;
  if (timed) {
    for (hipEvent_t *e : {&k->ev_start, &k->ev_comput, &k->ev_output, &k->ev_finish, &k->ev_mm0, &k->ev_mm1})
      ok = ok && (*e = event_get(true)) != nullptr;
  } else {
    ok = (k->ev_finish = event_get(false)) != nullptr;  // an untimed task has only its completion event
  }
  if (!ok) {
    destroy_task_events(k);
    delete k;
    return TALSH_FAILURE;
  }
  k->status = TALSH_TASK_SCHEDULED;
  tt->task_p = k;
  tt->task_error = -1;
  tt->dev_kind = DEV_AMD_GPU;
  tt->data_kind = kind;
  tt->coherence = (int)coh;
  tt->num_args = nargs;
  tt->data_vol = 0;
  tt->flops = 0;
  tt->exec_time = 0;
  return 0;
}

void destroy_task_events(Task *k)
{
  if (k->timed) {
    for (hipEvent_t e : {k->ev_start, k->ev_comput, k->ev_output, k->ev_finish, k->ev_mm0, k->ev_mm1})
      event_put(e, true);
  } else {
    event_put(k->ev_finish, false);
  }
  k->ev_start = k->ev_comput = k->ev_output = k->ev_finish = k->ev_mm0 = k->ev_mm1 = nullptr;
}

// Prepared operand of an operation.
struct Arg {
  talsh_tens_t *t = nullptr;
  unsigned coh = COPY_M;
  bool output = false, need_input = true;
  void *src = nullptr;       // source image body
  int src_dev = -1;
  void *exec = nullptr;      // body on the executing GPU
  bool created = false;      // exec image allocated for this task
  talsh_dev_rsc_t new_rsc{};
  int kind = NO_TYPE;
  void *attach = nullptr;    // caller memory to use as the executing-device image
};

// Chooses the source image (prefer one on the executing device) and allocates
// the executing-device image if needed. No GPU work is enqueued here.
int prepare(Arg &a, Device *d)
{
  if (!a.t || talshTensorIsEmpty(a.t) == YEP) return TALSH_OBJECT_IS_EMPTY;
  int idx = -1;
  for (int i = 0; i < a.t->ndev; ++i)
    if (a.t->dev_rsc[i].dev_id == d->flat_id && a.t->avail[i] == YEP) { idx = i; break; }
  if (idx < 0)
    for (int i = 0; i < a.t->ndev; ++i)
      if (a.t->avail[i] == YEP) { idx = i; break; }
  if (idx < 0) return a.t->ndev == 0 ? TALSH_OBJECT_IS_EMPTY : TALSH_IN_PROGRESS;
  a.kind = a.t->data_kind[idx];
  a.src = a.t->dev_rsc[idx].gmem_p;
  a.src_dev = a.t->dev_rsc[idx].dev_id;
  if (a.src_dev == d->flat_id) {
    a.exec = a.src;
    return 0;
  }
  if (a.attach) {
    a.new_rsc.dev_id = d->flat_id;
    a.new_rsc.gmem_p = a.attach;
    a.new_rsc.buf_entry = -1;
    a.new_rsc.mem_attached = 1;
  } else {
    const size_t bytes = tensor_volume(a.t) * (size_t)kind_size(a.kind);
    const int e = image_alloc(d->flat_id, bytes, &a.new_rsc);
    if (e) return e;
  }
  a.exec = a.new_rsc.gmem_p;
  a.created = true;
  return 0;
}

void unprepare(Arg &a)
{
  if (a.created) image_release(a.new_rsc, nullptr);
  a.created = false;
}

// Temporary device buffers of one task (released at completion).
struct Temps {
  Device *d;
  std::vector<talsh_dev_rsc_t> bufs;
  int alloc(size_t bytes, void **p)
  {
    talsh_dev_rsc_t r;
    int e = image_alloc(d->flat_id, bytes, &r);
    if (e) return e;
    bufs.push_back(r);
    *p = r.gmem_p;
    return 0;
  }
  void rollback()
  {
    for (auto &r : bufs) image_release(r, nullptr);
    bufs.clear();
  }
};

// Enqueues input transfers and registers the coherence actions for completion.
void enqueue_inputs(Task *k, std::vector<Arg> &args)
{
  for (Arg &a : args) {
    image_wait(k->stream, a.src);
    if (a.exec != a.src) image_wait(k->stream, a.exec);
    if (a.created && a.need_input) {
      const size_t bytes = tensor_volume(a.t) * (size_t)kind_size(a.kind);
      (void)hipMemcpyAsync(a.exec, a.src, bytes, a.src_dev == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                           k->stream);
      if (Device *d = gpu_by_flat(k->dev_flat)) d->stats.traffic_in += (double)bytes;
    }
  }
}

void enqueue_outputs(Task *k, std::vector<Arg> &args)
{
  for (Arg &a : args) {
    if (!a.output || !a.created) continue;
    if (a.coh == COPY_T || a.coh == COPY_K) {
      const size_t bytes = tensor_volume(a.t) * (size_t)kind_size(a.kind);
      (void)hipMemcpyAsync(a.src, a.exec, bytes, a.src_dev == 0 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                           k->stream);
      if (Device *d = gpu_by_flat(k->dev_flat)) d->stats.traffic_out += (double)bytes;
    }
  }
}

// Registers the new images with their tensors and marks the images that the
// task writes or will discard as unavailable until completion (inputs that
// stay valid remain usable by concurrent tasks).
void commit_args(Task *k, std::vector<Arg> &args, int exec_flat)
{
  for (Arg &a : args) {
    Task::ArgFinal f;
    f.tens = a.t;
    f.src = a.src;
    f.exec = a.exec;
    f.coh = a.coh;
    f.output = a.output;
    f.created = a.created;
    k->finals.push_back(f);
    if (a.created) {
      const int ei = tensor_add_image(a.t, a.new_rsc, a.kind);
      if (ei >= 0) a.t->avail[ei] = NOPE;
    }
    const bool lock_src = a.output || a.coh == COPY_D || (a.coh == COPY_M && a.src_dev != exec_flat);
    if (lock_src)
      for (int i = 0; i < a.t->ndev; ++i)
        if (a.t->dev_rsc[i].gmem_p == a.src) a.t->avail[i] = NOPE;
  }
}

}  // namespace

static void remove_by_ptr(talsh_tens_t *t, void *p)
{
  for (int i = 0; i < t->ndev; ++i)
    if (t->dev_rsc[i].gmem_p == p) { tensor_remove_image(t, i, nullptr); return; }
}

// Completion: coherence actions, temporaries, statistics.
int task_finalize(talsh_task_t *tt, int status)
{
  Task *k = task_of(tt);
  if (!k || k->finalized) return 0;
  for (size_t a = 0; a < k->finals.size(); ++a) {
    talsh_tens_t *t = k->finals[a].tens;
    const Task::ArgFinal &p = k->finals[a];
    if (!t || !t->shape_p) continue;
    for (int i = 0; i < t->ndev; ++i)
      if (t->dev_rsc[i].gmem_p == p.src || t->dev_rsc[i].gmem_p == p.exec) t->avail[i] = YEP;
    if (status != TALSH_TASK_COMPLETED) {
      if (p.created) remove_by_ptr(t, p.exec);  // failed task: drop the half-made image
      continue;
    }
    bool keep_src = true, keep_exec = true;
    if (p.created) {
      switch (p.coh) {
      case COPY_D: keep_src = keep_exec = false; break;
      case COPY_M: keep_src = false; break;
      case COPY_T: keep_exec = false; break;
      default: break;
      }
      if (p.output && p.coh == COPY_M) keep_src = false;
    } else if (p.coh == COPY_D) {
      keep_src = keep_exec = false;
    }
    if (p.output) {
      // every image other than the freshly written ones is stale
      for (int i = t->ndev - 1; i >= 0; --i) {
        void *q = t->dev_rsc[i].gmem_p;
        const bool fresh = (q == p.exec) || (q == p.src && p.created && (p.coh == COPY_T || p.coh == COPY_K));
        if (!fresh) tensor_remove_image(t, i, nullptr);
      }
      if (!p.created && p.coh == COPY_D) remove_by_ptr(t, p.exec);
      if (p.created && !keep_exec) remove_by_ptr(t, p.exec);
      if (p.created && !keep_src) remove_by_ptr(t, p.src);
    } else {
      if (!keep_exec && p.exec != p.src) remove_by_ptr(t, p.exec);
      if (!keep_src) remove_by_ptr(t, p.src);
      if (!p.created && !keep_exec) remove_by_ptr(t, p.exec);
    }
  }
  for (auto &f : k->frees) pending_free(f);
  k->frees.clear();
  Device *d = gpu_by_flat(k->dev_flat);
  if (d) {
    if (status == TALSH_TASK_COMPLETED) {
      d->stats.tasks_completed++;
      d->stats.flops += k->flops;
    } else {
      d->stats.tasks_failed++;
    }
  }
  k->status = status;
  k->finalized = true;
  tt->task_error = status == TALSH_TASK_COMPLETED ? 0 : 1;
  float ms = 0;
  if (status == TALSH_TASK_COMPLETED && k->timed && hipEventElapsedTime(&ms, k->ev_start, k->ev_finish) == hipSuccess)
    tt->exec_time = ms * 1e-3;
  return 0;
}

namespace {

// Shared scheduling skeleton: creates/validates the task handle, prepares the
// operands (allocations first, so TRY_LATER leaves everything unchanged),
// then runs `body` on the stream and finishes the task.
template <typename Body>
int run_task(talsh_task_t *user, Device *d, unsigned coh, std::vector<Arg> &args, int kind, double flops,
             Body body)
{
  Engine &E = engine();
  if (!E.on) return TALSH_NOT_INITIALIZED;
  talsh_task_t local;
  talsh_task_t *tt = user ? user : &local;
  if (!user) talshTaskClean(&local);
  else if (talshTaskIsEmpty(user) != YEP) return TALSH_OBJECT_NOT_EMPTY;
  (void)hipSetDevice(d->ordinal);
  const int nargs = (int)args.size();
  for (int a = 0; a < nargs; ++a) {
    args[a].coh = argument_coherence_get_value(coh, (unsigned)nargs, (unsigned)a);
    if (args[a].coh > 3) args[a].coh = COPY_M;
  }
  int e = 0;
  int prepared = 0;
  for (; prepared < nargs && !e; ++prepared) e = prepare(args[prepared], d);
  if (!e) {
    for (const Arg &a : args)
      if (a.kind != kind && kind != NO_TYPE) e = TALSH_INVALID_ARGS;
  }
  Temps temps{d, {}};
  if (!e) e = body.allocate(temps);
  if (e) {
    temps.rollback();
    for (int a = 0; a < prepared; ++a) unprepare(args[a]);
    if (e == TRY_LATER || e == DEVICE_UNABLE) d->stats.tasks_deferred++;
    return e;
  }
  // a blocking call runs where its output was last used, so it needs no
  // cross-stream wait; asynchronous tasks spread over the stream pool
  hipStream_t prefer = nullptr;
  if (!user)
    for (const Arg &a : args)
      if (a.output && (prefer = image_last_stream(a.exec)) != nullptr) break;
  // the task of a blocking call is never queried for its timings, so it records
  // only its completion event (every recorded event adds latency to the call)
  e = new_task(tt, d, kind, coh, nargs, prefer, user != nullptr);
  if (e) {
    temps.rollback();
    for (Arg &a : args) unprepare(a);
    return e;
  }
  Task *k = task_of(tt);
  k->flops = flops;
  for (int a = 0; a < nargs; ++a) {
    tt->tens_args[a].tens_p = args[a].t;
    tt->tens_args[a].source_image = 0;
  }
  d->stats.tasks_submitted++;
  if (k->timed) (void)hipEventRecord(k->ev_start, k->stream);
  enqueue_inputs(k, args);
  if (k->timed) (void)hipEventRecord(k->ev_comput, k->stream);
  e = body.enqueue(k, d);
  if (k->timed) (void)hipEventRecord(k->ev_output, k->stream);
  enqueue_outputs(k, args);
  (void)hipEventRecord(k->ev_finish, k->stream);
  {
    std::vector<void *> imgs;
    for (Arg &a : args) {
      imgs.push_back(a.exec);
      if (a.src != a.exec) imgs.push_back(a.src);
    }
    images_touch(k->stream, imgs.data(), (int)imgs.size());
  }
  for (auto &r : temps.bufs) k->frees.push_back(PendingFree{r.dev_id, r.buf_entry, nullptr});
  commit_args(k, args, d->flat_id);
  tt->flops = flops;
  if (e) {
    (void)hipStreamSynchronize(k->stream);
    task_finalize(tt, TALSH_TASK_ERROR);
    if (!user) talshTaskDestruct(&local);
    return TALSH_FAILURE;
  }
  if (!user) {
    int st = 0;
    e = talshTaskWait(&local, &st);
    if (!e && st != TALSH_TASK_COMPLETED) e = TALSH_TASK_ERROR;
    const int e2 = talshTaskDestruct(&local);
    if (!e) e = e2;
  }
  return e;
}

Device *exec_device(int dev_id, int dev_kind, int *err)
{
  Engine &E = engine();
  *err = 0;
  if (!E.on) { *err = TALSH_NOT_INITIALIZED; return nullptr; }
  const int flat = resolve_device(dev_id, dev_kind);
  if (flat == 0) { *err = TALSH_NOT_AVAILABLE; return nullptr; }  // tensor operations execute on GPUs
  Device *d = flat > 0 ? gpu_by_flat(flat) : nullptr;
  if (!d) *err = (E.gpus.empty() ? TALSH_NOT_AVAILABLE : TALSH_INVALID_ARGS);
  return d;
}

int tensor_kind(const talsh_tens_t *t)
{
  if (!t || !t->shape_p || t->ndev <= 0) return NO_TYPE;
  return t->data_kind[0];
}

std::vector<int> dims_of(const talsh_tens_t *t)
{
  return std::vector<int>(t->shape_p->dims, t->shape_p->dims + t->shape_p->num_dim);
}

struct NoTemps {
  int allocate(Temps &) { return 0; }
};

}  // namespace

}  // namespace xt

using namespace xt;

// Selected plan variant for the next contraction (-1: heuristic); set by the
// plan selector extension.
static int g_plan_variant = -1;
static int g_gemm_tile = -1;

// How GEMM-shaped contractions choose their plan (talshAmdSetPlanMode, or
// TALSH_PLAN_MODE = heuristic | learned | autotune at first use):
//   0 heuristic: fewest transpose bytes, default tile;
//   1 learned (default): the learned selector where one is installed and the
//     contraction lies in its training domain; outside that domain (with a
//     model installed) the exhaustive autotuner; no model: the heuristic;
//   2 autotune: the exhaustive autotuner for every shape.
// The autotuner's choices are memoized per shape (pattern, extents, kind,
// precision mode, accumulation).
static int g_plan_mode = -1;
static std::mutex g_tune_mu;
static std::unordered_map<std::string, std::pair<int, int>> g_tune_memo;

static int plan_mode()
{
  if (g_plan_mode < 0) {
    const char *e = std::getenv("TALSH_PLAN_MODE");
    g_plan_mode = 1;
    if (e && (!std::strcmp(e, "autotune") || !std::strcmp(e, "2"))) g_plan_mode = 2;
    if (e && (!std::strcmp(e, "heuristic") || !std::strcmp(e, "0"))) g_plan_mode = 0;
  }
  return g_plan_mode;
}

static int contract_impl(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens, double sr,
                         double si, int dev_id, int dev_kind, int copy_ctrl, int accumulative, talsh_task_t *task,
                         int force_variant, int force_tile);

namespace {

bool resident_on(const talsh_tens_t *t, int flat)
{
  for (int i = 0; i < t->ndev; ++i)
    if (t->dev_rsc[i].dev_id == flat && t->avail[i] == YEP) return true;
  return false;
}

// Exhaustive autotuner: every (matricization or fused gather, tile) candidate
// of the contraction runs once as a warm-up and once or twice timed (kernel time of
// its transposes and matrix product, HIP events on the launch stream) on the
// caller's L and R images into a scratch D on the executing GPU; the fastest
// is memoized under key. Blocking. false when it cannot tune now (L or R not
// resident on the GPU, no room for the scratch D, no candidate ran), leaving
// the choice to the heuristic for this call.
bool autotune_choose(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens, double sr,
                     double si, Device *d, bool acc, const std::vector<int> &dig, int dr, int lr, int rr,
                     const std::vector<int> &dd, const std::vector<int> &ld, const std::vector<int> &rd, int kind,
                     const std::string &key, int *variant, int *tile)
{
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune_memo.find(key);
    if (it != g_tune_memo.end()) {
      *variant = it->second.first;
      *tile = it->second.second;
      return true;
    }
  }
  if (!resident_on(ltens, d->flat_id) || !resident_on(rtens, d->flat_id)) return false;
  std::vector<PlanCandidate> cands;
  if (plan_candidates(dig, dr, lr, rr, dd, ld, rd, kind, d->r4mode, acc, d->num_cus, cands) || cands.empty())
    return false;
  talsh_tens_t scratch;
  talshTensorClean(&scratch);
  if (talshTensorConstruct(&scratch, kind, dr, dd.data(), d->flat_id, nullptr, -1, nullptr, 0.0, 0.0)) return false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> sink;
  double best = -1.0;
  int bv = -1, bt = -1;
  // Above the selector's flop domain, fused-gather candidates are left out
  // when their operands are not both staged along a contiguous run (each
  // staged element is then a separate access: 2.5-7 times the matricized
  // product's time on the 96/128 contraction) and for complex kinds (a large
  // complex product runs faster realified than in the complex kernel the
  // gather uses: 61 against 57 TFLOP/s in FP64 there).
  double lflops = 1.0;  // log2 of the real flop count, as the selector's domain is defined
  for (int i = 0; i < lr; ++i) lflops += std::log2((double)ld[i]);
  for (int j = 0; j < rr; ++j)
    if (dig[lr + j] > 0) lflops += std::log2((double)rd[j]);
  const bool large = lflops > kSelectorMaxLog2Flops;
  for (const PlanCandidate &c : cands) {
    // features 47 and 48: log2 of A's and B's contiguous run along their staging direction
    if (large && c.variant >= kFusedVariantBase && (kind_complex(kind) || c.f[47] <= 0.f || c.f[48] <= 0.f))
      continue;
    double tmin = -1.0;
    for (int rep = 0; rep < 3; ++rep) {
      prof_set_sink(&sink);
      const int e = contract_impl(cptrn, &scratch, ltens, rtens, sr, si, d->flat_id, DEV_DEFAULT, COPY_MTT,
                                  acc ? YEP : NOPE, nullptr, c.variant, c.tile);
      prof_set_sink(nullptr);
      const double ms = prof_sink_drain(sink);
      if (e || ms < 0) { tmin = -1.0; break; }
      if (rep == 0 && ms < 20.0) continue;  // warm-up (long candidates: the first run counts)
      if (tmin < 0 || ms < tmin) tmin = ms;
      if (ms >= 20.0) break;
    }
    if (tmin >= 0 && (best < 0 || tmin < best)) { best = tmin; bv = c.variant; bt = c.tile; }
  }
  talshTensorDestruct(&scratch);
  if (bv < 0) return false;
  std::lock_guard<std::mutex> g(g_tune_mu);
  if (g_tune_memo.size() >= 65536) g_tune_memo.clear();
  g_tune_memo[key] = {bv, bt};
  *variant = bv;
  *tile = bt;
  return true;
}

}  // namespace

extern "C" {

int talshAmdSetPlanVariant(int variant, int tile)
{
  g_plan_variant = variant;
  g_gemm_tile = tile;
  return 0;
}

// ============================================================================ tasks
int talshTaskCreate(talsh_task_t **t)
{
  if (!t) return TALSH_INVALID_ARGS;
  *t = (talsh_task_t *)std::malloc(sizeof(talsh_task_t));
  if (!*t) return TRY_LATER;
  return talshTaskClean(*t);
}

int talshTaskClean(talsh_task_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  t->task_p = nullptr;
  t->task_error = -1;
  t->dev_kind = DEV_NULL;
  t->data_kind = NO_TYPE;
  t->coherence = -1;
  t->num_args = 0;
  for (auto &a : t->tens_args) { a.tens_p = nullptr; a.source_image = -1; }
  t->data_vol = 0;
  t->flops = 0;
  t->exec_time = 0;
  return 0;
}

int talshTaskIsEmpty(const talsh_task_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  return t->task_p == nullptr ? YEP : NOPE;
}

int talshTaskDestruct(talsh_task_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  Task *k = task_of(t);
  if (k) {
    if (!k->finalized) {
      const hipError_t q = hipEventQuery(k->ev_finish);
      if (q == hipErrorNotReady) return TALSH_IN_PROGRESS;
      task_finalize(t, q == hipSuccess ? TALSH_TASK_COMPLETED : TALSH_TASK_ERROR);
    }
    destroy_task_events(k);
    delete k;
  }
  return talshTaskClean(t);
}

int talshTaskDestroy(talsh_task_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  const int e = talshTaskDestruct(t);
  if (e == TALSH_IN_PROGRESS) return e;
  std::free(t);
  return e;
}

int talshTaskDevId(talsh_task_t *t, int *dev_kind)
{
  Task *k = task_of(t);
  if (!k) { if (dev_kind) *dev_kind = DEV_NULL; return DEV_NULL; }
  int kind = DEV_NULL;
  const int n = decode_device_id(k->dev_flat, &kind);
  if (dev_kind) { *dev_kind = kind; return n; }
  return k->dev_flat;
}
int talshTaskDevId_(talsh_task_t *t, int *dev_kind) { return talshTaskDevId(t, dev_kind); }

int talshTaskArgCoherence(const talsh_task_t *t) { return t ? t->coherence : -1; }

const talshTensArg_t *talshTaskTensArgs(const talsh_task_t *t, int *num_args)
{
  if (!t) { if (num_args) *num_args = 0; return nullptr; }
  if (num_args) *num_args = t->num_args;
  return t->tens_args;
}

int talshTaskStatus(talsh_task_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  Task *k = task_of(t);
  if (!k) return TALSH_TASK_EMPTY;
  if (k->finalized) return k->status;
  const hipError_t q = hipEventQuery(k->ev_finish);
  if (q == hipErrorNotReady) {
    if (hipEventQuery(k->ev_output) == hipSuccess) return TALSH_TASK_OUTPUT_READY;
    if (hipEventQuery(k->ev_comput) == hipSuccess) return TALSH_TASK_INPUT_READY;
    if (hipEventQuery(k->ev_start) == hipSuccess) return TALSH_TASK_STARTED;
    return TALSH_TASK_SCHEDULED;
  }
  task_finalize(t, q == hipSuccess ? TALSH_TASK_COMPLETED : TALSH_TASK_ERROR);
  return k->status;
}

int talshTaskComplete(talsh_task_t *t, int *stats, int *ierr)
{
  if (ierr) *ierr = 0;
  if (!t) { if (ierr) *ierr = TALSH_INVALID_ARGS; return NOPE; }
  const int st = talshTaskStatus(t);
  if (stats) *stats = st;
  if (st == TALSH_TASK_EMPTY) { if (ierr) *ierr = TALSH_OBJECT_IS_EMPTY; return NOPE; }
  if (st == TALSH_TASK_ERROR) { if (ierr) *ierr = TALSH_TASK_ERROR; return YEP; }
  return st == TALSH_TASK_COMPLETED ? YEP : NOPE;
}

int talshTaskWait(talsh_task_t *t, int *stats)
{
  if (!t) return TALSH_INVALID_ARGS;
  Task *k = task_of(t);
  if (!k) { if (stats) *stats = TALSH_TASK_EMPTY; return TALSH_OBJECT_IS_EMPTY; }
  if (!k->finalized) {
    // polls the completion event, as the reference's wait loops on
    // talshTaskComplete (TALSH/talshc.cpp:2481): a blocking event wait adds
    // the runtime's wake-up latency to every synchronous call
    hipError_t e;
    while ((e = hipEventQuery(k->ev_finish)) == hipErrorNotReady) {}
    task_finalize(t, e == hipSuccess ? TALSH_TASK_COMPLETED : TALSH_TASK_ERROR);
  }
  if (stats) *stats = k->status;
  return 0;
}

int talshTasksWait(int ntasks, talsh_task_t tasks[], int stats[])
{
  if (ntasks <= 0 || !tasks) return TALSH_INVALID_ARGS;
  int err = 0;
  for (int i = 0; i < ntasks; ++i) {
    int st = TALSH_TASK_EMPTY;
    const int e = talshTaskWait(&tasks[i], &st);
    if (stats) stats[i] = st;
    if (e && !err) err = e;
  }
  return err;
}

int talshTaskTime(talsh_task_t *t, double *total, double *comput, double *input, double *output, double *mmul)
{
  if (!t || !total) return TALSH_INVALID_ARGS;
  Task *k = task_of(t);
  if (!k) return TALSH_OBJECT_IS_EMPTY;
  if (talshTaskStatus(t) != TALSH_TASK_COMPLETED) return TALSH_IN_PROGRESS;
  auto span = [](hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms * 1e-3 : -1.0;
  };
  *total = span(k->ev_start, k->ev_finish);
  if (input) *input = span(k->ev_start, k->ev_comput);
  if (comput) *comput = span(k->ev_comput, k->ev_output);
  if (output) *output = span(k->ev_output, k->ev_finish);
  if (mmul) *mmul = k->has_mm ? span(k->ev_mm0, k->ev_mm1) : -1.0;
  return 0;
}
int talshTaskTime_(talsh_task_t *t, double *total, double *comput, double *input, double *output, double *mmul)
{
  return talshTaskTime(t, total, comput, input, output, mmul);
}

void talshTaskPrint(const talsh_task_t *t)
{
  if (!t) return;
  std::printf("#MESSAGE: TAL-SH task: error %d, device kind %d, data kind %d, coherence %d, args %d, flops %.6e\n",
              t->task_error, t->dev_kind, t->data_kind, t->coherence, t->num_args, t->flops);
}

// ============================================================================ place / discard
int talshTensorPlace(talsh_tens_t *tens, int dev_id, int dev_kind, void *dev_mem, int copy_ctrl, talsh_task_t *task)
{
  Engine &E = engine();
  if (!E.on) return TALSH_NOT_INITIALIZED;
  if (!tens || talshTensorIsEmpty(tens) == YEP) return TALSH_INVALID_ARGS;
  const int flat = dev_kind == DEV_NULL ? canonical_dev(dev_id) : resolve_device(dev_id, dev_kind);
  if (flat < 0) return TALSH_INVALID_ARGS;
  if (find_image(tens, flat, NO_TYPE) >= 0) {
    if (copy_ctrl == COPY_M)
      for (int i = tens->ndev - 1; i >= 0; --i)
        if (tens->dev_rsc[i].dev_id != flat) tensor_remove_image(tens, i, nullptr);
    return 0;
  }
  if (tens->ndev <= 0) return TALSH_OBJECT_IS_EMPTY;
  if (flat == 0) {
    // device -> host: synchronous copy into a new host image
    int src = -1;
    for (int i = 0; i < tens->ndev && src < 0; ++i)
      if (tens->avail[i] == YEP) src = i;
    if (src < 0) return TALSH_IN_PROGRESS;
    const int kind = tens->data_kind[src];
    const size_t bytes = tensor_volume(tens) * (size_t)kind_size(kind);
    talsh_dev_rsc_t r;
    if (dev_mem) { r.dev_id = 0; r.gmem_p = dev_mem; r.buf_entry = -1; r.mem_attached = 1; }
    else {
      const int e = image_alloc(0, bytes, &r);
      if (e) return e;
    }
    Device *d = gpu_by_flat(tens->dev_rsc[src].dev_id);
    if (d) {
      (void)hipSetDevice(d->ordinal);
      for (auto s : d->streams) (void)hipStreamSynchronize(s);
      if (!XT_HIP_OK(hipMemcpy(r.gmem_p, tens->dev_rsc[src].gmem_p, bytes, hipMemcpyDeviceToHost))) {
        image_release(r, nullptr);
        return TALSH_FAILURE;
      }
      d->stats.traffic_out += (double)bytes;
    }
    void *srcp = tens->dev_rsc[src].gmem_p;
    tensor_add_image(tens, r, kind);
    if (copy_ctrl == COPY_M || copy_ctrl == COPY_D) remove_by_ptr(tens, srcp);
    if (task) {  // host placement completes synchronously
      talshTaskClean(task);
      task->task_error = 0;
    }
    return 0;
  }
  int err = 0;
  Device *d = exec_device(flat, DEV_NULL, &err);
  if (!d) return err;
  std::vector<Arg> args(1);
  args[0].t = tens;
  args[0].need_input = true;
  args[0].attach = dev_mem;
  struct Body {
    int allocate(Temps &) { return 0; }
    int enqueue(Task *, Device *) { return 0; }
  } body;
  const int kind = tensor_kind(tens);
  return run_task(task, d, (unsigned)(copy_ctrl & 3), args, kind, 0.0, body);
}
int talshTensorPlace_(talsh_tens_t *tens, int dev_id, int dev_kind, void *dev_mem, int copy_ctrl, talsh_task_t *task)
{
  return talshTensorPlace(tens, dev_id, dev_kind, dev_mem, copy_ctrl, task);
}

int talshTensorDiscard(talsh_tens_t *tens, int dev_id, int dev_kind)
{
  if (!engine().on) return TALSH_NOT_INITIALIZED;
  if (!tens || talshTensorIsEmpty(tens) == YEP) return TALSH_INVALID_ARGS;
  const int flat = dev_kind == DEV_NULL ? canonical_dev(dev_id) : resolve_device(dev_id, dev_kind);
  if (flat < 0) return TALSH_INVALID_ARGS;
  for (int i = tens->ndev - 1; i >= 0; --i) {
    if (tens->dev_rsc[i].dev_id != flat) continue;
    if (tens->avail[i] != YEP) return TALSH_IN_PROGRESS;
    if (Device *d = gpu_by_flat(flat)) {
      (void)hipSetDevice(d->ordinal);
      for (auto s : d->streams) (void)hipStreamSynchronize(s);
    }
    tensor_remove_image(tens, i, nullptr);
  }
  return 0;
}
int talshTensorDiscard_(talsh_tens_t *tens, int dev_id, int dev_kind) { return talshTensorDiscard(tens, dev_id, dev_kind); }

int talshTensorDiscardOther(talsh_tens_t *tens, int dev_id, int dev_kind)
{
  if (!engine().on) return TALSH_NOT_INITIALIZED;
  if (!tens || talshTensorIsEmpty(tens) == YEP) return TALSH_INVALID_ARGS;
  const int flat = dev_kind == DEV_NULL ? canonical_dev(dev_id) : resolve_device(dev_id, dev_kind);
  if (flat < 0) return TALSH_INVALID_ARGS;
  if (find_image(tens, flat, NO_TYPE) < 0) return TALSH_NOT_FOUND;
  for (int i = tens->ndev - 1; i >= 0; --i) {
    if (tens->dev_rsc[i].dev_id == flat) continue;
    if (tens->avail[i] != YEP) return TALSH_IN_PROGRESS;
    tensor_remove_image(tens, i, nullptr);
  }
  return 0;
}
int talshTensorDiscardOther_(talsh_tens_t *tens, int dev_id, int dev_kind)
{
  return talshTensorDiscardOther(tens, dev_id, dev_kind);
}

// ============================================================================ init / scale
int talshTensorInit(talsh_tens_t *dtens, double vr, double vi, int dev_id, int dev_kind, int copy_ctrl,
                    talsh_task_t *task)
{
  int err = 0;
  Device *d = exec_device(dev_id, dev_kind, &err);
  if (!d) return err;
  if (!dtens || talshTensorIsEmpty(dtens) == YEP) return TALSH_INVALID_ARGS;
  std::vector<Arg> args(1);
  args[0].t = dtens;
  args[0].output = true;
  args[0].need_input = false;
  struct Body {
    Arg *a; double re, im;
    int allocate(Temps &) { return 0; }
    int enqueue(Task *k, Device *) { return launch_fill(a->exec, a->kind, tensor_volume(a->t), re, im, k->stream); }
  } body{&args[0], vr, vi};
  return run_task(task, d, (unsigned)copy_ctrl, args, tensor_kind(dtens), 0.0, body);
}
int talshTensorInit_(talsh_tens_t *dtens, double vr, double vi, int dev_id, int dev_kind, int copy_ctrl,
                     talsh_task_t *task)
{
  return talshTensorInit(dtens, vr, vi, dev_id, dev_kind, copy_ctrl, task);
}

int talshTensorScale(talsh_tens_t *dtens, double vr, double vi, int dev_id, int dev_kind, int copy_ctrl,
                     talsh_task_t *task)
{
  int err = 0;
  Device *d = exec_device(dev_id, dev_kind, &err);
  if (!d) return err;
  if (!dtens || talshTensorIsEmpty(dtens) == YEP) return TALSH_INVALID_ARGS;
  std::vector<Arg> args(1);
  args[0].t = dtens;
  args[0].output = true;
  struct Body {
    Arg *a; double re, im;
    int allocate(Temps &) { return 0; }
    int enqueue(Task *k, Device *) { return launch_scale(a->exec, a->kind, tensor_volume(a->t), re, im, k->stream); }
  } body{&args[0], vr, vi};
  return run_task(task, d, (unsigned)copy_ctrl, args, tensor_kind(dtens), (double)tensor_volume(dtens), body);
}
int talshTensorScale_(talsh_tens_t *dtens, double vr, double vi, int dev_id, int dev_kind, int copy_ctrl,
                      talsh_task_t *task)
{
  return talshTensorScale(dtens, vr, vi, dev_id, dev_kind, copy_ctrl, task);
}

// ============================================================================ slice / insert (not on this engine)
}  // extern "C"

// Slice (D = L[offsets + x], D the box) or insert (D[offsets + x] = L, L the
// box), accumulating when asked: TALSH/talshc.cpp:4008-4300 (talshTensorSlice,
// talshTensorInsert) over TALSH/tensor_algebra_cpu.F90:2315-2620; the
// reference runs them on the host only (its GPU entries are stubs,
// TALSH/tensor_algebra_gpu_nvidia.cu:4788-4804).
static int box_op(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                  int copy_ctrl, int accumulative, talsh_task_t *task, bool insert)
{
  int err = 0;
  Device *d = exec_device(dev_id, dev_kind, &err);
  if (!d) return err;
  if (!dtens || !ltens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  const int rank = talshTensorRank(dtens);
  if (rank != talshTensorRank(ltens) || (rank > 0 && !offsets)) return TALSH_INVALID_ARGS;
  const std::vector<int> dd = dims_of(dtens), ld = dims_of(ltens);
  const std::vector<int> &big = insert ? dd : ld, &box = insert ? ld : dd;
  std::vector<int64_t> bigstr(rank), boxstr(rank);
  int64_t base = 0, sb = 1, sx = 1;
  for (int i = 0; i < rank; ++i) {
    if (offsets[i] < 0 || box[i] <= 0 || (int64_t)offsets[i] + box[i] > big[i]) return TALSH_INVALID_ARGS;
    bigstr[i] = sb;
    boxstr[i] = sx;
    base += (int64_t)offsets[i] * sb;
    sb *= big[i];
    sx *= box[i];
  }
  const bool acc = accumulative != NOPE;
  std::vector<Arg> args(2);
  args[0].t = dtens; args[0].output = true; args[0].need_input = acc || insert;  // insert keeps D outside the box
  args[1].t = ltens;
  struct Body {
    Arg *D, *L; std::vector<int> box; std::vector<int64_t> sstr, dstr; int64_t sbase, dbase; bool acc;
    int allocate(Temps &) { return 0; }
    int enqueue(Task *k, Device *)
    {
      return launch_box(L->exec, D->exec, L->kind, (int)box.size(), box.data(), sstr.data(), dstr.data(), sbase,
                        dbase, acc, k->stream);
    }
  } body{&args[0], &args[1], box, insert ? boxstr : bigstr, insert ? bigstr : boxstr, insert ? 0 : base,
         insert ? base : 0, acc};
  double vol = 1;
  for (int x : box) vol *= x;
  return run_task(task, d, (unsigned)copy_ctrl, args, tensor_kind(dtens), acc ? vol : 0.0, body);
}

extern "C" {

int talshTensorSlice(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                     int copy_ctrl, int accumulative, talsh_task_t *task)
{
  return box_op(dtens, ltens, offsets, dev_id, dev_kind, copy_ctrl, accumulative, task, false);
}
int talshTensorSlice_(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                      int copy_ctrl, int accumulative, talsh_task_t *task)
{
  return talshTensorSlice(dtens, ltens, offsets, dev_id, dev_kind, copy_ctrl, accumulative, task);
}
int talshTensorInsert(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                      int copy_ctrl, int accumulative, talsh_task_t *task)
{
  return box_op(dtens, ltens, offsets, dev_id, dev_kind, copy_ctrl, accumulative, task, true);
}
int talshTensorInsert_(talsh_tens_t *dtens, talsh_tens_t *ltens, const int *offsets, int dev_id, int dev_kind,
                       int copy_ctrl, int accumulative, talsh_task_t *task)
{
  return talshTensorInsert(dtens, ltens, offsets, dev_id, dev_kind, copy_ctrl, accumulative, task);
}

// ============================================================================ copy / add
int talshTensorCopy(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, int dev_id, int dev_kind,
                    int copy_ctrl, talsh_task_t *task)
{
  int err = 0;
  Device *d = exec_device(dev_id, dev_kind, &err);
  if (!d) return err;
  if (!cptrn || !dtens || !ltens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  int dig[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, dig, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (rr != 0 || dr != lr || dr != talshTensorRank(dtens) || lr != talshTensorRank(ltens)) return TALSH_INVALID_ARGS;
  const std::vector<int> ld = dims_of(ltens), dd = dims_of(dtens);
  std::vector<int> o2n(lr);
  for (int i = 0; i < lr; ++i) {
    o2n[i] = dig[i] - 1;
    if (o2n[i] < 0 || o2n[i] >= dr || dd[o2n[i]] != ld[i]) return TALSH_INVALID_ARGS;
  }
  bool lconj = cj & 2;
  if (cj & 1) lconj = !lconj;
  std::vector<Arg> args(2);
  args[0].t = dtens; args[0].output = true; args[0].need_input = false;
  args[1].t = ltens;
  struct Body {
    Arg *D, *L; std::vector<int> ld, o2n; bool conj;
    int allocate(Temps &) { return 0; }
    int enqueue(Task *k, Device *)
    {
      return launch_permute(L->exec, D->exec, L->kind, (int)ld.size(), ld.data(), o2n.data(), conj, k->stream);
    }
  } body{&args[0], &args[1], ld, o2n, lconj};
  const int kind = tensor_kind(dtens);
  return run_task(task, d, (unsigned)copy_ctrl, args, kind, 0.0, body);
}
int talshTensorCopy_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, int dev_id, int dev_kind,
                     int copy_ctrl, talsh_task_t *task)
{
  return talshTensorCopy(cptrn, dtens, ltens, dev_id, dev_kind, copy_ctrl, task);
}

int talshTensorAdd(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, double sr, double si, int dev_id,
                   int dev_kind, int copy_ctrl, talsh_task_t *task)
{
  int err = 0;
  Device *d = exec_device(dev_id, dev_kind, &err);
  if (!d) return err;
  if (!cptrn || !dtens || !ltens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  int dig[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, dig, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (rr != 0 || dr != lr || dr != talshTensorRank(dtens) || lr != talshTensorRank(ltens)) return TALSH_INVALID_ARGS;
  const std::vector<int> ld = dims_of(ltens), dd = dims_of(dtens);
  std::vector<int> o2n(lr);
  for (int i = 0; i < lr; ++i) {
    o2n[i] = dig[i] - 1;
    if (o2n[i] < 0 || o2n[i] >= dr || dd[o2n[i]] != ld[i]) return TALSH_INVALID_ARGS;
  }
  bool lconj = cj & 2;
  if (cj & 1) lconj = !lconj;
  std::vector<Arg> args(2);
  args[0].t = dtens; args[0].output = true;
  args[1].t = ltens;
  struct Body {
    Arg *D, *L; std::vector<int> ld, o2n; bool conj; double sr, si; void *tmp = nullptr;
    int allocate(Temps &t)
    {
      if (is_identity(o2n)) return 0;
      return t.alloc(tensor_volume(L->t) * (size_t)kind_size(tensor_kind(L->t)), &tmp);
    }
    int enqueue(Task *k, Device *)
    {
      const void *src = L->exec;
      if (tmp) {
        int e = launch_permute(L->exec, tmp, L->kind, (int)ld.size(), ld.data(), o2n.data(), false, k->stream);
        if (e) return e;
        src = tmp;
      }
      return launch_axpy(D->exec, src, D->kind, tensor_volume(D->t), sr, si, conj && kind_complex(D->kind), true,
                         k->stream);
    }
  } body{&args[0], &args[1], ld, o2n, lconj, sr, si};
  return run_task(task, d, (unsigned)copy_ctrl, args, tensor_kind(dtens), 2.0 * (double)tensor_volume(dtens), body);
}
int talshTensorAdd_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, double sr, double si, int dev_id,
                    int dev_kind, int copy_ctrl, talsh_task_t *task)
{
  return talshTensorAdd(cptrn, dtens, ltens, sr, si, dev_id, dev_kind, copy_ctrl, task);
}

// ============================================================================ contract
int talshTensorContract(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens, double sr,
                        double si, int dev_id, int dev_kind, int copy_ctrl, int accumulative, talsh_task_t *task)
{
  return contract_impl(cptrn, dtens, ltens, rtens, sr, si, dev_id, dev_kind, copy_ctrl, accumulative, task, -1, -1);
}

}  // extern "C"

// force_variant >= 0: plan with that matricization and GEMM tile (the
// autotuner's candidate runs); otherwise as talshAmdSetPlanVariant and the
// plan mode decide.
static int contract_impl(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens, double sr,
                         double si, int dev_id, int dev_kind, int copy_ctrl, int accumulative, talsh_task_t *task,
                         int force_variant, int force_tile)
{
  int err = 0;
  Device *d = exec_device(dev_id, dev_kind, &err);
  if (!d) return err;
  if (!cptrn || !dtens || !ltens || !rtens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP || talshTensorIsEmpty(rtens) == YEP)
    return TALSH_OBJECT_IS_EMPTY;
  int digc[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, digc, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (rr < 0 || dr != talshTensorRank(dtens) || lr != talshTensorRank(ltens) || rr != talshTensorRank(rtens))
    return TALSH_INVALID_ARGS;
  std::vector<int> dig(digc, digc + lr + rr);
  const std::vector<int> dd = dims_of(dtens), ld = dims_of(ltens), rd = dims_of(rtens);
  // extents must agree along every paired index
  for (int i = 0; i < lr; ++i) {
    const int v = dig[i];
    if (v > 0 ? (v > dr || dd[v - 1] != ld[i]) : (-v > rr || rd[-v - 1] != ld[i])) return TALSH_INVALID_ARGS;
  }
  for (int j = 0; j < rr; ++j) {
    const int v = dig[lr + j];
    if (v > 0 ? (v > dr || dd[v - 1] != rd[j]) : (-v > lr || ld[-v - 1] != rd[j])) return TALSH_INVALID_ARGS;
  }
  {  // every D index must come from L or R
    std::vector<int> cnt(dr + 1, 0);
    for (int v : dig)
      if (v > 0) cnt[v]++;
    for (int q = 1; q <= dr; ++q)
      if (cnt[q] == 0) return TALSH_INVALID_ARGS;
  }
  const int kind = tensor_kind(dtens);
  const bool cplx = kind_complex(kind);
  bool lconj = cplx && (cj & 2), rconj = cplx && (cj & 4);
  if (cplx && (cj & 1)) { lconj = !lconj; rconj = !rconj; }
  const bool acc = accumulative != NOPE;
  const int es = kind_size(kind);
  std::vector<Arg> args(3);
  args[0].t = dtens; args[0].output = true; args[0].need_input = acc;
  args[1].t = ltens;
  args[2].t = rtens;
  ContractPlan P;
  int plan_variant = g_plan_variant, plan_tile = g_gemm_tile;
  if (force_variant >= 0) {
    plan_variant = force_variant;
    plan_tile = force_tile;
  }
  if (dr == 0 && lr == 0 && rr == 0) P.kase = 0;
  else if ((lr == 0) != (rr == 0)) P.kase = 1;
  else if (dr == 0) P.kase = 2;
  else {
    if (plan_variant < 0) {
      const int mode = plan_mode();
      int v = -1, t = -1;
      bool chosen = mode == 1 && selector_choose(dig, dr, lr, rr, dd, ld, rd, kind, d->r4mode, acc, d->num_cus, &v, &t);
      if (!chosen && (mode == 2 || (mode == 1 && selector_installed()))) {
        const std::string key = plan_key(dig, lr, rr, dd, ld, rd, kind, d->r4mode, acc, d->num_cus);
        chosen = autotune_choose(cptrn, dtens, ltens, rtens, sr, si, d, acc, dig, dr, lr, rr, dd, ld, rd, kind, key,
                                 &v, &t);
      }
      if (chosen) {
        plan_variant = v;
        if (plan_tile < 0) plan_tile = t;
      }
    }
    int e = plan_contraction(dig, dr, lr, rr, dd, ld, rd, es, acc, plan_variant, P, nullptr);
    // the three-way BF16 split reads matricized operands: no fused gather
    if (!e && P.fused && (kind == R4 || kind == C4) && d->r4mode == R4Mode::kBF16x3)
      e = plan_contraction(dig, dr, lr, rr, dd, ld, rd, es, acc, -1, P, nullptr);
    if (e) return e;
  }
  size_t ll = 1, lrr = 1, lc = 1;
  for (int i = 0; i < lr; ++i) (dig[i] > 0 ? ll : lc) *= (size_t)ld[i];
  for (int j = 0; j < rr; ++j)
    if (dig[lr + j] > 0) lrr *= (size_t)rd[j];
  const double flops = (cplx ? 8.0 : 2.0) * (double)ll * (double)lrr * (double)lc;
  struct Body {
    Arg *D, *L, *R; ContractPlan P; std::vector<int> dig, dd, ld, rd; int lr, rr; bool lconj, rconj, acc;
    double are, aim; int tile;
    void *lt = nullptr, *rt = nullptr, *dt = nullptr, *scr = nullptr, *tt = nullptr, *ws = nullptr;
    size_t wsb = 0;
    // The matrix product as it is launched (operands exchanged when the plan
    // swaps them). The workspace is sized from the same arguments: the
    // three-way split pads rows of A and B to different multiples, so the
    // two orientations need different amounts of scratch.
    GemmArgs gemm_args(R4Mode r4mode) const
    {
      GemmArgs g{};
      g.kind = D->kind;
      g.r4mode = r4mode;
      g.tile = tile;
      g.beta_zero = !acc;
      g.alpha_re = are;
      g.alpha_im = kind_complex(D->kind) ? aim : 0.0;
      if (!P.swap) {
        g.m = P.m; g.n = P.n; g.conj_a = lconj; g.conj_b = rconj;
      } else {
        g.m = P.n; g.n = P.m; g.conj_a = rconj; g.conj_b = lconj;
      }
      g.k = P.k;
      g.a_kfast = P.a_kfast;
      g.b_kfast = P.b_kfast;
      return g;
    }
    int allocate(Temps &t)
    {
      const size_t es = kind_size(D->kind);
      if (P.kase == 1) {
        const bool l_is_tensor = lr > 0;
        const std::vector<int> &td = l_is_tensor ? ld : rd;
        std::vector<int> o2n(td.size());
        for (size_t i = 0; i < td.size(); ++i) o2n[i] = dig[(l_is_tensor ? 0 : lr) + i] - 1;
        if (!is_identity(o2n)) return t.alloc(tensor_volume(D->t) * es, &tt);
        return 0;
      }
      if (P.kase == 0) return t.alloc(1024 * sizeof(double2), &scr);
      if (P.kase == 2) {
        int e = t.alloc(1024 * sizeof(double2), &scr);
        if (e) return e;
        // R reordered into L's contracted order when they differ
        std::vector<int> o2n(rr);
        for (int j = 0; j < rr; ++j) o2n[j] = -dig[lr + j] - 1;
        if (!is_identity(o2n)) return t.alloc(tensor_volume(R->t) * es, &rt);
        return 0;
      }
      int e = 0;
      if (P.fused) {
        wsb = gemm_gather_workspace_bytes(gemm_args(t.d->r4mode));
        return wsb ? t.alloc(wsb, &ws) : 0;
      }
      if (!P.l_o2n.empty()) e = t.alloc(tensor_volume(L->t) * es, &lt);
      if (!e && !P.r_o2n.empty()) e = t.alloc(tensor_volume(R->t) * es, &rt);
      if (!e && !P.d_in_o2n.empty()) e = t.alloc(tensor_volume(D->t) * es, &dt);
      if (!e) {
        wsb = gemm_workspace_bytes(gemm_args(t.d->r4mode));
        if (wsb) e = t.alloc(wsb, &ws);
      }
      return e;
    }
    int enqueue(Task *k, Device *dev)
    {
      hipStream_t s = k->stream;
      const int kind = D->kind;
      hipEvent_t pr[2];
      if (P.kase == 0)
        return launch_dot(D->exec, L->exec, R->exec, kind, 1, are, aim, lconj, rconj, acc, scr,
                          1024 * sizeof(double2), s);
      if (P.kase == 1) {
        const bool l_is_tensor = lr > 0;
        Arg *T = l_is_tensor ? L : R, *S = l_is_tensor ? R : L;
        const std::vector<int> &td = l_is_tensor ? ld : rd;
        std::vector<int> o2n(td.size());
        for (size_t i = 0; i < td.size(); ++i) o2n[i] = dig[(l_is_tensor ? 0 : lr) + i] - 1;
        const void *src = T->exec;
        if (tt) {
          int e = launch_permute(T->exec, tt, kind, (int)td.size(), td.data(), o2n.data(), false, s);
          if (e) return e;
          src = tt;
        }
        return launch_axpy_dev(D->exec, src, S->exec, kind, tensor_volume(D->t), are, aim,
                               l_is_tensor ? lconj : rconj, l_is_tensor ? rconj : lconj, acc, s);
      }
      if (P.kase == 2) {
        const void *rsrc = R->exec;
        if (rt) {
          std::vector<int> o2n(rr);
          for (int j = 0; j < rr; ++j) o2n[j] = -dig[lr + j] - 1;
          int e = launch_permute(R->exec, rt, kind, rr, rd.data(), o2n.data(), false, s);
          if (e) return e;
          rsrc = rt;
        }
        return launch_dot(D->exec, L->exec, rsrc, kind, tensor_volume(L->t), are, aim, lconj, rconj, acc, scr,
                          1024 * sizeof(double2), s);
      }
      if (P.fused) {  // one product straight over the tensors
        GemmArgs g = gemm_args(dev->r4mode);
        g.a = P.swap ? R->exec : L->exec;
        g.b = P.swap ? L->exec : R->exec;
        g.c = D->exec;
        if (k->timed) (void)hipEventRecord(k->ev_mm0, s);
        prof_begin(s, kKGemm, pr);
        const int e = launch_gemm_gather(g, P.maps, ws, wsb, s);
        prof_end(s, kKGemm, pr);
        if (k->timed) (void)hipEventRecord(k->ev_mm1, s);
        k->has_mm = k->timed;
        return e;
      }
      // GEMM path
      const void *la = L->exec, *rb = R->exec;
      void *dc = D->exec;
      if (lt) {
        int e = launch_permute(L->exec, lt, kind, lr, ld.data(), P.l_o2n.data(), false, s);
        if (e) return e;
        la = lt;
      }
      if (rt) {
        int e = launch_permute(R->exec, rt, kind, rr, rd.data(), P.r_o2n.data(), false, s);
        if (e) return e;
        rb = rt;
      }
      if (dt) {
        if (acc) {
          int e = launch_permute(D->exec, dt, kind, (int)dd.size(), dd.data(), P.d_in_o2n.data(), false, s);
          if (e) return e;
        }
        dc = dt;
      }
      GemmArgs g = gemm_args(dev->r4mode);
      g.a = P.swap ? rb : la;
      g.b = P.swap ? la : rb;
      g.c = dc;
      if (k->timed) (void)hipEventRecord(k->ev_mm0, s);
      prof_begin(s, kKGemm, pr);
      int e = launch_gemm_ws(g, ws, wsb, s);
      prof_end(s, kKGemm, pr);
      if (k->timed) (void)hipEventRecord(k->ev_mm1, s);
      k->has_mm = k->timed;
      if (e) return e;
      if (dt) {
        e = launch_permute(dt, D->exec, kind, (int)P.dmat_dims.size(), P.dmat_dims.data(), P.d_out_o2n.data(), false,
                           s);
      }
      return e;
    }
  } body{&args[0], &args[1], &args[2], P, dig, dd, ld, rd, lr, rr, lconj, rconj, acc, sr, cplx ? si : 0.0,
         plan_tile};
  if (!cplx) body.aim = si;  // real kinds: GEMM paths use Re(alpha); the scaled-add case needs the full factor
  if (!cplx && P.kase != 1) body.aim = 0.0;
  return run_task(task, d, (unsigned)copy_ctrl, args, kind, flops, body);
}

extern "C" {

int talshAmdSetPlanMode(int mode)
{
  if (mode < 0 || mode > 2) return TALSH_INVALID_ARGS;
  g_plan_mode = mode;
  return 0;
}

int talshAmdGetPlanMode(void) { return plan_mode(); }

int talshAmdAutotuneChoice(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                           const int *rdims, int data_kind, int accumulative, int fast_math, int *variant, int *tile)
{
  if (!cptrn || !variant || !tile) return TALSH_INVALID_ARGS;
  int digc[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, digc, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr != drank || lr != lrank || rr != rrank) return TALSH_INVALID_ARGS;
  std::vector<int> dig(digc, digc + lr + rr), dd(ddims, ddims + dr), ld(ldims, ldims + lr), rd(rdims, rdims + rr);
  Engine &E = engine();
  const int cus = (E.on && !E.gpus.empty()) ? E.gpus[0]->num_cus : 256;
  const std::string key = plan_key(dig, lr, rr, dd, ld, rd, data_kind,
                                   fast_math ? R4Mode::kBF16x3 : R4Mode::kNativeF32, accumulative != NOPE, cus);
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tune_memo.find(key);
  if (it == g_tune_memo.end()) return TALSH_NOT_FOUND;
  *variant = it->second.first;
  *tile = it->second.second;
  return 0;
}

int talshAmdAutotuneClear(void)
{
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune_memo.clear();
  return 0;
}

int talshTensorContract_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens, double sr,
                         double si, int dev_id, int dev_kind, int copy_ctrl, int accumulative, talsh_task_t *task)
{
  return talshTensorContract(cptrn, dtens, ltens, rtens, sr, si, dev_id, dev_kind, copy_ctrl, accumulative, task);
}

}  // extern "C"

// ---- out-of-device-memory contraction (talshTensorContractXL) -------------------
static size_t g_xl_budget = 0;  // device bytes one sub-contraction may use (0: half the arena)

namespace {

void *image_body(const talsh_tens_t *t, int flat, int kind)
{
  for (int i = 0; i < t->ndev; ++i)
    if (t->dev_rsc[i].dev_id == flat && t->data_kind[i] == kind && t->avail[i] == YEP) return t->dev_rsc[i].gmem_p;
  return nullptr;
}

// Copies the box [offs, offs + box) of a host DLF body to (gather) or from a
// contiguous buffer, one run of box[0] elements at a time.
void host_box(char *full, const int *fdims, const size_t *offs, const int *box, int rank, size_t es, char *buf,
              bool gather)
{
  if (rank == 0) {
    if (gather) std::memcpy(buf, full, es);
    else std::memcpy(full, buf, es);
    return;
  }
  size_t rows = 1;
  for (int i = 1; i < rank; ++i) rows *= (size_t)box[i];
  std::vector<size_t> fstr(rank);
  size_t st = 1;
  for (int i = 0; i < rank; ++i) { fstr[i] = st; st *= (size_t)fdims[i]; }
  const size_t rowb = (size_t)box[0] * es;
  std::vector<int> idx(rank, 0);
  for (size_t r = 0; r < rows; ++r) {
    size_t off = offs[0];
    for (int i = 1; i < rank; ++i) off += (offs[i] + (size_t)idx[i]) * fstr[i];
    char *f = full + off * es, *b = buf + r * rowb;
    if (gather) std::memcpy(b, f, rowb);
    else std::memcpy(f, b, rowb);
    for (int i = 1; i < rank; ++i) {
      if (++idx[i] < box[i]) break;
      idx[i] = 0;
    }
  }
}

size_t slice_bytes(const talsh_tens_slice_t &s, size_t es) { return tensShape_volume(&s.shape) * es; }

}  // namespace

extern "C" {

// Contractions too large for the device run as a sequence of sub-contractions
// on slices: the operation is halved with talshTensorOpDecompose2 until every
// piece fits the budget, as in TALSH/talshc.cpp:5394-5642. Each piece gathers
// its L, R and current D slices (on the GPU with talshTensorSlice when the full
// tensor has a GPU image, else from the host image through pinned staging),
// accumulates L*R into the D slice on the GPU, and writes the D slice back.
// Pieces run in order, so pieces that share a D region (a split contracted
// index) accumulate one after the other.
int talshTensorContractXL(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens, double sr,
                          double si, int dev_id, int dev_kind, int accumulative)
{
  int err = 0;
  Device *d = exec_device(dev_id, dev_kind, &err);
  if (!d) return err;
  if (!cptrn || !dtens || !ltens || !rtens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP || talshTensorIsEmpty(rtens) == YEP)
    return TALSH_OBJECT_IS_EMPTY;
  const int kind = tensor_kind(dtens);
  const size_t es = kind_size(kind);
  const size_t dsz = tensor_volume(dtens) * es, lsz = tensor_volume(ltens) * es, rsz = tensor_volume(rtens) * es;
  const size_t budget = g_xl_budget ? g_xl_budget : d->arena.capacity() / 2;
  auto fits = [&](size_t a, size_t b, size_t c) { return 3 * (a + b + c) <= budget; };  // slices + temporaries
  if (fits(dsz, lsz, rsz))
    return talshTensorContract(cptrn, dtens, ltens, rtens, sr, si, d->flat_id, DEV_NULL, COPY_TTT, accumulative,
                               nullptr);
  // decompose
  std::vector<talsh_tens_op_t *> todo, ops;
  auto release = [&]() {
    for (auto *o : todo) talshTensorOpDestroy(o);
    for (auto *o : ops) talshTensorOpDestroy(o);
    todo.clear();
    ops.clear();
  };
  {
    talsh_tens_op_t *root = nullptr;
    int e = talshTensorOpCreate(&root);
    if (e) return e;
    todo.push_back(root);
    talsh_tens_t *args[3] = {dtens, ltens, rtens};
    for (talsh_tens_t *t : args) {
      const int r = talshTensorRank(t);
      std::vector<size_t> z(r > 0 ? r : 1, 0);
      e = talshTensorOpSetArgument(root, t, z.data(), t->shape_p->dims);
      if (e) { release(); return e; }
    }
    e = talshTensorOpSpecify(root, TALSH_TENSOR_CONTRACT, kind, cptrn, sr, si);
    if (e) { release(); return e; }
  }
  while (!todo.empty()) {
    talsh_tens_op_t *op = todo.back();
    todo.pop_back();
    if (fits(slice_bytes(op->tens_slice[0], es), slice_bytes(op->tens_slice[1], es),
             slice_bytes(op->tens_slice[2], es))) {
      ops.push_back(op);
      continue;
    }
    talsh_tens_op_t *c1 = nullptr, *c2 = nullptr;
    int e = talshTensorOpCreate(&c1);
    if (!e) e = talshTensorOpCreate(&c2);
    if (!e) e = talshTensorOpDecompose2(op, c1, c2);
    talshTensorOpDestroy(op);
    if (c2) todo.push_back(c2);
    if (c1) todo.push_back(c1);
    if (e || todo.size() + ops.size() > (1u << 16)) {
      release();
      return e ? e : DEVICE_UNABLE;
    }
  }
  // images used: a GPU image on the executing device when present, else the host image
  const int gpu = d->flat_id;
  talsh_tens_t *args[3] = {dtens, ltens, rtens};
  bool on_gpu[3];
  void *host[3];
  for (int a = 0; a < 3; ++a) {
    on_gpu[a] = image_body(args[a], gpu, kind) != nullptr;
    host[a] = on_gpu[a] ? nullptr : image_body(args[a], 0, kind);
    if (!on_gpu[a] && !host[a]) { release(); return TALSH_IN_PROGRESS; }
  }
  int e = talshTensorDiscardOther(dtens, on_gpu[0] ? gpu : 0, DEV_NULL);  // D's other images go stale
  if (!e && accumulative == NOPE) {
    if (on_gpu[0]) e = talshTensorInit(dtens, 0.0, 0.0, gpu, DEV_NULL, COPY_M, nullptr);
    else std::memset(host[0], 0, dsz);
  }
  size_t stage_bytes = 0;
  for (auto *op : ops)
    for (int a = 0; a < 3; ++a) stage_bytes = std::max(stage_bytes, slice_bytes(op->tens_slice[a], es));
  void *stage = nullptr;
  if (!e && !XT_HIP_OK(hipHostMalloc(&stage, std::max<size_t>(stage_bytes, 16), hipHostMallocDefault))) e = TRY_LATER;
  // One piece's slices on the GPU. Pieces run in order; while piece k's
  // contraction runs (an asynchronous task), the L and R slices of piece k+1
  // are loaded (GPU slice kernels on other streams, or host gathers and
  // copies), so two pieces are in flight as in the reference's progress loop
  // (TALSH/talshc.cpp:5524-5592). D slices load after the previous piece's D
  // slice is stored: pieces that split a contracted index share D regions
  // and accumulate in sequence.
  struct Piece {
    talsh_tens_t loc[3];
    std::vector<int> offs[3];
    bool loaded[3] = {false, false, false};
  };
  auto piece_clear = [&](Piece &P) {
    for (auto &t : P.loc) {
      if (talshTensorIsEmpty(&t) != YEP) talshTensorDestruct(&t);
      talshTensorClean(&t);
    }
    for (int a = 0; a < 3; ++a) { P.offs[a].clear(); P.loaded[a] = false; }
  };
  // constructs and fills argument a of piece P for operation op (blocking)
  auto piece_load = [&](talsh_tens_op_t *op, Piece &P, int a) -> int {
    const talsh_tens_slice_t &sl = op->tens_slice[a];
    const int r = sl.shape.num_dim;
    P.offs[a].clear();
    for (int i = 0; i < r; ++i) P.offs[a].push_back((int)sl.bases.offsets[i]);
    int err2 = talshTensorConstruct(&P.loc[a], kind, r, sl.shape.dims, gpu, nullptr, -1, nullptr, 0.0, 0.0);
    if (err2) { talshTensorClean(&P.loc[a]); return err2; }
    if (on_gpu[a]) {
      err2 = talshTensorSlice(&P.loc[a], args[a], P.offs[a].data(), gpu, DEV_NULL, COPY_MT, NOPE, nullptr);
    } else {
      host_box(static_cast<char *>(host[a]), args[a]->shape_p->dims, sl.bases.offsets, sl.shape.dims, r, es,
               static_cast<char *>(stage), true);
      if (!XT_HIP_OK(hipMemcpy(image_body(&P.loc[a], gpu, kind), stage, slice_bytes(sl, es), hipMemcpyHostToDevice)))
        err2 = TALSH_FAILURE;
    }
    if (!err2) P.loaded[a] = true;
    return err2;
  };
  Piece cur, nxt;
  for (auto *P : {&cur, &nxt})
    for (auto &t : P->loc) talshTensorClean(&t);
  for (size_t k = 0; !e && k < ops.size(); ++k) {
    talsh_tens_op_t *op = ops[k];
    for (int a = 1; a < 3 && !e; ++a)  // L and R unless loaded ahead
      if (!cur.loaded[a]) e = piece_load(op, cur, a);
    if (!e) e = piece_load(op, cur, 0);  // D: its current values (pieces accumulate)
    talsh_task_t tk;
    talshTaskClean(&tk);
    if (!e)
      e = talshTensorContract(cptrn, &cur.loc[0], &cur.loc[1], &cur.loc[2], sr, si, gpu, DEV_NULL, COPY_MTT, YEP,
                              &tk);
    if (!e && k + 1 < ops.size()) {  // load ahead: L and R of the next piece
      for (int a = 1; a < 3; ++a) {
        const int ea = piece_load(ops[k + 1], nxt, a);
        if (ea == TRY_LATER || ea == DEVICE_UNABLE) break;  // no room now: loaded in its own turn
        if (ea) { e = ea; break; }
      }
    }
    if (talshTaskIsEmpty(&tk) != YEP) {
      int st = 0;
      const int ew = talshTaskWait(&tk, &st);
      if (!e && (ew || st != TALSH_TASK_COMPLETED)) e = ew ? ew : TALSH_TASK_ERROR;
      talshTaskDestruct(&tk);
    }
    if (!e) {  // store the D slice
      const talsh_tens_slice_t &sl = op->tens_slice[0];
      if (on_gpu[0]) {
        e = talshTensorInsert(dtens, &cur.loc[0], cur.offs[0].data(), gpu, DEV_NULL, COPY_MT, NOPE, nullptr);
      } else if (XT_HIP_OK(hipMemcpy(stage, image_body(&cur.loc[0], gpu, kind), slice_bytes(sl, es),
                                     hipMemcpyDeviceToHost))) {
        host_box(static_cast<char *>(host[0]), dtens->shape_p->dims, sl.bases.offsets, sl.shape.dims,
                 sl.shape.num_dim, es, static_cast<char *>(stage), false);
      } else {
        e = TALSH_FAILURE;
      }
    }
    piece_clear(cur);
    std::swap(cur, nxt);
  }
  piece_clear(cur);
  piece_clear(nxt);
  if (stage) (void)hipHostFree(stage);
  release();
  return e;
}

int talshAmdSetXLBudget(size_t bytes)
{
  g_xl_budget = bytes;
  return 0;
}
int talshTensorContractXL_(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                           double sr, double si, int dev_id, int dev_kind, int accumulative)
{
  return talshTensorContractXL(cptrn, dtens, ltens, rtens, sr, si, dev_id, dev_kind, accumulative);
}

}  // extern "C"

// ============================================================================ decompositions
// D = L * S * R by a truncated SVD of D matricized per the pattern (open
// indices of L x open indices of R; TALSH/talshc.cpp:5650-6058 semantics):
// L holds the left singular vectors, R the conjugate-transposed right ones,
// S the nv largest singular values (nv = volume of the contracted indices,
// at most min(rows, columns)), absorbed per `absorb`. D is left unchanged
// (the reference destroys its copy).
static int svd_op(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                  talsh_tens_t *stens, char absorb, int dev_id, int dev_kind)
{
  int err = 0;
  Device *d = exec_device(dev_id, dev_kind, &err);
  if (!d) return err;
  if (!cptrn || !dtens || !ltens || !rtens || !stens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) == YEP || talshTensorIsEmpty(ltens) == YEP || talshTensorIsEmpty(rtens) == YEP)
    return TALSH_OBJECT_IS_EMPTY;
  int mode;
  switch (absorb) {
  case 'N': case 'n': mode = 0; break;
  case 'L': case 'l': mode = 1; break;
  case 'R': case 'r': mode = 2; break;
  case 'S': case 's': mode = 3; break;
  default: return TALSH_INVALID_ARGS;
  }
  int dig[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, dig, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr <= 0 || lr <= 0 || rr <= 0 || cj != 0) return TALSH_INVALID_ARGS;
  if (dr != talshTensorRank(dtens) || lr != talshTensorRank(ltens) || rr != talshTensorRank(rtens))
    return TALSH_INVALID_ARGS;
  int dprm[1 + MAX_TENSOR_RANK], lprm[1 + MAX_TENSOR_RANK], rprm[1 + MAX_TENSOR_RANK];
  int ncd = 0, nlu = 0, nru = 0, nhu = 0, ierr = 0;
  get_contraction_permutations(0, 0, lr, rr, dig, 0, dprm, lprm, rprm, &ncd, &nlu, &nru, &nhu, &ierr);
  if (ierr || nhu != 0 || nlu <= 0 || nru <= 0 || ncd <= 0) return TALSH_INVALID_ARGS;
  const std::vector<int> dd = dims_of(dtens), ld = dims_of(ltens), rd = dims_of(rtens);
  // matricized orders: D (N2O) -> [L-open | R-open]; L (O2N) -> [open | contracted];
  // R (O2N) -> [contracted | open]
  std::vector<int> dn(dr), ln(lr), rn(rr), d_o2n(dr), l_n2o(lr), r_n2o(rr);
  for (int i = 0; i < dr; ++i) { dn[i] = dd[dprm[1 + i] - 1]; d_o2n[dprm[1 + i] - 1] = i; }
  for (int i = 0; i < lr; ++i) { ln[lprm[1 + i] - 1] = ld[i]; l_n2o[lprm[1 + i] - 1] = i; }
  for (int i = 0; i < rr; ++i) { rn[rprm[1 + i] - 1] = rd[i]; r_n2o[rprm[1 + i] - 1] = i; }
  int64_t lu = 1, ru = 1, nv = 1;
  for (int i = 0; i < nlu; ++i) {
    if (ln[i] != dn[i]) return TALSH_INVALID_ARGS;
    lu *= dn[i];
  }
  for (int i = 0; i < nru; ++i) {
    if (rn[ncd + i] != dn[nlu + i]) return TALSH_INVALID_ARGS;
    ru *= dn[nlu + i];
  }
  std::vector<int> sdims(ncd);
  for (int i = 0; i < ncd; ++i) {
    if (ln[nlu + i] != rn[i]) return TALSH_INVALID_ARGS;
    sdims[i] = rn[i];
    nv *= rn[i];
  }
  if (nv > std::min(lu, ru)) return TALSH_INVALID_ARGS;  // more singular values than the rank allows
  const int kind = tensor_kind(dtens);
  if (tensor_kind(ltens) != kind || tensor_kind(rtens) != kind) return TALSH_INVALID_ARGS;
  // S: constructed on the executing device when empty, reshaped to the
  // contracted extents otherwise
  if (talshTensorIsEmpty(stens) == YEP) {
    const int e = talshTensorConstruct(stens, kind, ncd, sdims.data(), d->flat_id, nullptr, -1, nullptr, 0.0, 0.0);
    if (e != TALSH_SUCCESS && e != NOT_CLEAN) return e;
  } else {
    if (talshTensorRank(stens) != ncd || tensor_kind(stens) != kind) return TALSH_INVALID_ARGS;
    if (dims_of(stens) != sdims) {
      const int e = talshTensorReshape(stens, ncd, sdims.data());
      if (e != TALSH_SUCCESS) return e;
    }
  }
  auto trivial = [](const std::vector<int> &p) {
    for (size_t i = 0; i < p.size(); ++i)
      if (p[i] != (int)i) return false;
    return true;
  };
  std::vector<Arg> args(4);
  args[0].t = dtens;
  args[1].t = ltens; args[1].output = true; args[1].need_input = false;
  args[2].t = rtens; args[2].output = true; args[2].need_input = false;
  args[3].t = stens; args[3].output = true; args[3].need_input = false;
  struct Body {
    Arg *D, *L, *R, *S;
    int kind, mode;
    int64_t lu, ru, nv;
    std::vector<int> dd, d_o2n, ln, l_n2o, rn, r_n2o;
    bool dperm, lperm, rperm;
    void *z = nullptr, *lm = nullptr, *rm = nullptr, *work = nullptr;
    int allocate(Temps &t)
    {
      const size_t es = (size_t)kind_size(kind);
      int e = 0;
      if (dperm) e = t.alloc(es * (size_t)(lu * ru), &z);
      if (!e && lperm) e = t.alloc(es * (size_t)(lu * nv), &lm);
      if (!e && rperm) e = t.alloc(es * (size_t)(nv * ru), &rm);
      if (!e) e = t.alloc(svd_workspace_bytes(kind, lu, ru), &work);
      return e;
    }
    int enqueue(Task *k, Device *)
    {
      int e = 0;
      if (dperm)
        e = launch_permute(D->exec, z, kind, (int)dd.size(), dd.data(), d_o2n.data(), false, k->stream);
      if (e) return e;
      SvdArgs g{};
      g.kind = kind; g.lu = lu; g.ru = ru; g.nv = nv;
      g.a = dperm ? z : D->exec;
      g.lmat = lperm ? lm : L->exec;
      g.rmat = rperm ? rm : R->exec;
      g.svals = S->exec;
      g.work = work;
      g.absorb = mode;
      g.max_sweeps = 80;
      e = launch_svd(g, k->stream);
      if (!e && lperm) e = launch_permute(lm, L->exec, kind, (int)ln.size(), ln.data(), l_n2o.data(), false, k->stream);
      if (!e && rperm) e = launch_permute(rm, R->exec, kind, (int)rn.size(), rn.data(), r_n2o.data(), false, k->stream);
      return e;
    }
  } body{&args[0], &args[1], &args[2], &args[3], kind, mode, lu, ru, nv, dd, d_o2n, ln, l_n2o, rn, r_n2o,
         !trivial(d_o2n), !trivial(l_n2o), !trivial(r_n2o)};
  unsigned coh = 0;
  for (unsigned a = 0; a < 4; ++a) argument_coherence_set_value(&coh, 4, a, COPY_M);
  const double flops = 10.0 * (double)std::max(lu, ru) * (double)std::min(lu, ru) * (double)std::min(lu, ru);
  return run_task(nullptr, d, coh, args, kind, flops, body);
}

extern "C" {

int talshTensorDecomposeSVD(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                            talsh_tens_t *stens, const char absorb, int dev_id, int dev_kind)
{
  return svd_op(cptrn, dtens, ltens, rtens, stens, absorb, dev_id, dev_kind);
}

// the absorbing forms use a temporary S (TALSH/talshc.cpp:6060-6130)
static int svd_absorbed(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens, char absorb,
                        int dev_id, int dev_kind)
{
  talsh_tens_t stens;
  int e = talshTensorClean(&stens);
  if (e != TALSH_SUCCESS) return e;
  e = svd_op(cptrn, dtens, ltens, rtens, &stens, absorb, dev_id, dev_kind);
  if (talshTensorIsEmpty(&stens) != YEP) {
    const int e2 = talshTensorDestruct(&stens);
    if (e == TALSH_SUCCESS) e = e2;
  }
  return e;
}
int talshTensorDecomposeSVDL(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                             int dev_id, int dev_kind)
{
  return svd_absorbed(cptrn, dtens, ltens, rtens, 'L', dev_id, dev_kind);
}
int talshTensorDecomposeSVDR(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                             int dev_id, int dev_kind)
{
  return svd_absorbed(cptrn, dtens, ltens, rtens, 'R', dev_id, dev_kind);
}
int talshTensorDecomposeSVDLR(const char *cptrn, talsh_tens_t *dtens, talsh_tens_t *ltens, talsh_tens_t *rtens,
                              int dev_id, int dev_kind)
{
  return svd_absorbed(cptrn, dtens, ltens, rtens, 'S', dev_id, dev_kind);
}

// D := L * R with L, R the singular factors of D (singular values dropped):
// the isometry closest to D in its matricization. The pattern has exactly one
// contracted index; its extent is min(rows, columns) (TALSH/talshc.cpp:6111-6188).
int talshTensorOrthogonalizeSVD(const char *cptrn, talsh_tens_t *dtens, int dev_id, int dev_kind)
{
  int err = 0;
  Device *d = exec_device(dev_id, dev_kind, &err);
  if (!d) return err;
  if (!cptrn || !dtens) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(dtens) == YEP) return TALSH_OBJECT_IS_EMPTY;
  int dig[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, dig, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr <= 0 || lr <= 0 || rr <= 0 || dr != talshTensorRank(dtens)) return TALSH_INVALID_ARGS;
  const std::vector<int> dd = dims_of(dtens);
  std::vector<int> ldims(lr), rdims(rr);
  int lc = -1, rc = -1;
  int64_t lv = 1, rv = 1;
  for (int i = 0; i < lr; ++i) {
    if (dig[i] > 0) { ldims[i] = dd[dig[i] - 1]; lv *= ldims[i]; }
    else if (lc < 0) lc = i;
    else return TALSH_INVALID_ARGS;  // one contracted index only
  }
  for (int i = 0; i < rr; ++i) {
    if (dig[lr + i] > 0) { rdims[i] = dd[dig[lr + i] - 1]; rv *= rdims[i]; }
    else if (rc < 0) rc = i;
    else return TALSH_INVALID_ARGS;
  }
  if (lc < 0 || rc < 0) return TALSH_INVALID_ARGS;
  const int64_t k = std::min(lv, rv);
  if (k > INT32_MAX) return TALSH_LIMIT_EXCEEDED;
  ldims[lc] = rdims[rc] = (int)k;
  const int kind = tensor_kind(dtens);
  talsh_tens_t lt, rt, st;
  talshTensorClean(&lt);
  talshTensorClean(&rt);
  talshTensorClean(&st);
  int e = talshTensorConstruct(&lt, kind, lr, ldims.data(), d->flat_id, nullptr, -1, nullptr, 0.0, 0.0);
  if (e == NOT_CLEAN) e = TALSH_SUCCESS;
  if (e == TALSH_SUCCESS) {
    e = talshTensorConstruct(&rt, kind, rr, rdims.data(), d->flat_id, nullptr, -1, nullptr, 0.0, 0.0);
    if (e == NOT_CLEAN) e = TALSH_SUCCESS;
    if (e == TALSH_SUCCESS) {
      e = svd_op(cptrn, dtens, &lt, &rt, &st, 'N', dev_id, dev_kind);
      if (e == TALSH_SUCCESS)
        e = talshTensorContract(cptrn, dtens, &lt, &rt, 1.0, 0.0, dev_id, dev_kind, COPY_MTT, NOPE, nullptr);
      talshTensorDestruct(&rt);
    }
    talshTensorDestruct(&lt);
  }
  if (talshTensorIsEmpty(&st) != YEP) talshTensorDestruct(&st);
  return e;
}

// ============================================================================ not provided by this engine
// (nor by the reference: TALSH/talshc.cpp:6193-6204 returns TALSH_NOT_IMPLEMENTED)
int talshTensorOrthogonalizeMGS(talsh_tens_t *, int, int *, int, int) { return TALSH_NOT_IMPLEMENTED; }

}  // extern "C"

// ============================================================================ engine extensions
extern "C" {

int talshAmdProfile(int on)
{
  engine().profiling = on ? 1 : 0;
  return 0;
}

int talshAmdProfileQuery(int kernel_class, double *total_ms, long long *launches)
{
  if (!total_ms || !launches) return TALSH_INVALID_ARGS;
  return prof_collect(kernel_class, total_ms, launches);
}

int talshAmdPlanCandidates(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                           const int *rdims, int data_kind, int accumulative, int cap, double *moved, int *ncand)
{
  if (!cptrn || !ncand) return TALSH_INVALID_ARGS;
  int digc[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, digc, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr != drank || lr != lrank || rr != rrank || dr == 0 || lr == 0 || rr == 0) return TALSH_INVALID_ARGS;
  std::vector<int> dig(digc, digc + lr + rr), dd(ddims, ddims + dr), ld(ldims, ldims + lr), rd(rdims, rdims + rr);
  ContractPlan P;
  std::vector<ContractPlan> all;
  const int e = plan_contraction(dig, dr, lr, rr, dd, ld, rd, kind_size(data_kind), accumulative != NOPE, -1, P, &all);
  if (e) return e;
  *ncand = (int)all.size();
  for (int i = 0; i < (int)all.size() && i < cap; ++i)
    if (moved) moved[i] = all[i].bytes_moved;
  return 0;
}

int talshAmdPlanFusedCandidates(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims,
                                int rrank, const int *rdims, int *nfused)
{
  if (!cptrn || !nfused) return TALSH_INVALID_ARGS;
  int digc[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, digc, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr != drank || lr != lrank || rr != rrank || dr == 0 || lr == 0 || rr == 0) return TALSH_INVALID_ARGS;
  std::vector<int> dig(digc, digc + lr + rr), dd(ddims, ddims + dr), ld(ldims, ldims + lr), rd(rdims, rdims + rr);
  std::vector<ContractPlan> fz;
  const int e = plan_fused(dig, dr, lr, rr, dd, ld, rd, fz);
  if (e) return e;
  *nfused = (int)fz.size();
  return 0;
}

int talshAmdPlanFusedMaps(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                          const int *rdims, int candidate, int *info, int *ext, long long *str)
{
  if (!cptrn || !info || !ext || !str) return TALSH_INVALID_ARGS;
  int digc[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, digc, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr != drank || lr != lrank || rr != rrank || dr == 0 || lr == 0 || rr == 0) return TALSH_INVALID_ARGS;
  std::vector<int> dig(digc, digc + lr + rr), dd(ddims, ddims + dr), ld(ldims, ldims + lr), rd(rdims, rdims + rr);
  std::vector<ContractPlan> fz;
  const int e = plan_fused(dig, dr, lr, rr, dd, ld, rd, fz);
  if (e) return e;
  if (candidate < 0 || candidate >= (int)fz.size()) return TALSH_INVALID_ARGS;
  const ContractPlan &c = fz[candidate];
  const GatherGroup *g[6] = {&c.maps.a_row, &c.maps.a_k, &c.maps.b_row, &c.maps.b_k, &c.maps.c_row, &c.maps.c_col};
  info[0] = c.swap ? 1 : 0;
  info[1] = c.a_kfast ? 1 : 0;
  info[2] = c.b_kfast ? 1 : 0;
  for (int t = 0; t < 6; ++t) {
    info[3 + t] = g[t]->nd;
    for (int j = 0; j < kGatherMaxDims; ++j) {
      ext[kGatherMaxDims * t + j] = j < g[t]->nd ? g[t]->ext[j] : 1;
      str[kGatherMaxDims * t + j] = j < g[t]->nd ? (long long)g[t]->str[j] : 0;
    }
  }
  return 0;
}

int talshAmdPermutePlan(int data_kind, int rank, const int *dims, const int *o2n, int *info)
{
  if (!dims || !o2n || !info) return TALSH_INVALID_ARGS;
  PermuteInfo pi{};
  const int e = plan_permute_info(data_kind, rank, dims, o2n, &pi);
  if (e) return e;
  info[0] = pi.fused_rank; info[1] = pi.path; info[2] = pi.tile_volume;
  info[3] = pi.in_run; info[4] = pi.out_run; info[5] = pi.grid; info[6] = pi.vec;
  return 0;
}

static GemmArgs gemm_args(int data_kind, int fast_math, size_t m, size_t n, size_t k, const void *a, int a_kfast,
                          const void *b, int b_kfast, void *c, double alpha_re, double alpha_im, int beta_zero, int tile)
{
  GemmArgs g{};
  g.kind = data_kind;
  g.r4mode = fast_math ? R4Mode::kBF16x3 : R4Mode::kNativeF32;
  g.m = m; g.n = n; g.k = k; g.a = a; g.b = b; g.c = c;
  g.a_kfast = a_kfast; g.b_kfast = b_kfast; g.beta_zero = beta_zero;
  g.alpha_re = alpha_re; g.alpha_im = alpha_im; g.tile = tile;
  return g;
}

int talshAmdGemm(int data_kind, int fast_math, size_t m, size_t n, size_t k, const void *a, int a_kfast,
                 const void *b, int b_kfast, void *c, double alpha_re, double alpha_im, int beta_zero, int tile,
                 void *workspace, size_t workspace_bytes)
{
  Engine &E = engine();
  if (!E.on || E.gpus.empty()) return TALSH_NOT_INITIALIZED;
  Device *d = E.gpus[0];
  const GemmArgs g = gemm_args(data_kind, fast_math, m, n, k, a, a_kfast, b, b_kfast, c, alpha_re, alpha_im,
                               beta_zero, tile);
  const size_t need = gemm_workspace_bytes(g);
  // the three-way split cannot run without its workspace; a real product
  // without room for split-K partials runs unsplit
  if (gemm_uses_split3(g) && need > workspace_bytes) return TALSH_INVALID_ARGS;
  (void)hipSetDevice(d->ordinal);
  hipStream_t s = d->streams[0];
  hipEvent_t pr[2];
  prof_begin(s, kKGemm, pr);
  const int e = need <= workspace_bytes ? launch_gemm_ws(g, workspace, workspace_bytes, s)
                                        : launch_gemm_ws(g, nullptr, 0, s);
  prof_end(s, kKGemm, pr);
  if (e) return e;
  return XT_HIP_OK(hipStreamSynchronize(s)) ? 0 : TALSH_FAILURE;
}

// The same product with its workspace in the GPU's fixed scratch array: no
// allocation, no timing events. Products whose workspace exceeds the array
// are refused before anything is launched.
int talshAmdGemmNoAlloc(int data_kind, int fast_math, size_t m, size_t n, size_t k, const void *a, int a_kfast,
                        const void *b, int b_kfast, void *c, double alpha_re, double alpha_im, int beta_zero, int tile)
{
  Engine &E = engine();
  if (!E.on || E.gpus.empty()) return TALSH_NOT_INITIALIZED;
  Device *d = E.gpus[0];
  if (!d->noalloc_scratch) return TALSH_FAILURE;
  const GemmArgs g = gemm_args(data_kind, fast_math, m, n, k, a, a_kfast, b, b_kfast, c, alpha_re, alpha_im,
                               beta_zero, tile);
  const size_t need = gemm_workspace_bytes(g);
  if (need > kGemmNoAllocScratchBytes) {
    log_error("talshAmdGemmNoAlloc: the product needs %zu bytes of scratch, more than the fixed %zu", need,
              kGemmNoAllocScratchBytes);
    return TALSH_LIMIT_EXCEEDED;
  }
  (void)hipSetDevice(d->ordinal);
  hipStream_t s = d->streams[0];
  const int e = launch_gemm_ws(g, d->noalloc_scratch, kGemmNoAllocScratchBytes, s);
  if (e) return e;
  return XT_HIP_OK(hipStreamSynchronize(s)) ? 0 : TALSH_FAILURE;
}

size_t talshAmdGemmWorkspace(int data_kind, int fast_math, size_t m, size_t n, size_t k)
{
  GemmArgs g{};
  g.kind = data_kind;
  g.r4mode = fast_math ? R4Mode::kBF16x3 : R4Mode::kNativeF32;
  g.m = m; g.n = n; g.k = k;
  size_t bytes = 0;
  for (int tile = -1; tile < 3; ++tile) {  // enough for every tile choice
    g.tile = tile;
    bytes = std::max(bytes, gemm_workspace_bytes(g));
  }
  return bytes;
}

int talshAmdDeviceSynchronize(int dev_num)
{
  Engine &E = engine();
  if (dev_num < 0 || dev_num >= (int)E.gpus.size()) return TALSH_INVALID_ARGS;
  Device *d = E.gpus[dev_num];
  (void)hipSetDevice(d->ordinal);
  for (auto s : d->streams)
    if (!XT_HIP_OK(hipStreamSynchronize(s))) return TALSH_FAILURE;
  return 0;
}

}  // extern "C"
