// talsh.h entry points for library control, tensor blocks and tasks.
//
// Behaviour follows TALSH/talshc.cpp (init :546-621, construct :1167-1302,
// tensor queries :1304-1700, tasks :1935-2597); the runtime underneath is this
// engine's (arena per GPU, stream pool, per-image ordering events).
#include <dlfcn.h>

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>

#include <algorithm>

#include "kernels.hpp"
#include "runtime_impl.hpp"
#include "talsh_amd.h"

namespace xt {

void log_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  std::fprintf(stderr, "#ERROR(talsh-mi355x): ");
  std::vfprintf(stderr, fmt, ap);
  std::fprintf(stderr, "\n");
  va_end(ap);
}

Engine &engine()
{
  static Engine e;
  return e;
}

bool is_gpu_flat(int flat)
{
  int kind = DEV_NULL;
  const int num = decode_device_id(flat, &kind);
  return (kind == DEV_AMD_GPU || kind == DEV_NVIDIA_GPU) && num >= 0;
}

Device *gpu_by_flat(int flat)
{
  int kind = DEV_NULL;
  const int num = decode_device_id(flat, &kind);
  if ((kind != DEV_AMD_GPU && kind != DEV_NVIDIA_GPU) || num < 0) return nullptr;
  auto &g = engine().gpus;
  return num < (int)g.size() ? g[num] : nullptr;
}

int gpu_flat_id(int dev_num) { return encode_device_id(DEV_AMD_GPU, dev_num); }

int canonical_dev(int flat)
{
  if (flat == 0) return 0;
  Device *d = gpu_by_flat(flat);
  return d ? d->flat_id : -1;
}

// Resolves (dev_id, dev_kind) of an API call to a canonical flat id: 0 for the
// host, the DEV_AMD_GPU flat id for a GPU, -1 if invalid. DEV_DEFAULT picks GPU 0.
int resolve_device(int dev_id, int dev_kind)
{
  Engine &E = engine();
  if (dev_kind == DEV_DEFAULT) {
    if (dev_id == DEV_DEFAULT) return E.gpus.empty() ? -1 : E.gpus[0]->flat_id;
    return canonical_dev(dev_id);
  }
  if (dev_kind == DEV_HOST) return (dev_id == 0 || dev_id == DEV_DEFAULT) ? 0 : -1;
  if (dev_kind == DEV_AMD_GPU || dev_kind == DEV_NVIDIA_GPU) {
    const int num = dev_id == DEV_DEFAULT ? 0 : dev_id;
    if (num < 0 || num >= (int)E.gpus.size()) return -1;
    return E.gpus[num]->flat_id;
  }
  return -1;
}

// ---- host memory ---------------------------------------------------------------
static std::mutex g_host_mu;
static std::unordered_map<void *, int> g_host_allocs;  // 1 pinned, 0 malloc

void *host_alloc(size_t bytes)
{
  void *p = nullptr;
  int pinned = 0;
  if (!engine().gpus.empty() && bytes > 0) {
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess) pinned = 1;
    else { (void)hipGetLastError(); p = nullptr; }
  }
  if (!p) p = std::malloc(bytes ? bytes : 1);
  if (!p) return nullptr;
  std::lock_guard<std::mutex> g(g_host_mu);
  g_host_allocs[p] = pinned;
  return p;
}

void host_free(void *p)
{
  if (!p) return;
  int pinned = 0;
  {
    std::lock_guard<std::mutex> g(g_host_mu);
    auto it = g_host_allocs.find(p);
    if (it == g_host_allocs.end()) return;
    pinned = it->second;
    g_host_allocs.erase(it);
  }
  if (pinned) (void)hipHostFree(p);
  else std::free(p);
}

// ---- image helpers ---------------------------------------------------------------
int image_alloc(int dev_flat, size_t bytes, talsh_dev_rsc_t *rsc)
{
  rsc->dev_id = dev_flat;
  rsc->mem_attached = 0;
  rsc->buf_entry = -1;
  rsc->gmem_p = nullptr;
  if (dev_flat == 0) {
    rsc->gmem_p = host_alloc(bytes);
    return rsc->gmem_p ? 0 : TRY_LATER;
  }
  Device *d = gpu_by_flat(dev_flat);
  if (!d) return TALSH_INVALID_ARGS;
  void *p = nullptr;
  int h = -1;
  int e = d->arena.alloc(bytes, &p, &h);
  if (e) return e;
  rsc->gmem_p = p;
  rsc->buf_entry = h;
  return 0;
}

void image_release(const talsh_dev_rsc_t &r, Task *defer)
{
  if (r.mem_attached || !r.gmem_p) return;
  if (defer) {
    defer->frees.push_back(PendingFree{r.dev_id, r.buf_entry, r.dev_id == 0 ? r.gmem_p : nullptr});
    return;
  }
  if (r.dev_id == 0) host_free(r.gmem_p);
  else if (Device *d = gpu_by_flat(r.dev_id)) d->arena.release(r.buf_entry);
}

void pending_free(const PendingFree &f)
{
  if (f.dev_flat == 0) host_free(f.host_ptr);
  else if (Device *d = gpu_by_flat(f.dev_flat)) d->arena.release(f.handle);
}

int tensor_add_image(talsh_tens_t *t, const talsh_dev_rsc_t &r, int kind)
{
  if (t->ndev >= t->dev_rsc_len) return TALSH_LIMIT_EXCEEDED;
  const int i = t->ndev++;
  t->dev_rsc[i] = r;
  t->data_kind[i] = kind;
  t->avail[i] = YEP;
  return i;
}

void tensor_remove_image(talsh_tens_t *t, int i, Task *defer)
{
  if (i < 0 || i >= t->ndev) return;
  image_release(t->dev_rsc[i], defer);
  for (int j = i; j + 1 < t->ndev; ++j) {
    t->dev_rsc[j] = t->dev_rsc[j + 1];
    t->data_kind[j] = t->data_kind[j + 1];
    t->avail[j] = t->avail[j + 1];
  }
  --t->ndev;
}

int find_image(const talsh_tens_t *t, int dev_flat, int kind)
{
  for (int i = 0; i < t->ndev; ++i)
    if (t->dev_rsc[i].dev_id == dev_flat && (kind == NO_TYPE || t->data_kind[i] == kind)) return i;
  return -1;
}

size_t tensor_volume(const talsh_tens_t *t) { return tensShape_volume(t->shape_p); }

// ---- ordering between tasks on different streams --------------------------------
// device image -> event recorded after the last task that touched it, and the
// stream it was recorded on; one event is shared by all images of a task
// (reference counted, returned to the pool when no image refers to it)
struct LastUse { hipEvent_t ev; hipStream_t s; };
static std::mutex g_sync_mu;
static std::unordered_map<void *, LastUse> g_last_use;
static std::unordered_map<hipEvent_t, int> g_ev_refs;

// ---- event pool: hipEventCreate costs microseconds per call, and every task
// uses several events; released events are kept per (device, timing) ------------
static std::mutex g_evpool_mu;
static std::map<std::pair<int, bool>, std::vector<hipEvent_t>> g_evpool;

hipEvent_t event_get(bool timing)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> g(g_evpool_mu);
    auto &v = g_evpool[{dev, timing}];
    if (!v.empty()) {
      hipEvent_t e = v.back();
      v.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, timing ? hipEventDefault : hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

void event_put(hipEvent_t e, bool timing)
{
  if (!e) return;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> g(g_evpool_mu);
  g_evpool[{dev, timing}].push_back(e);
}

void event_pool_clear()
{
  std::lock_guard<std::mutex> g(g_evpool_mu);
  for (auto &kv : g_evpool)
    for (hipEvent_t e : kv.second) (void)hipEventDestroy(e);
  g_evpool.clear();
}

static void ev_unref(hipEvent_t e)
{
  auto it = g_ev_refs.find(e);
  if (it != g_ev_refs.end() && --it->second == 0) {
    g_ev_refs.erase(it);
    event_put(e, false);  // waits already enqueued on it keep their record
  }
}

void image_wait(hipStream_t s, void *img)
{
  std::lock_guard<std::mutex> g(g_sync_mu);
  auto it = g_last_use.find(img);
  if (it == g_last_use.end() || it->second.s == s) return;  // same stream: queue order suffices
  if (hipEventQuery(it->second.ev) == hipSuccess) return;    // already complete
  (void)hipStreamWaitEvent(s, it->second.ev, 0);
}

hipStream_t image_last_stream(void *img)
{
  std::lock_guard<std::mutex> g(g_sync_mu);
  auto it = g_last_use.find(img);
  return it == g_last_use.end() ? nullptr : it->second.s;
}

void images_touch(hipStream_t s, void *const *imgs, int n)
{
  if (n <= 0) return;
  hipEvent_t ev = event_get(false);
  if (!ev) return;
  (void)hipEventRecord(ev, s);
  std::lock_guard<std::mutex> g(g_sync_mu);
  for (int i = 0; i < n; ++i) {
    auto it = g_last_use.find(imgs[i]);
    if (it != g_last_use.end()) {
      if (it->second.ev == ev) continue;  // listed twice
      ev_unref(it->second.ev);
      it->second = LastUse{ev, s};
    } else {
      g_last_use[imgs[i]] = LastUse{ev, s};
    }
    ++g_ev_refs[ev];
  }
  if (g_ev_refs.find(ev) == g_ev_refs.end()) event_put(ev, false);
}

void image_touch(hipStream_t s, void *img) { images_touch(s, &img, 1); }

void image_forget_all()
{
  std::lock_guard<std::mutex> g(g_sync_mu);
  for (auto &kv : g_ev_refs) (void)hipEventDestroy(kv.first);
  g_ev_refs.clear();
  g_last_use.clear();
}

// ---- kernel profiling --------------------------------------------------------------
struct ProfRec { int id; hipEvent_t a, b; };
static std::mutex g_prof_mu;
static std::vector<ProfRec> g_prof;
// When set on this thread, kernel timings go to the sink instead of the
// engine's profile (the autotuner times candidates without touching the
// caller's profiling records).
static thread_local std::vector<std::pair<hipEvent_t, hipEvent_t>> *tl_prof_sink = nullptr;

void prof_set_sink(std::vector<std::pair<hipEvent_t, hipEvent_t>> *sink) { tl_prof_sink = sink; }

double prof_sink_drain(std::vector<std::pair<hipEvent_t, hipEvent_t>> &sink)
{
  double t = 0;
  bool ok = true;
  for (auto &r : sink) {
    float ms = 0;
    if (hipEventSynchronize(r.second) == hipSuccess && hipEventElapsedTime(&ms, r.first, r.second) == hipSuccess)
      t += ms;
    else
      ok = false;
    event_put(r.first, true);
    event_put(r.second, true);
  }
  sink.clear();
  return ok ? t : -1.0;
}

void prof_begin(hipStream_t s, int id, hipEvent_t *pair)
{
  pair[0] = pair[1] = nullptr;
  if (!engine().profiling && !tl_prof_sink) return;
  pair[0] = event_get(true);
  pair[1] = event_get(true);
  if (!pair[0] || !pair[1]) {
    event_put(pair[0], true);
    event_put(pair[1], true);
    pair[0] = pair[1] = nullptr;
    return;
  }
  (void)hipEventRecord(pair[0], s);
  (void)id;
}

void prof_end(hipStream_t s, int id, hipEvent_t *pair)
{
  if (!pair[0] || !pair[1]) return;
  (void)hipEventRecord(pair[1], s);
  if (tl_prof_sink) {
    tl_prof_sink->emplace_back(pair[0], pair[1]);
    return;
  }
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof.push_back(ProfRec{id, pair[0], pair[1]});
}

int prof_collect(int id, double *total_ms, long long *count)
{
  std::lock_guard<std::mutex> g(g_prof_mu);
  double t = 0;
  long long c = 0;
  std::vector<ProfRec> keep;
  for (auto &r : g_prof) {
    if (r.id != id) { keep.push_back(r); continue; }
    float ms = 0;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      t += ms;
      ++c;
    }
    event_put(r.a, true);
    event_put(r.b, true);
  }
  g_prof.swap(keep);
  *total_ms = t;
  *count = c;
  return 0;
}

}  // namespace xt

using namespace xt;

extern "C" {

// ============================================================================ control
int talshInit(size_t *host_buf_size, int *host_arg_max, int ngpus, int gpu_list[], int nmics, int mic_list[],
              int namds, int amd_list[])
{
  (void)mic_list;
  Engine &E = engine();
  std::lock_guard<std::mutex> g(E.mu);
  if (E.on) return TALSH_ALREADY_INITIALIZED;
  if (nmics > 0) return TALSH_NOT_AVAILABLE;
  if (ngpus < 0 || namds < 0 || ngpus > MAX_GPUS_PER_NODE || namds > MAX_AMDS_PER_NODE) return TALSH_INVALID_ARGS;
  // the GPU lists name HIP device ordinals; DEV_NVIDIA_GPU entries are an alias
  std::vector<int> ords;
  for (int i = 0; i < ngpus; ++i) ords.push_back(gpu_list[i]);
  for (int i = 0; i < namds; ++i)
    if (std::find(ords.begin(), ords.end(), amd_list[i]) == ords.end()) ords.push_back(amd_list[i]);
  for (size_t i = 1; i < ords.size(); ++i)
    if (ords[i] != ords[i - 1] + 1) return TALSH_INVALID_ARGS;  // consecutive, as the reference requires
  int have = 0;
  if (!ords.empty()) {
    if (hipGetDeviceCount(&have) != hipSuccess) { (void)hipGetLastError(); have = 0; }
    for (int o : ords)
      if (o < 0 || o >= have) return TALSH_INVALID_ARGS;
  }
  const char *env_bytes = std::getenv("TALSH_GPU_ARENA_BYTES");
  const char *env_frac = std::getenv("TALSH_GPU_MEM_FRACTION");
  for (size_t i = 0; i < ords.size(); ++i) {
    Device *d = new Device();
    d->ordinal = ords[i];
    d->flat_id = gpu_flat_id((int)i);
    if (!XT_HIP_OK(hipSetDevice(d->ordinal))) { delete d; return TALSH_FAILURE; }
    hipDeviceProp_t prop;
    if (!XT_HIP_OK(hipGetDeviceProperties(&prop, d->ordinal))) { delete d; return TALSH_FAILURE; }
    d->num_cus = prop.multiProcessorCount;
    size_t fr = 0, tot = 0;
    if (!XT_HIP_OK(hipMemGetInfo(&fr, &tot))) { delete d; return TALSH_FAILURE; }
    d->mem_total = tot;
    size_t want = (size_t)((double)fr * (env_frac ? std::atof(env_frac) : 0.5));
    if (env_bytes) want = (size_t)std::strtoull(env_bytes, nullptr, 10);
    want = want / Arena::kAlign * Arena::kAlign;
    while (want >= (64ull << 20)) {
      if (hipMalloc(&d->arena_mem, want) == hipSuccess) break;
      (void)hipGetLastError();
      d->arena_mem = nullptr;
      want = want / 10 * 9 / Arena::kAlign * Arena::kAlign;
    }
    if (!d->arena_mem) { delete d; return TALSH_FAILURE; }
    d->arena.reset(d->arena_mem, want);
    for (int k = 0; k < kStreamsPerDevice; ++k)
      if (!XT_HIP_OK(hipStreamCreateWithFlags(&d->streams[k], hipStreamNonBlocking))) { delete d; return TALSH_FAILURE; }
    // this GPU's address of the allocation-free product's scratch array (device current)
    if (gemm_noalloc_scratch(&d->noalloc_scratch))
      log_error("GPU %d: scratch of talshAmdGemmNoAlloc not found; that entry point will fail", d->ordinal);
    E.gpus.push_back(d);
  }
  if (!E.gpus.empty()) (void)hipSetDevice(E.gpus[0]->ordinal);
  if (host_arg_max) *host_arg_max = 1 << 20;
  (void)host_buf_size;  // host images are allocated per tensor (pinned when a GPU is present)
  E.on = true;
  {  // learned plan selector: TALSH_PLAN_MODEL, else the model shipped beside the library
    std::string path;
    const char *pm = std::getenv("TALSH_PLAN_MODEL");
    bool must = false;
    if (pm) {
      path = pm;
      must = !path.empty();
    } else {
      Dl_info info;
      if (dladdr(reinterpret_cast<void *>(&talshInit), &info) && info.dli_fname) {
        path = info.dli_fname;
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/plan_model.txt";
        FILE *f = std::fopen(path.c_str(), "r");
        if (f) std::fclose(f);
        else path.clear();
      }
    }
    void *h = nullptr;
    if (!path.empty() && talshAmdMlpLoad(path.c_str(), &h) == 0) {
      talshAmdSelectorInstall(h);
      talshAmdMlpDestroy(h);
    } else if (must) {
      log_error("TALSH_PLAN_MODEL=%s could not be loaded; contractions use the byte-count heuristic", pm);
    }
  }
  return TALSH_SUCCESS;
}

int talshShutdown()
{
  Engine &E = engine();
  std::lock_guard<std::mutex> g(E.mu);
  if (!E.on) return TALSH_NOT_INITIALIZED;
  int err = 0;
  for (Device *d : E.gpus) {
    (void)hipSetDevice(d->ordinal);
    if (hipDeviceSynchronize() != hipSuccess) err = TALSH_FAILURE;
  }
  image_forget_all();
  event_pool_clear();
  for (Device *d : E.gpus) {
    for (auto &s : d->streams)
      if (s) (void)hipStreamDestroy(s);
    if (d->arena.live_blocks() > 0) err = err ? err : NOT_CLEAN;
    if (d->arena_mem) (void)hipFree(d->arena_mem);
    delete d;
  }
  E.gpus.clear();
  E.on = false;
  return err;
}

void talshSetMemAllocPolicyHost(int mem_policy, int fallback, int *ierr)
{
  Engine &E = engine();
  if (mem_policy != MEM_ALLOC_REGULAR && mem_policy != MEM_ALLOC_TMP_BUF && mem_policy != MEM_ALLOC_ALL_BUF) {
    if (ierr) *ierr = TALSH_INVALID_ARGS;
    return;
  }
  E.host_mem_policy = mem_policy;
  E.host_mem_fallback = fallback;
  if (ierr) *ierr = 0;
}

static int fast_math_target(int dev_kind, int dev_id, std::vector<Device *> &out)
{
  Engine &E = engine();
  if (!E.on) return TALSH_NOT_INITIALIZED;
  if (dev_kind != DEV_AMD_GPU && dev_kind != DEV_NVIDIA_GPU) return TALSH_NOT_AVAILABLE;
  if (dev_id == DEV_DEFAULT) out = E.gpus;
  else if (dev_id >= 0 && dev_id < (int)E.gpus.size()) out.push_back(E.gpus[dev_id]);
  else return TALSH_INVALID_ARGS;
  return 0;
}

int talshEnableFastMath(int dev_kind, int dev_id)
{
  std::vector<Device *> v;
  int e = fast_math_target(dev_kind, dev_id, v);
  if (e) return e;
  for (Device *d : v) d->r4mode = R4Mode::kBF16x3;
  return 0;
}

int talshDisableFastMath(int dev_kind, int dev_id)
{
  std::vector<Device *> v;
  int e = fast_math_target(dev_kind, dev_id, v);
  if (e) return e;
  for (Device *d : v) d->r4mode = R4Mode::kNativeF32;
  return 0;
}

int talshQueryFastMath(int dev_kind, int dev_id)
{
  std::vector<Device *> v;
  if (fast_math_target(dev_kind, dev_id, v) || v.empty()) return NOPE;
  return v[0]->r4mode == R4Mode::kBF16x3 ? YEP : NOPE;
}

int talshDeviceCount(int dev_kind, int *dev_count)
{
  if (!dev_count) return TALSH_INVALID_ARGS;
  *dev_count = 0;
  switch (dev_kind) {
  case DEV_HOST: *dev_count = 1; return 0;
  case DEV_AMD_GPU:
  case DEV_NVIDIA_GPU: {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) { (void)hipGetLastError(); n = 0; }
    *dev_count = std::min(n, MAX_AMDS_PER_NODE);
    return 0;
  }
  case DEV_INTEL_MIC: return 0;
  default: return TALSH_INVALID_ARGS;
  }
}

int talshFlatDevId(int dev_kind, int dev_num) { return encode_device_id(dev_kind, dev_num); }

int talshKindDevId(int dev_id, int *dev_kind)
{
  int k = DEV_NULL;
  const int n = decode_device_id(dev_id, &k);
  if (dev_kind) *dev_kind = k;
  return n;
}

int talshDeviceState(int dev_num, int dev_kind)
{
  const int flat = dev_kind == DEV_NULL ? canonical_dev(dev_num) : resolve_device(dev_num, dev_kind);
  if (flat < 0) return DEV_OFF;
  return DEV_ON;
}
int talshDeviceState_(int dev_num, int dev_kind) { return talshDeviceState(dev_num, dev_kind); }

int talshDeviceBusyLeast(int dev_kind)
{
  Engine &E = engine();
  if (dev_kind == DEV_HOST) return 0;
  if (E.gpus.empty()) return dev_kind == DEV_NULL ? 0 : DEV_NULL;
  size_t best = 0;
  int bi = 0;
  for (size_t i = 0; i < E.gpus.size(); ++i) {
    const size_t f = E.gpus[i]->arena.free_bytes();
    if (f > best) { best = f; bi = (int)i; }
  }
  return dev_kind == DEV_NULL ? E.gpus[bi]->flat_id : bi;
}
int talshDeviceBusyLeast_(int dev_kind) { return talshDeviceBusyLeast(dev_kind); }

int talshDetermineOptimalDevice(const talsh_tens_t *t0, const talsh_tens_t *t1, const talsh_tens_t *t2)
{
  // Operations run on GPUs: prefer the GPU that already holds the most bytes.
  Engine &E = engine();
  if (E.gpus.empty()) return DEV_NULL;
  std::vector<double> score(E.gpus.size(), 0.0);
  for (const talsh_tens_t *t : {t0, t1, t2}) {
    if (!t || !t->shape_p) continue;
    for (int i = 0; i < t->ndev; ++i) {
      int k = DEV_NULL;
      const int n = decode_device_id(t->dev_rsc[i].dev_id, &k);
      if ((k == DEV_AMD_GPU || k == DEV_NVIDIA_GPU) && n >= 0 && n < (int)score.size())
        score[n] += (double)tensor_volume(t) * kind_size(t->data_kind[i]);
    }
  }
  int best = 0;
  for (size_t i = 1; i < score.size(); ++i)
    if (score[i] > score[best]) best = (int)i;
  return E.gpus[best]->flat_id;
}

static Device *device_of(int dev_num, int dev_kind)
{
  const int flat = dev_kind == DEV_NULL ? canonical_dev(dev_num) : resolve_device(dev_num, dev_kind);
  return flat > 0 ? gpu_by_flat(flat) : nullptr;
}

size_t talshDeviceMemorySize(int dev_num, int dev_kind)
{
  Device *d = device_of(dev_num, dev_kind);
  return d ? d->mem_total : 0;
}
size_t talshDeviceMemorySize_(int dev_num, int dev_kind) { return talshDeviceMemorySize(dev_num, dev_kind); }
size_t talshDeviceBufferSize(int dev_num, int dev_kind)
{
  Device *d = device_of(dev_num, dev_kind);
  return d ? d->arena.capacity() : 0;
}
size_t talshDeviceBufferSize_(int dev_num, int dev_kind) { return talshDeviceBufferSize(dev_num, dev_kind); }
size_t talshDeviceTensorSize(int dev_num, int dev_kind)
{
  Device *d = device_of(dev_num, dev_kind);
  return d ? d->arena.largest_free() : 0;
}
size_t talshDeviceTensorSize_(int dev_num, int dev_kind) { return talshDeviceTensorSize(dev_num, dev_kind); }
size_t talshDeviceBufferFreeSize(int dev_num, int dev_kind)
{
  Device *d = device_of(dev_num, dev_kind);
  return d ? d->arena.free_bytes() : 0;
}
size_t talshDeviceBufferFreeSize_(int dev_num, int dev_kind) { return talshDeviceBufferFreeSize(dev_num, dev_kind); }
void *talshDeviceBufferBasePtr(int dev_num, int dev_kind)
{
  Device *d = device_of(dev_num, dev_kind);
  return d ? d->arena.base() : nullptr;
}
void *talshDeviceBufferBasePtr_(int dev_num, int dev_kind) { return talshDeviceBufferBasePtr(dev_num, dev_kind); }

double talshDeviceGetFlops(int dev_kind, int dev_id)
{
  Engine &E = engine();
  double f = 0;
  for (size_t i = 0; i < E.gpus.size(); ++i) {
    if ((dev_kind == DEV_DEFAULT || dev_kind == DEV_AMD_GPU || dev_kind == DEV_NVIDIA_GPU) &&
        (dev_id == DEV_DEFAULT || dev_id == (int)i))
      f += E.gpus[i]->stats.flops;
  }
  return f;
}

void talshMemManagerLogStart() { engine().mem_log = true; }
void talshMemManagerLogFinish() { engine().mem_log = false; }
void talshTensorOpLogStart() { engine().op_log = true; }
void talshTensorOpLogFinish() { engine().op_log = false; }

int talshStats(int dev_id, int dev_kind)
{
  Engine &E = engine();
  if (!E.on) return TALSH_NOT_INITIALIZED;
  for (size_t i = 0; i < E.gpus.size(); ++i) {
    if (dev_id >= 0 && dev_id != (int)i) continue;
    (void)dev_kind;
    const DeviceStats &s = E.gpus[i]->stats;
    std::printf("#MSG(TAL-SH MI355X): GPU %zu: tasks submitted %lld completed %lld deferred %lld failed %lld; "
                "flops %.6e; in %.6e B; out %.6e B; arena free %zu / %zu B\n",
                i, s.tasks_submitted, s.tasks_completed, s.tasks_deferred, s.tasks_failed, s.flops, s.traffic_in,
                s.traffic_out, E.gpus[i]->arena.free_bytes(), E.gpus[i]->arena.capacity());
  }
  return 0;
}
int talshStats_(int dev_id, int dev_kind) { return talshStats(dev_id, dev_kind); }

// ============================================================================ tensors
int talshTensorCreate(talsh_tens_t **tens_block)
{
  if (!tens_block) return TALSH_INVALID_ARGS;
  *tens_block = (talsh_tens_t *)std::malloc(sizeof(talsh_tens_t));
  if (!*tens_block) return TRY_LATER;
  return talshTensorClean(*tens_block);
}

int talshTensorClean(talsh_tens_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  t->shape_p = nullptr;
  t->dev_rsc = nullptr;
  t->data_kind = nullptr;
  t->avail = nullptr;
  t->dev_rsc_len = 0;
  t->ndev = 0;
  return 0;
}

int talshTensorIsEmpty(const talsh_tens_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  return t->shape_p == nullptr ? YEP : NOPE;
}

int talshTensorConstruct(talsh_tens_t *t, int data_kind, int tens_rank, const int tens_dims[], int dev_id,
                         void *ext_mem, int in_hab, talsh_tens_init_i init_method, double init_val_real,
                         double init_val_imag)
{
  Engine &E = engine();
  if (!E.on) return TALSH_NOT_INITIALIZED;
  if (!t) return TALSH_INVALID_ARGS;
  if (talshTensorIsEmpty(t) != YEP) return TALSH_OBJECT_NOT_EMPTY;
  int es = 0;
  if (tens_valid_data_kind(data_kind, &es) != YEP) return TALSH_INVALID_ARGS;
  if (tens_rank < 0 || tens_rank > MAX_TENSOR_RANK) return TALSH_INVALID_ARGS;
  const int flat = canonical_dev(dev_id);
  if (flat < 0) return TALSH_INVALID_ARGS;
  (void)in_hab;
  int e = tensShape_create(&t->shape_p);
  if (e) { talshTensorClean(t); return e; }
  e = tensShape_construct(t->shape_p, NOPE, tens_rank, tens_dims, nullptr, nullptr);
  if (e) { tensShape_destroy(t->shape_p); talshTensorClean(t); return e; }
  t->dev_rsc_len = TALSH_MAX_DEV_PRESENT;
  t->dev_rsc = (talsh_dev_rsc_t *)std::calloc(TALSH_MAX_DEV_PRESENT, sizeof(talsh_dev_rsc_t));
  t->data_kind = (int *)std::calloc(TALSH_MAX_DEV_PRESENT, sizeof(int));
  t->avail = (int *)std::calloc(TALSH_MAX_DEV_PRESENT, sizeof(int));
  t->ndev = 0;
  if (!t->dev_rsc || !t->data_kind || !t->avail) { talshTensorDestruct(t); return TRY_LATER; }
  if (data_kind == NO_TYPE) return 0;  // shape only
  const size_t vol = tensor_volume(t);
  talsh_dev_rsc_t r;
  if (ext_mem) {
    r.dev_id = flat; r.gmem_p = ext_mem; r.buf_entry = -1; r.mem_attached = 1;
    tensor_add_image(t, r, data_kind);
    return 0;  // external memory is not initialized
  }
  const size_t bytes = vol * (size_t)es;
  e = image_alloc(flat, bytes, &r);
  if (e) { talshTensorDestruct(t); return e; }
  tensor_add_image(t, r, data_kind);
  // initialization
  if (flat == 0) {
    if (init_method) {
      talsh_tens_data_t td{r.gmem_p, vol, data_kind};
      if (init_method(&td, t->shape_p, nullptr)) return NOT_CLEAN;
      return 0;
    }
    switch (data_kind) {
    case R4: { float *p = (float *)r.gmem_p; const float v = (float)init_val_real; for (size_t i = 0; i < vol; ++i) p[i] = v; break; }
    case R8: { double *p = (double *)r.gmem_p; for (size_t i = 0; i < vol; ++i) p[i] = init_val_real; break; }
    case C4: { float *p = (float *)r.gmem_p; for (size_t i = 0; i < vol; ++i) { p[2 * i] = (float)init_val_real; p[2 * i + 1] = (float)init_val_imag; } break; }
    case C8: { double *p = (double *)r.gmem_p; for (size_t i = 0; i < vol; ++i) { p[2 * i] = init_val_real; p[2 * i + 1] = init_val_imag; } break; }
    }
    return 0;
  }
  if (init_method) return NOT_CLEAN;  // device bodies cannot run a host initializer in place
  e = talshTensorInit(t, init_val_real, init_val_imag, flat, DEV_NULL, COPY_M, nullptr);
  return e ? NOT_CLEAN : 0;
}

int talshTensorConstruct_(talsh_tens_t *t, int data_kind, int tens_rank, const int tens_dims[], int dev_id,
                          void *ext_mem, int in_hab, talsh_tens_init_i init_method, double init_val_real,
                          double init_val_imag)
{
  return talshTensorConstruct(t, data_kind, tens_rank, tens_dims, dev_id, ext_mem, in_hab, init_method, init_val_real,
                              init_val_imag);
}

int talshTensorImportData(talsh_tens_t *t, int data_kind, const void *ext_data)
{
  if (!engine().on) return TALSH_NOT_INITIALIZED;
  if (!t || !ext_data || talshTensorIsEmpty(t) == YEP) return TALSH_INVALID_ARGS;
  int idx = -1;
  for (int k = 0; k < t->ndev && idx < 0; ++k)
    if (t->data_kind[k] == data_kind) idx = k;
  if (idx < 0) return TALSH_NOT_FOUND;
  const talsh_dev_rsc_t keep = t->dev_rsc[idx];
  // every other image becomes stale
  for (int k = t->ndev - 1; k >= 0; --k)
    if (t->dev_rsc[k].gmem_p != keep.gmem_p) tensor_remove_image(t, k, nullptr);
  const size_t bytes = tensor_volume(t) * (size_t)kind_size(data_kind);
  if (keep.dev_id == 0) {
    std::memcpy(keep.gmem_p, ext_data, bytes);
    return 0;
  }
  Device *d = gpu_by_flat(keep.dev_id);
  if (!d) return TALSH_FAILURE;
  (void)hipSetDevice(d->ordinal);
  for (auto s : d->streams) (void)hipStreamSynchronize(s);
  return XT_HIP_OK(hipMemcpy(keep.gmem_p, ext_data, bytes, hipMemcpyHostToDevice)) ? 0 : TALSH_FAILURE;
}

int talshTensorDestruct(talsh_tens_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  int err = 0;
  if (t->shape_p) {
    // wait for outstanding device work on the images before releasing them
    for (int i = 0; i < t->ndev; ++i) {
      if (t->dev_rsc[i].dev_id != 0)
        if (Device *d = gpu_by_flat(t->dev_rsc[i].dev_id)) {
          (void)hipSetDevice(d->ordinal);
          for (auto s : d->streams) (void)hipStreamSynchronize(s);
        }
    }
    while (t->ndev > 0) tensor_remove_image(t, t->ndev - 1, nullptr);
    tensShape_destroy(t->shape_p);
  }
  std::free(t->dev_rsc);
  std::free(t->data_kind);
  std::free(t->avail);
  talshTensorClean(t);
  return err;
}

int talshTensorDestroy(talsh_tens_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  const int e = talshTensorDestruct(t);
  std::free(t);
  return e;
}

int talshTensorRank(const talsh_tens_t *t)
{
  if (!t || !t->shape_p) return -1;
  return t->shape_p->num_dim;
}

size_t talshTensorVolume(const talsh_tens_t *t)
{
  if (!t || !t->shape_p) return 0;
  return tensor_volume(t);
}

size_t talshTensorSizeAllImages(const talsh_tens_t *t, int *num_images)
{
  if (num_images) *num_images = 0;
  if (!t || !t->shape_p) return 0;
  size_t s = 0;
  for (int i = 0; i < t->ndev; ++i) s += tensor_volume(t) * (size_t)kind_size(t->data_kind[i]);
  if (num_images) *num_images = t->ndev;
  return s;
}

const int *talshTensorDimExtents(const talsh_tens_t *t, int *rank)
{
  if (!t || !t->shape_p) { if (rank) *rank = -1; return nullptr; }
  if (rank) *rank = t->shape_p->num_dim;
  return t->shape_p->dims;
}

int talshTensorShape(const talsh_tens_t *t, talsh_tens_shape_t *s)
{
  if (!t || !s || !t->shape_p) return TALSH_INVALID_ARGS;
  if (s->num_dim >= 0) tensShape_destruct(s);
  tensShape_clean(s);
  return tensShape_construct(s, NOPE, t->shape_p->num_dim, t->shape_p->dims, t->shape_p->divs, t->shape_p->grps);
}

int talshTensorDataKind(const talsh_tens_t *t, int *num_images, int *data_kinds)
{
  if (!t || !num_images || !t->shape_p) return TALSH_INVALID_ARGS;
  *num_images = t->ndev;
  for (int i = 0; i < t->ndev; ++i)
    if (data_kinds) data_kinds[i] = t->data_kind[i];
  return 0;
}

int talshTensorReshape(talsh_tens_t *t, int tens_rank, const int tens_dims[])
{
  if (!t || !t->shape_p) return TALSH_INVALID_ARGS;
  if (talshTensorInUse(t) == YEP) return TALSH_IN_PROGRESS;
  return tensShape_reshape(t->shape_p, tens_rank, tens_dims, nullptr, nullptr);
}

int talshTensorInUse(const talsh_tens_t *t)
{
  if (!t || !t->shape_p) return NOPE;
  for (int i = 0; i < t->ndev; ++i)
    if (t->avail[i] != YEP) return YEP;
  return NOPE;
}

int talshTensorPresence(const talsh_tens_t *t, int *ncopies, int copies[], int data_kinds[], int dev_kind, int dev_id)
{
  if (!t || !ncopies || !t->shape_p) return TALSH_INVALID_ARGS;
  *ncopies = 0;
  for (int i = 0; i < t->ndev; ++i) {
    int k = DEV_NULL;
    const int n = decode_device_id(t->dev_rsc[i].dev_id, &k);
    const bool kind_ok = dev_kind == DEV_NULL || k == dev_kind ||
                         ((dev_kind == DEV_NVIDIA_GPU || dev_kind == DEV_AMD_GPU) &&
                          (k == DEV_NVIDIA_GPU || k == DEV_AMD_GPU));
    if (!kind_ok || (dev_id >= 0 && n != dev_id)) continue;
    if (copies) copies[*ncopies] = t->dev_rsc[i].dev_id;
    if (data_kinds) data_kinds[*ncopies] = t->data_kind[i];
    ++*ncopies;
  }
  return 0;
}
int talshTensorPresence_(const talsh_tens_t *t, int *ncopies, int copies[], int data_kinds[], int dev_kind, int dev_id)
{
  return talshTensorPresence(t, ncopies, copies, data_kinds, dev_kind, dev_id);
}

int talshTensorGetBodyAccess(talsh_tens_t *t, void **body_p, int data_kind, int dev_id, int dev_kind)
{
  if (!t || !body_p || !t->shape_p) return TALSH_INVALID_ARGS;
  const int flat = dev_kind == DEV_NULL ? canonical_dev(dev_id) : resolve_device(dev_id, dev_kind);
  if (flat < 0) return TALSH_INVALID_ARGS;
  const int i = find_image(t, flat, data_kind);
  if (i < 0) return TALSH_NOT_FOUND;
  if (t->dev_rsc[i].dev_id != 0) {
    // device work on the body must be complete before the caller touches it
    if (Device *d = gpu_by_flat(t->dev_rsc[i].dev_id)) {
      (void)hipSetDevice(d->ordinal);
      for (auto s : d->streams) (void)hipStreamSynchronize(s);
    }
  }
  *body_p = t->dev_rsc[i].gmem_p;
  return 0;
}
int talshTensorGetBodyAccess_(talsh_tens_t *t, void **body_p, int data_kind, int dev_id, int dev_kind)
{
  return talshTensorGetBodyAccess(t, body_p, data_kind, dev_id, dev_kind);
}

int talshTensorGetBodyAccessConst(const talsh_tens_t *t, const void **body_p, int data_kind, int dev_id, int dev_kind)
{
  void *p = nullptr;
  const int e = talshTensorGetBodyAccess(const_cast<talsh_tens_t *>(t), &p, data_kind, dev_id, dev_kind);
  if (body_p) *body_p = p;
  return e;
}

int talshTensorGetScalar(talsh_tens_t *t, double *re, double *im)
{
  if (!t || !re || !im || !t->shape_p || t->ndev <= 0) return TALSH_INVALID_ARGS;
  if (tensor_volume(t) != 1) return TALSH_INVALID_REQUEST;
  const talsh_dev_rsc_t &r = t->dev_rsc[0];
  double buf[2] = {0, 0};
  const int es = kind_size(t->data_kind[0]);
  if (r.dev_id == 0) std::memcpy(buf, r.gmem_p, es);
  else {
    Device *d = gpu_by_flat(r.dev_id);
    if (!d) return TALSH_FAILURE;
    (void)hipSetDevice(d->ordinal);
    for (auto s : d->streams) (void)hipStreamSynchronize(s);
    if (!XT_HIP_OK(hipMemcpy(buf, r.gmem_p, es, hipMemcpyDeviceToHost))) return TALSH_FAILURE;
  }
  switch (t->data_kind[0]) {
  case R4: *re = ((float *)buf)[0]; *im = 0; break;
  case R8: *re = buf[0]; *im = 0; break;
  case C4: *re = ((float *)buf)[0]; *im = ((float *)buf)[1]; break;
  case C8: *re = buf[0]; *im = buf[1]; break;
  default: return TALSH_FAILURE;
  }
  return 0;
}

void talshTensorPrint(const talsh_tens_t *t)
{
  if (!t || !t->shape_p) { std::printf("#MESSAGE: Empty tensor\n"); return; }
  std::printf("{");
  for (int q = 0; q < t->shape_p->num_dim; ++q) std::printf(q ? ",%d" : "%d", t->shape_p->dims[q]);
  std::printf("}");
}

void talshTensorPrintInfo(const talsh_tens_t *t)
{
  if (!t || !t->shape_p) { std::printf("#MESSAGE: Empty tensor\n"); return; }
  std::printf("#MESSAGE: Tensor block info: rank %d, volume %zu, images %d\n", t->shape_p->num_dim, tensor_volume(t),
              t->ndev);
  for (int i = 0; i < t->ndev; ++i)
    std::printf("  image %d: device %d, kind %d, avail %d, ptr %p\n", i, t->dev_rsc[i].dev_id, t->data_kind[i],
                t->avail[i], t->dev_rsc[i].gmem_p);
}

void talshTensorPrintBody(const talsh_tens_t *t, double thresh)
{
  if (!t || !t->shape_p) return;
  for (int i = 0; i < t->ndev; ++i) {
    if (t->dev_rsc[i].dev_id != 0) continue;
    const size_t vol = tensor_volume(t);
    for (size_t k = 0; k < vol; ++k) {
      double re = 0, im = 0;
      switch (t->data_kind[i]) {
      case R4: re = ((float *)t->dev_rsc[i].gmem_p)[k]; break;
      case R8: re = ((double *)t->dev_rsc[i].gmem_p)[k]; break;
      case C4: re = ((float *)t->dev_rsc[i].gmem_p)[2 * k]; im = ((float *)t->dev_rsc[i].gmem_p)[2 * k + 1]; break;
      case C8: re = ((double *)t->dev_rsc[i].gmem_p)[2 * k]; im = ((double *)t->dev_rsc[i].gmem_p)[2 * k + 1]; break;
      }
      if (std::sqrt(re * re + im * im) > thresh) std::printf("  [%zu] %.15e %.15e\n", k, re, im);
    }
    return;
  }
}

double talshTensorImageNorm1_cpu(const talsh_tens_t *t)
{
  if (!t || !t->shape_p) return -1.0;
  for (int i = 0; i < t->ndev; ++i) {
    if (t->dev_rsc[i].dev_id != 0) continue;
    const size_t vol = tensor_volume(t);
    double s = 0;
    for (size_t k = 0; k < vol; ++k) {
      switch (t->data_kind[i]) {
      case R4: s += std::fabs(((float *)t->dev_rsc[i].gmem_p)[k]); break;
      case R8: s += std::fabs(((double *)t->dev_rsc[i].gmem_p)[k]); break;
      case C4: s += std::hypot(((float *)t->dev_rsc[i].gmem_p)[2 * k], ((float *)t->dev_rsc[i].gmem_p)[2 * k + 1]); break;
      case C8: s += std::hypot(((double *)t->dev_rsc[i].gmem_p)[2 * k], ((double *)t->dev_rsc[i].gmem_p)[2 * k + 1]); break;
      }
    }
    return s;
  }
  return -1.0;
}

// ============================================================================ slices (not on this engine's path)
int talshTensorSliceCreate(talsh_tens_slice_t **slice)
{
  if (!slice) return TALSH_INVALID_ARGS;
  *slice = (talsh_tens_slice_t *)std::malloc(sizeof(talsh_tens_slice_t));
  if (!*slice) return TRY_LATER;
  return talshTensorSliceClean(*slice);
}
int talshTensorSliceClean(talsh_tens_slice_t *slice)
{
  if (!slice) return TALSH_INVALID_ARGS;
  slice->tensor = nullptr;
  tensSignature_clean(&slice->bases);
  tensShape_clean(&slice->shape);
  return 0;
}
int talshTensorSliceConstruct(talsh_tens_slice_t *slice, const talsh_tens_t *tensor, const size_t *offsets,
                              const int *dims, const int *divs, const int *grps)
{
  if (!slice || !tensor || !tensor->shape_p) return TALSH_INVALID_ARGS;
  const int rank = tensor->shape_p->num_dim;
  for (int q = 0; q < rank; ++q)
    if (!offsets || !dims || offsets[q] + (size_t)dims[q] > (size_t)tensor->shape_p->dims[q]) return TALSH_INVALID_ARGS;
  int e = tensSignature_construct(&slice->bases, rank, offsets);
  if (e) return e;
  e = tensShape_construct(&slice->shape, NOPE, rank, dims, divs, grps);
  if (e) { tensSignature_destruct(&slice->bases); return e; }
  slice->tensor = const_cast<talsh_tens_t *>(tensor);
  return 0;
}
size_t talshTensorSliceVolume(const talsh_tens_slice_t *slice) { return slice ? tensShape_volume(&slice->shape) : 0; }
int talshTensorSliceDestruct(talsh_tens_slice_t *slice)
{
  if (!slice) return TALSH_INVALID_ARGS;
  tensSignature_destruct(&slice->bases);
  tensShape_destruct(&slice->shape);
  slice->tensor = nullptr;
  return 0;
}
int talshTensorSliceDestroy(talsh_tens_slice_t *slice)
{
  if (!slice) return TALSH_INVALID_ARGS;
  talshTensorSliceDestruct(slice);
  std::free(slice);
  return 0;
}

}  // extern "C"

// ============================================================================ device resources
extern "C" {

int tensDevRsc_create(talsh_dev_rsc_t **drsc)
{
  if (!drsc) return TALSH_INVALID_ARGS;
  *drsc = (talsh_dev_rsc_t *)std::malloc(sizeof(talsh_dev_rsc_t));
  if (!*drsc) return TRY_LATER;
  return tensDevRsc_clean(*drsc);
}

int tensDevRsc_clean(talsh_dev_rsc_t *r)
{
  if (!r) return TALSH_INVALID_ARGS;
  r->dev_id = DEV_NULL;
  r->gmem_p = nullptr;
  r->buf_entry = -1;
  r->mem_attached = 0;
  return 0;
}

int tensDevRsc_is_empty(talsh_dev_rsc_t *r)
{
  if (!r) return TALSH_INVALID_ARGS;
  return (r->dev_id < 0 && r->gmem_p == nullptr) ? YEP : NOPE;
}

int tensDevRsc_same(const talsh_dev_rsc_t *a, const talsh_dev_rsc_t *b)
{
  if (!a || !b) return NOPE;
  return (a->dev_id == b->dev_id && a->gmem_p == b->gmem_p) ? YEP : NOPE;
}

int tensDevRsc_clone(const talsh_dev_rsc_t *in, talsh_dev_rsc_t *out)
{
  if (!in || !out) return TALSH_INVALID_ARGS;
  *out = *in;
  return 0;
}

int tensDevRsc_attach_mem(talsh_dev_rsc_t *r, int dev_id, void *mem_p, int buf_entry)
{
  if (!r || !mem_p || dev_id < 0) return TALSH_INVALID_ARGS;
  if (tensDevRsc_is_empty(r) != YEP) return TALSH_OBJECT_NOT_EMPTY;
  r->dev_id = dev_id;
  r->gmem_p = mem_p;
  r->buf_entry = buf_entry;
  r->mem_attached = 1;
  return 0;
}

int tensDevRsc_detach_mem(talsh_dev_rsc_t *r)
{
  if (!r || !r->mem_attached) return TALSH_INVALID_ARGS;
  return tensDevRsc_clean(r);
}

int tensDevRsc_allocate_mem(talsh_dev_rsc_t *r, int dev_id, size_t mem_size, int in_arg_buf)
{
  (void)in_arg_buf;
  if (!r) return TALSH_INVALID_ARGS;
  if (tensDevRsc_is_empty(r) != YEP) return TALSH_OBJECT_NOT_EMPTY;
  const int flat = canonical_dev(dev_id);
  if (flat < 0) return TALSH_INVALID_ARGS;
  const int e = image_alloc(flat, mem_size, r);
  if (e) tensDevRsc_clean(r);
  return e;
}

int tensDevRsc_free_mem(talsh_dev_rsc_t *r)
{
  if (!r || r->mem_attached) return TALSH_INVALID_ARGS;
  image_release(*r, nullptr);
  return tensDevRsc_clean(r);
}

int tensDevRsc_get_gmem_ptr(talsh_dev_rsc_t *r, void **p)
{
  if (!r || !p || tensDevRsc_is_empty(r) == YEP) return TALSH_INVALID_ARGS;
  *p = r->gmem_p;
  return 0;
}

int tensDevRsc_device_id(talsh_dev_rsc_t *r) { return r ? r->dev_id : DEV_NULL; }

int tensDevRsc_release_all(talsh_dev_rsc_t *r)
{
  if (!r) return TALSH_INVALID_ARGS;
  if (r->mem_attached) return tensDevRsc_clean(r);
  if (r->gmem_p) image_release(*r, nullptr);
  return tensDevRsc_clean(r);
}

int tensDevRsc_destroy(talsh_dev_rsc_t *r)
{
  if (!r) return TALSH_INVALID_ARGS;
  tensDevRsc_release_all(r);
  std::free(r);
  return 0;
}

}  // extern "C"
