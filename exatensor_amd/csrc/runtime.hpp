// Engine runtime: per-GPU state (arena, stream pool, statistics), engine task
// objects, and the device-kernel launchers used by the tensor operations.
#pragma once

#include <atomic>
#include <mutex>
#include <vector>

#include "arena.hpp"
#include "common.hpp"

namespace xt {

constexpr int kStreamsPerDevice = 4;  // matches the box's hardware queues per process

struct DeviceStats {
  long long tasks_submitted = 0, tasks_completed = 0, tasks_deferred = 0, tasks_failed = 0;
  double flops = 0, traffic_in = 0, traffic_out = 0, time_active = 0;
};

// Precision policy for single-precision (R4/C4) contractions.
enum class R4Mode : int { kNativeF32 = 0, kBF16x3 = 1 };

struct Device {
  int ordinal = -1;                  // HIP device ordinal
  int flat_id = -1;                  // TAL-SH flat id (DEV_AMD_GPU numbering)
  int num_cus = 0;
  size_t mem_total = 0;
  void *arena_mem = nullptr;
  Arena arena;
  hipStream_t streams[kStreamsPerDevice] = {};
  std::atomic<unsigned> next_stream{0};
  DeviceStats stats;
  R4Mode r4mode = R4Mode::kNativeF32;
  void *noalloc_scratch = nullptr;   // this GPU's copy of the allocation-free product's scratch array
};

struct Engine {
  bool on = false;
  std::vector<Device *> gpus;        // index = kind-specific GPU number
  std::mutex mu;
  int host_mem_policy = MEM_ALLOC_TMP_BUF;
  int host_mem_fallback = 1;
  bool op_log = false, mem_log = false;
  int profiling = 0;
};

Engine &engine();
Device *gpu_by_flat(int flat_id);
// Flat id of a GPU given a kind-specific request; DEV_NVIDIA_GPU is an alias.
int gpu_flat_id(int dev_num);
bool is_gpu_flat(int flat_id);

// ---------------------------------------------------------------------------
// Engine task (what talsh_task_t::task_p points to).
struct Image {                         // one body image of a tensor
  int dev_id;
  void *ptr;
  int handle;                          // arena handle, -1 for host or attached memory
  int attached;
  int kind;
};

struct PendingFree { int dev_flat; int handle; void *host_ptr; };

struct Task {
  int dev_flat = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_comput = nullptr, ev_output = nullptr, ev_finish = nullptr, ev_mm0 = nullptr,
             ev_mm1 = nullptr;
  bool has_mm = false;
  bool timed = true;  // start/input/output/GEMM events recorded (a blocking call's task records only its finish)
  int status = TALSH_TASK_EMPTY;
  bool finalized = false;
  std::vector<PendingFree> frees;      // temporaries released at completion
  // coherence actions applied at completion, one entry per argument; images
  // are identified by their body pointers (indices shift as images come and go)
  struct ArgFinal {
    talsh_tens_t *tens = nullptr;
    void *src = nullptr, *exec = nullptr;
    unsigned coh = 1;
    bool output = false, created = false;
  };
  std::vector<ArgFinal> finals;
  double flops = 0;
};

int task_finalize(talsh_task_t *t, int status);

// Pooled HIP events of the current device (timing or not); put returns one
// for reuse once nothing will query it again.
hipEvent_t event_get(bool timing);
void event_put(hipEvent_t e, bool timing);
void event_pool_clear();

// ---------------------------------------------------------------------------
// Profiling of the dominant kernels (event pairs on the launch stream).
// Redirects this thread's kernel timings into sink (nullptr: back to the
// engine's profile); prof_sink_drain returns their summed milliseconds (-1 if
// an event failed) and returns the events to the pool.
void prof_set_sink(std::vector<std::pair<hipEvent_t, hipEvent_t>> *sink);
double prof_sink_drain(std::vector<std::pair<hipEvent_t, hipEvent_t>> &sink);
void prof_begin(hipStream_t s, int kernel_id, hipEvent_t *pair);
void prof_end(hipStream_t s, int kernel_id, hipEvent_t *pair);
enum KernelId { kKTranspose = 0, kKGemm = 1, kKGrouped = 2, kKElementwise = 3, kKNum = 4 };

}  // namespace xt
