// Internal helpers shared by the C-ABI translation units.
#pragma once

#include "runtime.hpp"

namespace xt {

int canonical_dev(int flat);
int resolve_device(int dev_id, int dev_kind);
void *host_alloc(size_t bytes);
void host_free(void *p);
int image_alloc(int dev_flat, size_t bytes, talsh_dev_rsc_t *rsc);
void image_release(const talsh_dev_rsc_t &r, Task *defer);
void pending_free(const PendingFree &f);
int tensor_add_image(talsh_tens_t *t, const talsh_dev_rsc_t &r, int kind);
void tensor_remove_image(talsh_tens_t *t, int i, Task *defer);
int find_image(const talsh_tens_t *t, int dev_flat, int kind);
size_t tensor_volume(const talsh_tens_t *t);
void image_wait(hipStream_t s, void *img);
void image_touch(hipStream_t s, void *img);
// One event recorded on s after the current work, as the last use of every
// listed image.
void images_touch(hipStream_t s, void *const *imgs, int n);
// Stream of the last recorded use of an image (nullptr if none).
hipStream_t image_last_stream(void *img);
int prof_collect(int id, double *total_ms, long long *count);

}  // namespace xt
