/* talsh_amd.h — MI355X engine extensions of the talsh.h C API.
 *
 * None of these entry points has a counterpart in the reference TAL-SH; they
 * expose what this engine adds: kernel timing for roofline reporting, the
 * contraction plan variants used by the autotuner and the learned plan
 * selector, and the batched block-sparse contraction.
 */
#ifndef TALSH_AMD_H_
#define TALSH_AMD_H_

#include "talsh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel classes timed with HIP events on their launch stream. */
#define TALSH_AMD_KERNEL_TRANSPOSE 0
#define TALSH_AMD_KERNEL_GEMM 1
#define TALSH_AMD_KERNEL_GROUPED 2

/* Turn per-launch timing of the dominant kernels on (1) or off (0). */
int talshAmdProfile(int on);
/* Sum of the recorded launch durations (ms) and launch count of a kernel
 * class since the previous query; waits for the recorded launches. */
int talshAmdProfileQuery(int kernel_class, double *total_ms, long long *launches);

/* Plan of subsequent talshTensorContract calls on GEMM-shaped patterns:
 * variant >= 0 selects the variant-th candidate matricization (ranked by the
 * built-in byte-count heuristic; 0 is the heuristic's choice), -1 restores
 * the heuristic; tile selects the GEMM tile shape (-1 default). */
int talshAmdSetPlanVariant(int variant, int tile);
/* Number of distinct matricization candidates of a contraction and the bytes
 * each candidate moves in transposes (moved[] of capacity cap). */
int talshAmdPlanCandidates(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                           const int *rdims, int data_kind, int accumulative, int cap, double *moved, int *ncand);

/* Synchronize all engine streams of a GPU (kind-specific number). */
int talshAmdDeviceSynchronize(int dev_num);

#ifdef __cplusplus
}
#endif

#endif /* TALSH_AMD_H_ */
