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
 * built-in byte-count heuristic; 0 is the heuristic's choice), variant
 * TALSH_AMD_FUSED_VARIANT + i the i-th fused-gather candidate (no transposes:
 * the matrix product reads L and R and updates D through index offset
 * tables; native precision, so R4/C4 under fast math keep the heuristic's
 * matricization), -1 restores the heuristic; tile selects the GEMM tile
 * shape (-1 default). */
#define TALSH_AMD_FUSED_VARIANT 1000
int talshAmdSetPlanVariant(int variant, int tile);
/* Number of distinct matricization candidates of a contraction and the bytes
 * each candidate moves in transposes (moved[] of capacity cap). */
int talshAmdPlanCandidates(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                           const int *rdims, int data_kind, int accumulative, int cap, double *moved, int *ncand);
/* Number of fused-gather candidates of a contraction (0 when an index group
 * spans more than 16 dimensions after unit extents are dropped and contiguous
 * runs merged). No counterpart in the reference, which always transposes. */
int talshAmdPlanFusedCandidates(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims,
                                int rrank, const int *rdims, int *nfused);
/* Index groups of fused-gather candidate `candidate` (for inspection and host
 * tests): info[0] 1 if the product computes D^T (A = R, B = L), info[1] and
 * info[2] 1 if A / B are staged along k, info[3 + g] the number of dimensions
 * of group g in the order A rows, A k, B rows, B k, C rows, C columns; ext[16 g
 * + j] and str[16 g + j] the extent and element stride of dimension j of
 * group g (element i of a group sits at sum_j digit_j(i) * str[j], digits of
 * i over ext[], first fastest). No counterpart in the reference. */
int talshAmdPlanFusedMaps(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                          const int *rdims, int candidate, int *info, int *ext, long long *str);

/* Plan of the DLF permutation kernel for a tensor of data_kind with extents
 * dims[rank] and 0-based O2N permutation o2n[rank]: info[0] fused rank,
 * info[1] path (0 copy, 1 LDS-tiled, 2 scatter), info[2] tile volume,
 * info[3] input run, info[4] output run (elements), info[5] workgroups,
 * info[6] elements per global access. */
int talshAmdPermutePlan(int data_kind, int rank, const int *dims, const int *o2n, int *info);

/* Matrix product on GPU 0 with the engine's kernels (device pointers):
 * C(m x n, column-major) = alpha * op(A) * op(B) (+ C unless beta_zero);
 * A is k x m k-fastest (a_kfast) or m x k m-fastest, B is k x n k-fastest
 * (b_kfast) or n x k n-fastest. fast_math selects the three-way BF16 split
 * for R4 and C4 (C4 realified: one real product over 2m x 2k and n x 2k
 * split operands), which needs talshAmdGemmWorkspace() bytes of device workspace; real
 * FP64/FP32 products with few output tiles split K when given that workspace
 * (talshAmdGemmWorkspace covers every tile choice). Synchronous; used by the
 * autotuner and the tests. */
int talshAmdGemm(int data_kind, int fast_math, size_t m, size_t n, size_t k, const void *a, int a_kfast,
                 const void *b, int b_kfast, void *c, double alpha_re, double alpha_im, int beta_zero, int tile,
                 void *workspace, size_t workspace_bytes);
size_t talshAmdGemmWorkspace(int data_kind, int fast_math, size_t m, size_t n, size_t k);
/* talshAmdGemm for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference. The workspace is a fixed 64 MiB device
 * array per GPU (held by each GPU the library runs on until the process
 * ends); a product whose talshAmdGemmWorkspace() need exceeds 64 MiB returns
 * TALSH_LIMIT_EXCEEDED before anything runs. Calls no allocator and records
 * no timing events. Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other host threads) race. Synchronous. */
int talshAmdGemmNoAlloc(int data_kind, int fast_math, size_t m, size_t n, size_t k, const void *a, int a_kfast,
                        const void *b, int b_kfast, void *c, double alpha_re, double alpha_im, int beta_zero,
                        int tile);

/* Batched block-sparse contraction: nops block contractions
 *   D_i (+)= alpha * L_i * R_i   (one symbolic pattern, R8, R4, C8 or C4 with
 *   the pattern's conjugation flags, real alpha; R4/C4 under fast math use
 *   the three-way BF16 split)
 * whose bodies are all resident on GPU dev_num. Blocks sharing a D tensor
 * accumulate in op order. Create builds the work queue once; Execute runs it
 * synchronously as one persistent grouped launch (use_graph: replayed from a
 * captured hipGraph); Destroy releases it. accumulative = NOPE overwrites
 * each D block with the chain's sum. */
int talshAmdBatchCreate(const char *cptrn, int nops, talsh_tens_t *const *dtens, talsh_tens_t *const *ltens,
                        talsh_tens_t *const *rtens, double alpha_re, int accumulative, int dev_num, void **batch);
int talshAmdBatchExecute(void *batch, int use_graph);
int talshAmdBatchInfo(void *batch, int *nops, int *nitems, double *flops);
int talshAmdBatchDestroy(void *batch);

/* ---- learned contraction-plan selector ----
 * A multilayer perceptron (nin inputs, 1-3 hidden ReLU layers of width <= 64,
 * one linear output) that predicts the log of a contraction candidate's
 * kernel time from its features; trained on the GPU with the engine's own
 * forward/backward/Adam kernels. Parameters are FP32, stored layer by layer
 * as W[in][out] (out fastest) followed by b[out]. */
int talshAmdMlpCreate(int nin, int hidden_layers, int width, unsigned long long seed, void **mlp);
int talshAmdMlpDestroy(void *mlp);
int talshAmdMlpNumParams(void *mlp, int *nparams);
int talshAmdMlpGetParams(void *mlp, float *params);
int talshAmdMlpSetParams(void *mlp, const float *params);
/* Inputs are normalized as (x - x_mean) / x_std, outputs de-normalized as
 * y * y_std + y_mean (training sets both from its data). */
int talshAmdMlpSetNormalization(void *mlp, const float *x_mean, const float *x_std, float y_mean, float y_std);
/* Mean squared error (in normalized output units) over n rows and its
 * gradient with respect to every parameter, computed on the GPU. */
int talshAmdMlpGradient(void *mlp, const float *x, const float *y, int n, float *grad, float *loss);
/* GPU forward of n rows of raw features into y[n]. */
int talshAmdMlpPredict(void *mlp, const float *x, long long n, float *y);
/* Adam training on the GPU over n rows (x[n][nin], y[n]); epoch_loss[epochs]
 * receives the mean training loss of every epoch (may be NULL). */
int talshAmdMlpTrain(void *mlp, const float *x, const float *y, long long n, int epochs, int batch, float lr,
                     unsigned long long seed, float *epoch_loss);
int talshAmdMlpSave(void *mlp, const char *path);
int talshAmdMlpLoad(const char *path, void **mlp);
/* Features of every (matricization variant, GEMM tile) candidate of a
 * contraction: features[ncand][talshAmdPlanFeatureCount()], candidate i runs
 * with talshAmdSetPlanVariant(variants[i], tiles[i]). */
int talshAmdPlanFeatureCount(void);
int talshAmdPlanFeatures(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                         const int *rdims, int data_kind, int accumulative, int fast_math, int cap, int *variants,
                         int *tiles, float *features, int *ncand);
/* Install a copy of a trained model as the planner of talshTensorContract
 * (NULL removes it; without a model the byte-count heuristic plans; with one,
 * contractions beyond the model's training domain of 2^35.5 flops,
 * 2^26.5-element operands and the kinds R8, R4, C8 go to the exhaustive
 * autotuner, see talshAmdSetPlanMode). talshInit loads the model file named by the
 * environment variable TALSH_PLAN_MODEL, or else plan_model.txt next to
 * libtalsh.so when present; TALSH_PLAN_MODEL set to an empty string disables
 * the selector. */
int talshAmdSelectorInstall(void *mlp);
/* The candidate talshTensorContract would run: the installed selector's
 * choice, or the heuristic's (variant 0 with its default GEMM tile). */
int talshAmdSelectorChoice(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                           const int *rdims, int data_kind, int accumulative, int fast_math, int *variant, int *tile);

/* How GEMM-shaped contractions choose their plan (matricization and GEMM
 * tile) when no variant is forced with talshAmdSetPlanVariant:
 *   0 heuristic: fewest transpose bytes, default tile;
 *   1 learned (default): the installed selector inside its training domain,
 *     the exhaustive autotuner outside it; with no model installed, the
 *     heuristic;
 *   2 autotune: the exhaustive autotuner for every shape.
 * The autotuner runs every candidate on the caller's L and R (resident on the
 * executing GPU; otherwise the heuristic plans that call) into a scratch D,
 * times its transposes and product with HIP events, and memoizes the fastest
 * per shape (pattern, extents, kind, fast-math mode, accumulation); the first
 * call of a shape blocks for the search. The initial mode comes from the
 * environment variable TALSH_PLAN_MODE (heuristic | learned | autotune). */
int talshAmdSetPlanMode(int mode);
int talshAmdGetPlanMode(void);
/* The autotuner's memoized choice for a contraction shape, or
 * TALSH_NOT_FOUND if it has not tuned that shape; Clear forgets them all. */
int talshAmdAutotuneChoice(const char *cptrn, int drank, const int *ddims, int lrank, const int *ldims, int rrank,
                           const int *rdims, int data_kind, int accumulative, int fast_math, int *variant, int *tile);
int talshAmdAutotuneClear(void);

/* ---- hierarchical block-sparse tensors ----
 * Subspace aggregation tree of a space of dimension dim (HSpaceCtorSimple,
 * INTRAVIRT/subspaces.F90:2098-2372): ids 0..dim-1 are the single basis
 * vectors, dim the full space (the root); level l cuts every aggregate into
 * min(length, branch_fac[l]) contiguous pieces (2 beyond nlevels), the first
 * length % m pieces one longer; pieces longer than one get the next id. */
int talshAmdHSpaceCreate(int dim, int nlevels, const int *branch_fac, void **hspace);
int talshAmdHSpaceDestroy(void *hspace);
int talshAmdHSpaceInfo(void *hspace, int *num_subspaces, int *root);
/* Range [offset, offset + extent) of subspace sid, its tree level, parent
 * (-1 for the root) and number of children; any output may be NULL. */
int talshAmdHSpaceSubspace(void *hspace, int sid, int *offset, int *extent, int *level, int *parent, int *nchildren);
int talshAmdHSpaceChildren(void *hspace, int sid, int cap, int *children);
/* The contraction split of recursive tensors (TensContractionSplitFunc,
 * INTRAVIRT/tensor_recursive.F90:9202-9784; its caller is
 * INTERVIRT/tavp_worker.F90:9467): D, L and R are given by the space of every
 * dimension (xspaces[rank]) and their subtensors (xsub[n][rank], the subspace
 * id of each dimension; every subtensor of an argument at one tree level per
 * dimension). Writes the block contractions as triples (d, l, r) of
 * subtensor indices into triples[cap][3] and their number into *ntriples;
 * TALSH_LIMIT_EXCEEDED (with *ntriples set) when cap is too small. Order: D
 * subtensors by ascending ids (dimension 0 most significant), then contracted
 * ids, R subtensor, L subtensor. Conjugation flags are rejected. */
int talshAmdContractionSplit(const char *cptrn, int drank, void *const *dspaces, int nd, const int *dsub, int lrank,
                             void *const *lspaces, int nl, const int *lsub, int rrank, void *const *rspaces, int nr,
                             const int *rsub, long long cap, int *triples, long long *ntriples);
/* talshAmdBatchCreate over the triples of a split: block contraction i is
 * dblocks[d_i] (+)= alpha * lblocks[l_i] * rblocks[r_i]. */
int talshAmdBatchCreateFromSplit(const char *cptrn, long long ntriples, const int *triples, talsh_tens_t *const *dblocks,
                                 int nd, talsh_tens_t *const *lblocks, int nl, talsh_tens_t *const *rblocks, int nr,
                                 double alpha_re, int accumulative, int dev_num, void **batch);

/* Device bytes one piece of talshTensorContractXL may occupy (slices plus
 * temporaries); contractions above it are decomposed into pieces. 0 restores
 * the default (half of the device arena). */
int talshAmdSetXLBudget(size_t bytes);

/* Synchronize all engine streams of a GPU (kind-specific number). */
int talshAmdDeviceSynchronize(int dev_num);

#ifdef __cplusplus
}
#endif

#endif /* TALSH_AMD_H_ */
