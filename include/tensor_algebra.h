/* tensor_algebra.h — lower-level TAL-SH types and constants for the MI355X
 * tensor-block engine.
 *
 * Drop-in for TALSH/tensor_algebra.h: every constant value, struct layout and
 * function signature below matches the reference header (the ABI that
 * existing callers compile against); the contents are written for this
 * engine. Reference locations:
 *   constants        TALSH/tensor_algebra.h:57-203
 *   derived types    TALSH/tensor_algebra.h:211-249
 *   prototypes       TALSH/tensor_algebra.h:259-316
 */
#ifndef TENSOR_ALGEBRA_H_
#define TENSOR_ALGEBRA_H_

#include <stddef.h>
#include <stdint.h>
#include <time.h>

/* ---- global limits (TALSH/tensor_algebra.h:57-60) ---- */
#define MAX_TENSOR_RANK 56
#define MAX_TENSOR_OPERANDS 4
#define MAX_CONTRACTION_PATTERN_LEN 1024
#define MAX_MLNDS_PER_TENS 4

/* ---- data kinds (TALSH/tensor_algebra.h:63-73) ---- */
#define NO_TYPE 0
#define R4 4   /* float */
#define R8 8   /* double */
#define C4 14  /* float complex, interleaved (re,im) */
#define C8 18  /* double complex, interleaved (re,im) */

/* ---- device kinds and flat numbering (TALSH/tensor_algebra.h:75-86) ----
 * Flat ids: 0 host; 1..8 DEV_NVIDIA_GPU; 9..16 DEV_INTEL_MIC; 17..24 DEV_AMD_GPU.
 * This engine serves MI355X GPUs as DEV_AMD_GPU and accepts DEV_NVIDIA_GPU as
 * an alias of the same physical devices for existing callers. */
#define MAX_GPUS_PER_NODE 8
#define MAX_MICS_PER_NODE 8
#define MAX_AMDS_PER_NODE 8
#define DEV_NULL -1
#define DEV_DEFAULT DEV_NULL
#define DEV_HOST 0
#define DEV_NVIDIA_GPU 1
#define DEV_INTEL_MIC 2
#define DEV_AMD_GPU 3
#define DEV_MAX (1 + MAX_GPUS_PER_NODE + MAX_MICS_PER_NODE + MAX_AMDS_PER_NODE)

/* ---- host memory allocation policy ---- */
#define MEM_ALLOC_REGULAR 0
#define MEM_ALLOC_TMP_BUF 1
#define MEM_ALLOC_ALL_BUF 2

/* ---- aliases and status codes (TALSH/tensor_algebra.h:93-117) ---- */
#define NOPE 0
#define YEP 1
#define BLAS_ON 0
#define BLAS_OFF 1
#define EFF_TRN_OFF 0
#define EFF_TRN_ON 1
#define EFF_TRN_ON_CUTT 2
#define DEVICE_UNABLE -546372819  /* not an error: the op can never run on that device */
#define TRY_LATER -918273645      /* not an error: resources temporarily short */
#define NOT_CLEAN -192837465      /* warning: a resource was not released cleanly */
#define NUM_INSTABILITY -374628195
#define EVERYTHING 0
#define SOURCE 1
#define DESTINATION 2
#define TEMPORARY 3
#define DEV_OFF 0
#define DEV_ON 1
#define DEV_ON_BLAS 2
#define GPU_OFF 0
#define GPU_MINE 1
#define GPU_MINE_CUBLAS 2
#define NO_COPY_BACK 0
#define COPY_BACK 1

/* ---- coherence (copy) control (TALSH/tensor_algebra.h:119-203) ----
 * One base-4 digit per tensor argument, the destination being the most
 * significant: D=0 discard, M=1 move, T=2 temporary, K=3 keep. */
#define COPY_D 0
#define COPY_M 1
#define COPY_T 2
#define COPY_K 3
#define COPY_DD 0
#define COPY_DM 1
#define COPY_DT 2
#define COPY_DK 3
#define COPY_MD 4
#define COPY_MM 5
#define COPY_MT 6
#define COPY_MK 7
#define COPY_TD 8
#define COPY_TM 9
#define COPY_TT 10
#define COPY_TK 11
#define COPY_KD 12
#define COPY_KM 13
#define COPY_KT 14
#define COPY_KK 15
#define COPY_DDD 0
#define COPY_DDM 1
#define COPY_DDT 2
#define COPY_DDK 3
#define COPY_DMD 4
#define COPY_DMM 5
#define COPY_DMT 6
#define COPY_DMK 7
#define COPY_DTD 8
#define COPY_DTM 9
#define COPY_DTT 10
#define COPY_DTK 11
#define COPY_DKD 12
#define COPY_DKM 13
#define COPY_DKT 14
#define COPY_DKK 15
#define COPY_MDD 16
#define COPY_MDM 17
#define COPY_MDT 18
#define COPY_MDK 19
#define COPY_MMD 20
#define COPY_MMM 21
#define COPY_MMT 22
#define COPY_MMK 23
#define COPY_MTD 24
#define COPY_MTM 25
#define COPY_MTT 26
#define COPY_MTK 27
#define COPY_MKD 28
#define COPY_MKM 29
#define COPY_MKT 30
#define COPY_MKK 31
#define COPY_TDD 32
#define COPY_TDM 33
#define COPY_TDT 34
#define COPY_TDK 35
#define COPY_TMD 36
#define COPY_TMM 37
#define COPY_TMT 38
#define COPY_TMK 39
#define COPY_TTD 40
#define COPY_TTM 41
#define COPY_TTT 42
#define COPY_TTK 43
#define COPY_TKD 44
#define COPY_TKM 45
#define COPY_TKT 46
#define COPY_TKK 47
#define COPY_KDD 48
#define COPY_KDM 49
#define COPY_KDT 50
#define COPY_KDK 51
#define COPY_KMD 52
#define COPY_KMM 53
#define COPY_KMT 54
#define COPY_KMK 55
#define COPY_KTD 56
#define COPY_KTM 57
#define COPY_KTT 58
#define COPY_KTK 59
#define COPY_KKD 60
#define COPY_KKM 61
#define COPY_KKT 62
#define COPY_KKK 63

#ifndef MIN
#define MIN(a, b) (((a) < (b)) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) (((a) > (b)) ? (a) : (b))
#endif

/* ---- derived types (layouts fixed by TALSH/tensor_algebra.h:211-249) ---- */
typedef struct {
  int num_dim;      /* rank >= 0; -1 empty */
  size_t *offsets;  /* base offsets per dimension */
} talsh_tens_signature_t;

typedef struct {
  int num_dim;  /* rank >= 0; -1 empty */
  int *dims;    /* extents */
  int *divs;    /* dividers */
  int *grps;    /* groups */
} talsh_tens_shape_t;

typedef struct {
  void *base;
  size_t volume;
  int data_kind;
} talsh_tens_data_t;

typedef struct {
  int num_dim;
  int data_kind;
  void *body;
  size_t bases[MAX_TENSOR_RANK];
  size_t dims[MAX_TENSOR_RANK];
} talsh_tens_dense_t;

typedef struct {
  int dev_id;        /* flat device id (>=0), -1 none */
  void *gmem_p;      /* body address on that device */
  int buf_entry;     /* arena handle (>=0), -1 none */
  int mem_attached;  /* 1: external memory, never freed by the engine */
} talsh_dev_rsc_t;

typedef int (*talsh_tens_init_i)(const talsh_tens_data_t *tens_data, const talsh_tens_shape_t *tens_shape,
                                 const talsh_tens_signature_t *tens_signature);

/* Default arguments exist for C++ callers only (the reference header is C++). */
#ifdef __cplusplus
#define TALSH_DEF(x) = x
#else
#define TALSH_DEF(x)
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* No-op initializer usable as a talsh_tens_init_i. */
int talsh_tens_no_init(const talsh_tens_data_t *, const talsh_tens_shape_t *, const talsh_tens_signature_t *);

/* ---- generic helpers (TALSH/tensor_algebra.h:261-278) ---- */
size_t fortran_cptr_int(void *cptr);
int tens_valid_data_kind(int datk, int *datk_size TALSH_DEF(NULL));
int tens_valid_data_kind_(int datk, int *datk_size);
int permutation_trivial(const int perm_len, const int *perm, const int base TALSH_DEF(0));
void get_contr_pattern_sym(const int *rank_left, const int *rank_right, const int *conj_bits,
                           const int *cptrn_dig, char *cptrn_sym, int *cpl, int *ierr);
void get_contr_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank, const int *cptrn, int conj_bits,
                            int *dprm, int *lprm, int *rprm, int *ncd, int *nlu, int *nru, int *ierr);
void get_contraction_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank, const int *cptrn,
                                  int conj_bits, int *dprm, int *lprm, int *rprm, int *ncd, int *nlu, int *nru,
                                  int *nhu, int *ierr);
size_t tens_elem_offset_f(unsigned int num_dim, const unsigned int *dims, const unsigned int *mlndx);
void tens_elem_mlndx_f(size_t offset, unsigned int num_dim, const unsigned int *dims, unsigned int *mlndx);
unsigned int argument_coherence_get_value(unsigned int coh_ctrl, unsigned int tot_args, unsigned int arg_num);
int argument_coherence_set_value(unsigned int *coh_ctrl, unsigned int tot_args, unsigned int arg_num,
                                 unsigned int coh_val);
/* ---- device id codec (TALSH/tensor_algebra.h:279-282) ---- */
int valid_device_kind(int dev_kind);
int encode_device_id(int dev_kind, int dev_num);
int decode_device_id(int dev_id, int *dev_kind TALSH_DEF(NULL));
/* ---- device resources (TALSH/tensor_algebra.h:283-296) ---- */
int tensDevRsc_create(talsh_dev_rsc_t **drsc);
int tensDevRsc_clean(talsh_dev_rsc_t *drsc);
int tensDevRsc_is_empty(talsh_dev_rsc_t *drsc);
int tensDevRsc_same(const talsh_dev_rsc_t *drsc0, const talsh_dev_rsc_t *drsc1);
int tensDevRsc_clone(const talsh_dev_rsc_t *drsc_in, talsh_dev_rsc_t *drsc_out);
int tensDevRsc_attach_mem(talsh_dev_rsc_t *drsc, int dev_id, void *mem_p, int buf_entry TALSH_DEF(-1));
int tensDevRsc_detach_mem(talsh_dev_rsc_t *drsc);
int tensDevRsc_allocate_mem(talsh_dev_rsc_t *drsc, int dev_id, size_t mem_size, int in_arg_buf TALSH_DEF(NOPE));
int tensDevRsc_free_mem(talsh_dev_rsc_t *drsc);
int tensDevRsc_get_gmem_ptr(talsh_dev_rsc_t *drsc, void **gmem_p);
int tensDevRsc_device_id(talsh_dev_rsc_t *drsc);
int tensDevRsc_release_all(talsh_dev_rsc_t *drsc);
int tensDevRsc_destroy(talsh_dev_rsc_t *drsc);
/* ---- signatures and shapes (TALSH/tensor_algebra.h:297-315) ---- */
int tensSignature_create(talsh_tens_signature_t **tsigna);
int tensSignature_clean(talsh_tens_signature_t *tsigna);
int tensSignature_construct(talsh_tens_signature_t *tsigna, int rank, const size_t *offsets TALSH_DEF(NULL));
int tensSignature_destruct(talsh_tens_signature_t *tsigna);
int tensSignature_destroy(talsh_tens_signature_t *tsigna);
int tensShape_create(talsh_tens_shape_t **tshape);
int tensShape_clean(talsh_tens_shape_t *tshape);
int tensShape_construct(talsh_tens_shape_t *tshape, int pinned, int rank, const int *dims TALSH_DEF(NULL),
                        const int *divs TALSH_DEF(NULL), const int *grps TALSH_DEF(NULL));
int tensShape_destruct(talsh_tens_shape_t *tshape);
int tensShape_destroy(talsh_tens_shape_t *tshape);
size_t tensShape_volume(const talsh_tens_shape_t *tshape);
int tensShape_rank(const talsh_tens_shape_t *tshape);
int tensShape_reshape(talsh_tens_shape_t *tshape, int rank, const int *dims TALSH_DEF(NULL),
                      const int *divs TALSH_DEF(NULL), const int *grps TALSH_DEF(NULL));
void tensShape_print(const talsh_tens_shape_t *tshape);

#ifdef __cplusplus
}
#endif

#endif /* TENSOR_ALGEBRA_H_ */
