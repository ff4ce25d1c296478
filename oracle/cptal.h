/* CP-TAL oracle: a plain-C restatement of TAL-SH's CPU tensor-block path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it. The product library
 * (exatensor_amd/libtalsh.so) never links, loads or calls anything here.
 *
 * Every function cites the reference routine it restates (paths relative to the
 * ExaTENSOR source tree). Conventions follow TAL-SH:
 *  - DLF storage: dimension 0 is the fastest-running index.
 *  - Digital contraction pattern dig[0 .. lrank+rrank-1]: value > 0 is an
 *    uncontracted index and names its 1-based position in D; value < 0 is a
 *    contracted index and names the 1-based position in the other operand.
 *  - Permutations are 1-based with element [0] holding the sign (O2N / N2O as
 *    named per function).
 *  - Data kinds: R4=4, R8=8, C4=14, C8=18 (complex stored interleaved re,im).
 */
#ifndef CPTAL_ORACLE_H
#define CPTAL_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPTAL_R4 4
#define CPTAL_R8 8
#define CPTAL_C4 14
#define CPTAL_C8 18
#define CPTAL_MAX_RANK 56

/* Size in bytes of one element of a data kind (0 if invalid). */
int cptal_kind_size(int kind);

/* Symbolic pattern -> digital pattern, TALSH/talshf.F90:619-667 (fake "*R()"
 * for unary patterns, scalar-factor stripping) over
 * TALSH/tensor_algebra_cpu.F90:5307-5531 (get_contr_pattern_dig).
 * Returns 0 on success, the reference's error code otherwise. */
int cptal_str2dig(const char *cptrn, int *dig, int *drank, int *lrank, int *rrank, int *conj_bits);

/* Digital -> symbolic pattern, TALSH/tensor_algebra_cpu.F90:5533-5611. */
int cptal_dig2str(int lrank, int rrank, int conj_bits, const int *dig, char *out, int out_cap);

/* GEMM matricization permutations, TALSH/tensor_algebra_cpu.F90:5613-5755
 * (get_contr_permutations: dprm N2O, lprm/rprm O2N, no hyper-indices). */
void cptal_get_contr_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank, const int *cptrn,
                                  int conj_bits, int *dprm, int *lprm, int *rprm,
                                  int *ncd, int *nlu, int *nru, int *ierr);

/* Hyper-index aware variant, TALSH/tensor_algebra_cpu.F90:5757-5956
 * (get_contraction_permutations). */
void cptal_get_contraction_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank,
                                        const int *cptrn, int conj_bits, int *dprm, int *lprm,
                                        int *rprm, int *ncd, int *nlu, int *nru, int *nhu,
                                        int *ierr);

/* DLF permuted copy out[o2n(idx)] = in[idx] (optionally conjugated),
 * TALSH/tensor_algebra_cpu.F90:3151-3310 (tensor_block_copy) and
 * :7453-8502 (tensor_block_copy_dlf_*). o2n is 1-based with o2n[0] the sign. */
int cptal_copy_dlf(int kind, int rank, const int *dims_in, const int *o2n, const void *in,
                   void *out, int conj);

/* talshTensorCopy on the host: dig from cptal_str2dig, dims of L. */
int cptal_tensor_copy(int kind, int rank, const int *ldims, const int *dig, const void *l, void *d,
                      int conj_bits);

/* Slice extraction, TALSH/tensor_algebra_cpu.F90:2315-2465 (tensor_block_slice)
 * and :6900-6974 (tensor_block_slice_dlf_*): with beta = accumulative ? 1 : 0,
 *   slice[j] = slice[j] * beta + tens[offsets + j]   (a slice of the full
 * extent on every dimension is a copy or an add, as the reference does).
 * tdims: extents of tens; sdims: extents of slice; offsets 0-based.
 * Returns 0, or 17 (the reference's code) when the slice exceeds tens. */
int cptal_slice(int kind, int rank, const int *tdims, const void *tens, const int *sdims, void *slice,
                const int *offsets, int accumulative);

/* Slice insertion, TALSH/tensor_algebra_cpu.F90:2466-2620 (tensor_block_insert)
 * and :7216-7290 (tensor_block_insert_dlf_*):
 *   tens[offsets + j] = tens[offsets + j] * beta + slice[j]. */
int cptal_insert(int kind, int rank, const int *tdims, void *tens, const int *sdims, const void *slice,
                 const int *offsets, int accumulative);

/* Fill, TALSH/tensor_algebra_cpu.F90:955 (tensor_block_init). */
int cptal_init(int kind, size_t vol, void *d, double re, double im);

/* Scale, TALSH/tensor_algebra_cpu.F90:1821 (tensor_block_scale). */
int cptal_scale(int kind, size_t vol, void *d, double re, double im);

/* D += scale * perm(L), TALSH/talshf.F90:1857-1903 (cpu_tensor_block_add)
 * over TALSH/tensor_algebra_cpu.F90:3312-3674 (tensor_block_add). */
int cptal_tensor_add(int kind, int rank, const int *ldims, const int *dig, const void *l, void *d,
                     double sre, double sim, int conj_bits);

/* Tensor contraction D (+)= alpha * L * R with the reference's case logic,
 * TALSH/tensor_algebra_cpu.F90:3675-4358 (tensor_block_contract), in-house GEMM
 * semantics of :8984-10183 (tensor_block_pcontract_dlf_*) and the full
 * contraction of :8795-8983 (tensor_block_fcontract_dlf_*). Rank-0 operands
 * are one-element arrays. Returns 0 or the reference's error code. */
int cptal_contract(const int *dig, int kind, int drank, const int *ddims, void *d, int lrank,
                   const int *ldims, const void *l, int rrank, const int *rdims, const void *r,
                   double are, double aim, int conj_bits, int accumulative);

/* High-precision contraction for error reporting: every element of D is
 * computed as a direct sum in long double (no transposes), same pattern
 * semantics as cptal_contract. Output is written as long double pairs
 * (re, im) into out[2*vol(D)], with D's prior contents scaled by beta. */
int cptal_contract_ld(const int *dig, int kind, int drank, const int *ddims, const void *d,
                      int lrank, const int *ldims, const void *l, int rrank, const int *rdims,
                      const void *r, double are, double aim, int conj_bits, int accumulative,
                      long double *out);

/* Sampled high-precision check: for ns D positions (linear DLF offsets in
 * pos[]), compute the exact-ish value in long double into out[2*ns]. */
int cptal_contract_ld_sample(const int *dig, int kind, int drank, const int *ddims, const void *d,
                             int lrank, const int *ldims, const void *l, int rrank,
                             const int *rdims, const void *r, double are, double aim,
                             int conj_bits, int accumulative, size_t ns, const size_t *pos,
                             long double *out);

/* Seeded inputs used by every test and benchmark: SplitMix64 stream,
 * 53-bit mantissa mapped to U[lo,hi). kind R4/C4 rounds to float. */
void cptal_fill_uniform(int kind, size_t vol, void *d, unsigned long long seed, double lo, double hi);
/* Integer-valued R8 variant in [-2^20, 2^20]. */
void cptal_fill_int(size_t vol, double *d, unsigned long long seed);
/* N(0,1) by Box-Muller on the same stream (real and imaginary parts). */
void cptal_fill_normal(int kind, size_t vol, void *d, unsigned long long seed);

#ifdef __cplusplus
}
#endif

#endif
