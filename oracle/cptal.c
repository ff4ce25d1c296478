/* CP-TAL oracle: plain-C restatement of TAL-SH's CPU tensor-block path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). Loaded by tests/, the smoke check and
 * the CPU-baseline leg of bench.py; never by the product library.
 *
 * Reference routines restated here (paths relative to the ExaTENSOR tree):
 *   TALSH/talshf.F90:619-667            talsh_get_contr_ptrn_str2dig
 *   TALSH/tensor_algebra_cpu.F90:5307-5531  get_contr_pattern_dig
 *   TALSH/tensor_algebra_cpu.F90:5533-5611  get_contr_pattern_sym
 *   TALSH/tensor_algebra_cpu.F90:5613-5755  get_contr_permutations
 *   TALSH/tensor_algebra_cpu.F90:5757-5956  get_contraction_permutations
 *   TALSH/combinatoric.F90:1316-1368        merge_sort_key_int (stable key sort)
 *   TALSH/tensor_algebra_cpu.F90:3151-3310  tensor_block_copy (+ copy_dlf 7453-8502)
 *   TALSH/tensor_algebra_cpu.F90:3312-3674  tensor_block_add
 *   TALSH/tensor_algebra_cpu.F90:3675-4358  tensor_block_contract
 *   TALSH/tensor_algebra_cpu.F90:8795-8983  tensor_block_fcontract_dlf_*
 *   TALSH/tensor_algebra_cpu.F90:8984-10183 tensor_block_pcontract_dlf_*
 *   TALSH/tensor_algebra_cpu.F90:955, 1821  tensor_block_init / tensor_block_scale
 */
#include "cptal.h"

#include <complex.h>
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXR CPTAL_MAX_RANK

int cptal_kind_size(int kind)
{
  switch (kind) {
  case CPTAL_R4: return 4;
  case CPTAL_R8: return 8;
  case CPTAL_C4: return 8;
  case CPTAL_C8: return 16;
  default: return 0;
  }
}

static int is_complex(int kind) { return kind == CPTAL_C4 || kind == CPTAL_C8; }

/* Real approximant of a complex factor used for real data kinds: modulus with
 * the sign of the real part, or the imaginary part when the real part is zero
 * (cmplx8_to_real8, TALSH/tensor_algebra_cpu.F90:430-445). */
static double c2r(double re, double im)
{
  if (re != 0.0) return hypot(re, im) * (re > 0 ? 1.0 : -1.0);
  return im;
}

/* ------------------------------------------------------------------------- */
/* Pattern parser                                                            */
/* ------------------------------------------------------------------------- */

/* One index occurrence in the pattern, in order of appearance. */
typedef struct {
  int arg;        /* 0:D 1:L 2:R */
  int pos;        /* 1-based position inside the argument */
  int off, len;   /* label span in the string (without a trailing '+') */
  int used;       /* erased (already paired) */
} occ_t;

/* Label check of get_contr_pattern_dig's index_label_ok: alphanumeric, an
 * optional trailing '+' (contravariant) that may not be the first character. */
static int label_ok(const char *s, int n)
{
  if (n <= 0) return 0;
  for (int i = 0; i < n; ++i) {
    unsigned char c = (unsigned char)s[i];
    if (!(isalnum(c) && c < 128)) {
      if (i < n - 1 || i == 0 || c != '+') return 0;
    }
  }
  return 1;
}

static int alnum_us(char c) { return (isalnum((unsigned char)c) && (unsigned char)c < 128) || c == '_'; }

/* get_contr_pattern_dig (TALSH/tensor_algebra_cpu.F90:5307-5531). */
static int pattern_dig(const char *s, int l, int *drank, int *lrank, int *rrank, int *dig, int *conj_bits)
{
  static occ_t occ[3 * MAXR + 8];
  int nocc = 0, adims[3] = {0, 0, 0}, n = -1, conj = 0, i = 0;
  *drank = -1; *lrank = -1; *rrank = -1;
  if (l <= 0) { *conj_bits = 0; return 16; }
  while (i < l) {
    while (s[i] != '(') { ++i; if (i >= l) goto parsed; }
    if (i >= 2 && s[i - 1] == '+' && alnum_us(s[i - 2])) conj += 1 << (n + 1);
    ++n;
    if (n > 2) return 1;
    int start = i + 1, k = 1, after_sep = 0, closed = 0;
    ++i;
    while (i < l) {
      if (s[i] == ',' || s[i] == '|') {
        int len = i - start;
        if (!label_ok(s + start, len)) return 2;
        if (nocc >= 3 * MAXR + 8) return 2;
        int cv = (s[start + len - 1] == '+');
        occ[nocc].arg = n; occ[nocc].pos = k; occ[nocc].off = start; occ[nocc].len = len - cv;
        occ[nocc].used = 0; ++nocc;
        ++k; start = i + 1; after_sep = 1;
      } else if (s[i] == ')') {
        int len = i - start;
        if (len <= 0) {
          if (after_sep) return 3;
        } else {
          if (!label_ok(s + start, len)) return 4;
          int cv = (s[start + len - 1] == '+');
          occ[nocc].arg = n; occ[nocc].pos = k; occ[nocc].off = start; occ[nocc].len = len - cv;
          occ[nocc].used = 0; ++nocc;
          ++k;
        }
        closed = 1; ++i;
        break;
      }
      ++i;
    }
    if (!closed) return 5;
    adims[n] = k - 1;
  }
parsed:
  if (n == 2) {
    *drank = adims[0]; *lrank = adims[1]; *rrank = adims[2];
  } else if (n == 1) {
    *drank = adims[0]; *lrank = adims[1];
  } else {
    *conj_bits = conj;
    return 15;
  }
  const int L = adims[1];
  for (int a = 0; a < nocc; ++a) {
    if (occ[a].used) continue;
    /* the reference cannot open a pairing on the very last label (:5390-5391) */
    if (a == nocc - 1) return n == 2 ? 7 : 12;
    int k1 = occ[a].arg, k2 = occ[a].pos;
    int rounds = (n == 2) ? 2 : 1;
    int from = a + 1;
    for (int kf = 1; kf <= rounds; ++kf) {
      int b = -1;
      for (int c = from; c < nocc; ++c) {
        if (occ[c].used) continue;
        if (occ[c].len == occ[a].len && memcmp(s + occ[c].off, s + occ[a].off, (size_t)occ[a].len) == 0) { b = c; break; }
      }
      if (b < 0) {
        if (kf == 1) return n == 2 ? 11 : 14;
        break;
      }
      int k3 = occ[b].arg, k4 = occ[b].pos;
      if (n == 2) {
        if (k1 == 0 && k3 == 1) {
          if (kf != 1) return 8;
          dig[k4 - 1] = k2;
        } else if (k1 == 0 && k3 == 2) {
          dig[L + k4 - 1] = k2;
        } else if (k1 == 1 && k3 == 2) {
          if (kf != 1) return 9;
          dig[k2 - 1] = -k4; dig[L + k4 - 1] = -k2;
        } else {
          return 10;
        }
      } else {
        if (k1 == 0 && k3 == 1) dig[k4 - 1] = k2; else return 13;
      }
      occ[b].used = 1;
    }
    occ[a].used = 1;
  }
  *conj_bits = conj;
  return 0;
}

int cptal_str2dig(const char *c_str, int *dig, int *drank, int *lrank, int *rrank, int *conj_bits)
{
  /* talsh_get_contr_ptrn_str2dig (TALSH/talshf.F90:619-667) */
  char buf[1024 + 8];
  int csl = 0, star = -1;
  *drank = -1; *lrank = -1; *rrank = -1; *conj_bits = 0;
  while (c_str[csl] != 0) {
    if (csl >= 1024) return -1;
    if (c_str[csl] == '*' && star < 0) star = csl;
    buf[csl] = c_str[csl];
    ++csl;
  }
  if (star >= 0) {
    int has_paren = 0;
    for (int i = star; i < csl; ++i) if (buf[i] == '(') { has_paren = 1; break; }
    if (!has_paren) { /* scalar factor: strip "*<number>" */
      csl = star; star = -1;
      if (csl <= 0) return -2;
      if (buf[csl - 1] != ')') return -3;
    }
  }
  if (star < 0) { memcpy(buf + csl, "*R()", 4); csl += 4; }
  /* Fortran len_trim drops trailing blanks */
  while (csl > 0 && buf[csl - 1] == ' ') --csl;
  int tmp[2 * MAXR + 2];
  int ierr = pattern_dig(buf, csl, drank, lrank, rrank, tmp, conj_bits);
  if (ierr != 0) return ierr;
  /* the wrapper copies lrank+rrank entries; a unary parse (rrank=-1) copies one fewer */
  int dgl = *lrank + *rrank;
  for (int i = 0; i < dgl; ++i) dig[i] = tmp[i];
  return 0;
}

int cptal_dig2str(int lrank, int rrank, int conj_bits, const int *dig, char *out, int cap)
{
  /* get_contr_pattern_sym (TALSH/tensor_algebra_cpu.F90:5533-5611) */
  char b[4 * MAXR + 64];
  int p = 0;
  char left_lbl[MAXR + 1];
  if (lrank < 0 || rrank < 0) return 1;
  int nu = 0;
  for (int i = 0; i < lrank + rrank; ++i) if (dig[i] > 0) ++nu;
  p += sprintf(b + p, (conj_bits & 1) ? "D+(" : "D(");
  int m = 'a';
  for (int i = 0; i < nu; ++i) { b[p++] = (char)m++; b[p++] = ','; }
  if (b[p - 1] == '(') ++p;
  b[p - 1] = ')';
  p += sprintf(b + p, (conj_bits & 2) ? "+=L+(" : "+=L(");
  for (int i = 0; i < lrank; ++i) {
    if (dig[i] > 0) b[p++] = (char)('a' + dig[i] - 1);
    else { left_lbl[i] = (char)m; b[p++] = (char)m++; }
    b[p++] = ',';
  }
  if (b[p - 1] == '(') ++p;
  b[p - 1] = ')';
  p += sprintf(b + p, (conj_bits & 4) ? "*R+(" : "*R(");
  for (int i = 0; i < rrank; ++i) {
    int v = dig[lrank + i];
    b[p++] = v > 0 ? (char)('a' + v - 1) : left_lbl[-v - 1];
    b[p++] = ',';
  }
  if (b[p - 1] == '(') ++p;
  b[p - 1] = ')';
  b[p] = 0;
  if (p + 1 > cap) return 2;
  memcpy(out, b, (size_t)p + 1);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Permutations                                                              */
/* ------------------------------------------------------------------------- */

/* merge_sort_key_int (TALSH/combinatoric.F90:1316-1368): stable sort of the
 * items trn[1..n] by key[item-1]; trn[0] carries the permutation sign. */
static void sort_by_key(int n, const int *key, int *trn)
{
  for (int i = 2; i <= n; ++i) {       /* stable insertion sort, parity tracked */
    int v = trn[i], j = i - 1;
    while (j >= 1 && key[trn[j] - 1] > key[v - 1]) { trn[j + 1] = trn[j]; --j; trn[0] = -trn[0]; }
    trn[j + 1] = v;
  }
}

static int pattern_ok_simple(int lrank, int rrank, const int *c)
{
  int jl = lrank + rrank, jc = 0, seen[2 * MAXR + 2];
  if (jl <= 0) return 1;
  memset(seen, 0, sizeof(seen));
  for (int j0 = 1; j0 <= jl; ++j0) {
    int j1 = c[j0 - 1];
    if (j1 < 0) {
      if (j0 <= lrank) {
        if (-j1 > rrank || c[lrank - j1 - 1] != -j0) return 0;
      } else {
        if (-j1 > lrank || c[-j1 - 1] != -(j0 - lrank)) return 0;
      }
    } else if (j1 > 0 && j1 <= jl) {
      ++jc;
      if (seen[j1]) return 0;
      seen[j1] = 1;
    } else {
      return 0;
    }
  }
  for (int j0 = 1; j0 <= jc; ++j0) if (seen[j0] != 1) return 0;
  return 1;
}

void cptal_get_contr_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank, const int *cptrn,
                                  int conj_bits, int *dprm, int *lprm, int *rprm,
                                  int *ncd, int *nlu, int *nru, int *ierr)
{
  *ierr = 0;
  if (!(pattern_ok_simple(lrank, rrank, cptrn) && lrank >= 0 && rrank >= 0)) { *ierr = 1; return; }
  int left_conj = (conj_bits >> 1) & 1, right_conj = (conj_bits >> 2) & 1;
  if (conj_bits & 1) { left_conj = !left_conj; right_conj = !right_conj; }
  if (gemm_tl == 0 && left_conj) { *ierr = 2; return; }
  int drank = 0, simple;
  dprm[0] = 1;
  for (int i = 1; i <= lrank; ++i) if (cptrn[i - 1] > 0) dprm[++drank] = cptrn[i - 1];
  *nlu = drank; simple = (drank >= 2);
  for (int i = lrank + 1; i <= lrank + rrank; ++i) {
    if (cptrn[i - 1] > 0) {
      dprm[++drank] = cptrn[i - 1];
      if (cptrn[i - 1] <= *nlu) simple = 0;
    }
  }
  if (simple) for (int j = 1; j <= drank; ++j) dprm[j] = j;
  rprm[0] = 1; *nru = 0; *ncd = 0;
  int j = 0;
  if (rrank > 0) {
    for (int i = 1; i <= rrank; ++i) if (cptrn[lrank + i - 1] < 0) rprm[i] = ++j;
    *ncd = j; *nru = rrank - *ncd;
    if (simple && *nru >= 2) {
      for (int i = 1; i <= rrank; ++i) if (cptrn[lrank + i - 1] > 0) rprm[i] = *ncd + (cptrn[lrank + i - 1] - *nlu);
    } else {
      for (int i = 1; i <= rrank; ++i) if (cptrn[lrank + i - 1] > 0) rprm[i] = ++j;
    }
  }
  lprm[0] = 1;
  if (lrank > 0) {
    int jkey[2 * MAXR + 2], jtrn0[2 * MAXR + 2], jtrn1[2 * MAXR + 2];
    j = 0;
    for (int i = 1; i <= lrank; ++i) if (cptrn[i - 1] < 0) { ++j; jtrn1[j] = i; jkey[j - 1] = -cptrn[i - 1]; }
    jtrn0[0] = 1;
    for (int k = 1; k <= j; ++k) jtrn0[k] = k;
    if (j >= 2) sort_by_key(j, jkey, jtrn0);
    for (int i = 1; i <= j; ++i) lprm[jtrn1[jtrn0[i]]] = i;
    if (simple && *nlu >= 2) {
      for (int i = 1; i <= lrank; ++i) if (cptrn[i - 1] > 0) lprm[i] = *ncd + cptrn[i - 1];
    } else {
      for (int i = 1; i <= lrank; ++i) if (cptrn[i - 1] > 0) lprm[i] = ++j;
    }
  }
  if (gemm_tl == 0) {
    for (int i = 1; i <= lrank; ++i) lprm[i] = (lprm[i] <= *ncd) ? lprm[i] + *nlu : lprm[i] - *ncd;
  }
  if (gemm_tr != 0 || right_conj) {
    for (int i = 1; i <= rrank; ++i) rprm[i] = (rprm[i] <= *ncd) ? rprm[i] + *nru : rprm[i] - *ncd;
  }
}

static int pattern_ok_hyper(int lrank, int rrank, const int *c)
{
  int jl = lrank + rrank, jc = 0, jd[2 * MAXR + 2];
  if (jl <= 0) return 1;
  memset(jd, 0, sizeof(jd));
  for (int j0 = 1; j0 <= jl; ++j0) {
    int j1 = c[j0 - 1];
    if (j1 < 0) {
      if (j0 <= lrank) {
        if (-j1 > rrank || c[lrank - j1 - 1] != -j0) return 0;
      } else {
        if (-j1 > lrank || c[-j1 - 1] != -(j0 - lrank)) return 0;
      }
    } else if (j1 > 0 && j1 <= jl) {
      if (j0 <= lrank) {
        if (jd[j1] == 0) { jd[j1] = 1; ++jc; } else return 0;
      } else {
        if (jd[j1] == 0) { jd[j1] = 1; ++jc; }
        else if (jd[j1] == 1) jd[j1] = 2;
        else return 0;
      }
    } else {
      return 0;
    }
  }
  for (int j0 = 1; j0 <= jc; ++j0) if (jd[j0] != 1 && jd[j0] != 2) return 0;
  return 1;
}

void cptal_get_contraction_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank,
                                        const int *cptrn, int conj_bits, int *dprm, int *lprm,
                                        int *rprm, int *ncd, int *nlu, int *nru, int *nhu,
                                        int *ierr)
{
  int n = lrank + rrank, drank = 0;
  int dtens[2 * MAXR + 2], ltens[MAXR + 1], rtens[MAXR + 1];
  *ierr = 0; *ncd = 0; *nlu = 0; *nru = 0; *nhu = 0;
  dprm[0] = 1; lprm[0] = 1; rprm[0] = 1;
  if (!pattern_ok_hyper(lrank, rrank, cptrn)) { *ierr = 3; return; }
  if (lrank < 0 || rrank < 0) { *ierr = 1; return; }
  int left_conj = (conj_bits >> 1) & 1, right_conj = (conj_bits >> 2) & 1;
  if (conj_bits & 1) { left_conj = !left_conj; right_conj = !right_conj; }
  if (gemm_tl == 0 && left_conj) { *ierr = 2; return; }
  if (n > 0) {
    for (int i = 1; i <= n; ++i) dtens[i] = 0;
    for (int i = 1; i <= n; ++i) {
      int j = cptrn[i - 1];
      if (j > 0) {
        if (j > drank) drank = j;
        dtens[j] += 1;
        if (i <= lrank) ++*nlu; else ++*nru;
      } else if (j < 0) {
        ++*ncd;
      }
    }
    *ncd /= 2;
    int j = 0;
    for (int i = 1; i <= drank; ++i) {
      if (dtens[i] == 1) dtens[i] = ++j;
      else if (dtens[i] == 2) { ++*nhu; dtens[i] = drank + *nhu; }
    }
    *nlu -= *nhu; *nru -= *nhu;
    j = 0;
    for (int i = 1; i <= lrank; ++i) {
      if (cptrn[i - 1] < 0) { ltens[i] = -(*ncd - j); ++j; }
      else ltens[i] = dtens[cptrn[i - 1]];
    }
    for (int i = 1; i <= rrank; ++i) {
      int v = cptrn[lrank + i - 1];
      rtens[i] = (v < 0) ? ltens[-v] : dtens[v];
    }
    int k = 0;
    if (lrank > 0) {
      dprm[0] = 1;
      for (int q = 1; q <= lrank; ++q) dprm[q] = q;
      if (lrank > 1) sort_by_key(lrank, ltens + 1, dprm);
      for (int i = 1; i <= lrank; ++i) {
        int jj = dprm[i];
        if (ltens[jj] >= 1 && ltens[jj] <= drank) dtens[++k] = ltens[jj];
      }
      lprm[0] = dprm[0];
      for (int i = 1; i <= lrank; ++i) lprm[dprm[i]] = i;
    }
    if (rrank > 0) {
      dprm[0] = 1;
      for (int q = 1; q <= rrank; ++q) dprm[q] = q;
      if (rrank > 1) sort_by_key(rrank, rtens + 1, dprm);
      for (int i = 1; i <= rrank; ++i) {
        int jj = dprm[i];
        if (rtens[jj] >= 1 && rtens[jj] <= drank) { ++k; dtens[k] = rtens[jj]; }
        else if (rtens[jj] > drank) { ++k; dtens[k] = k; }
      }
      rprm[0] = dprm[0];
      for (int i = 1; i <= rrank; ++i) rprm[dprm[i]] = i;
    }
    if (drank > 0) {
      dprm[0] = 1;
      for (int i = 1; i <= drank; ++i) dprm[i] = dtens[i];
    }
  }
  if (gemm_tl == 0) {
    for (int i = 1; i <= lrank; ++i) {
      if (lprm[i] <= *ncd) lprm[i] += *nlu;
      else if (lprm[i] <= *ncd + *nlu) lprm[i] -= *ncd;
    }
  }
  if (gemm_tr != 0 || right_conj) {
    for (int i = 1; i <= rrank; ++i) {
      if (rprm[i] <= *ncd) rprm[i] += *nru;
      else if (rprm[i] <= *ncd + *nru) rprm[i] -= *ncd;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Element-wise helpers                                                      */
/* ------------------------------------------------------------------------- */

static size_t volume(int rank, const int *dims)
{
  size_t v = 1;
  for (int i = 0; i < rank; ++i) v *= (size_t)dims[i];
  return v;
}

static int perm_valid(int n, const int *o2n)
{
  int seen[MAXR + 1] = {0};
  for (int i = 1; i <= n; ++i) {
    if (o2n[i] < 1 || o2n[i] > n || seen[o2n[i]]) return 0;
    seen[o2n[i]] = 1;
  }
  return 1;
}

static int perm_trivial(int n, const int *p)
{
  for (int i = 1; i <= n; ++i) if (p[i] != i) return 0;
  return 1;
}

int cptal_copy_dlf(int kind, int rank, const int *dims_in, const int *o2n, const void *in, void *out, int conj)
{
  int es = cptal_kind_size(kind);
  if (es == 0) return 1;
  if (rank < 0 || rank > MAXR) return 2;
  if (rank > 0 && !perm_valid(rank, o2n)) return 3;
  size_t vol = volume(rank, dims_in);
  int doconj = conj && is_complex(kind);
  /* output strides: output dim o2n[i] has extent dims_in[i] */
  size_t ostr_of_out[MAXR + 1], ostr[MAXR];
  int odims[MAXR];
  for (int i = 0; i < rank; ++i) odims[o2n[i + 1] - 1] = dims_in[i];
  size_t s = 1;
  for (int q = 0; q < rank; ++q) { ostr_of_out[q] = s; s *= (size_t)odims[q]; }
  for (int i = 0; i < rank; ++i) ostr[i] = ostr_of_out[o2n[i + 1] - 1];
  if (vol == 0) return 0;
  /* Parallel over the input in chunks of the leading dimension. */
  size_t d0 = rank > 0 ? (size_t)dims_in[0] : 1;
  size_t nrows = vol / d0;
  const char *src = (const char *)in;
  char *dst = (char *)out;
#pragma omp parallel for schedule(static)
  for (size_t row = 0; row < nrows; ++row) {
    size_t rem = row, obase = 0;
    for (int i = 1; i < rank; ++i) {
      size_t c = rem % (size_t)dims_in[i];
      rem /= (size_t)dims_in[i];
      obase += c * ostr[i];
    }
    size_t ibase = row * d0, st = rank > 0 ? ostr[0] : 1;
    for (size_t c0 = 0; c0 < d0; ++c0) {
      const char *a = src + (ibase + c0) * (size_t)es;
      char *b = dst + (obase + c0 * st) * (size_t)es;
      memcpy(b, a, (size_t)es);
      if (doconj) {
        if (kind == CPTAL_C4) { float *f = (float *)b; f[1] = -f[1]; }
        else { double *f = (double *)b; f[1] = -f[1]; }
      }
    }
  }
  return 0;
}

int cptal_tensor_copy(int kind, int rank, const int *ldims, const int *dig, const void *l, void *d, int conj_bits)
{
  /* cpu_tensor_block_copy (TALSH/talshf.F90:1828-1855): transp = dig (O2N);
   * conjugation bits follow tensor_block_copy (bit0 D flips bit1 L). */
  int trn[MAXR + 1];
  trn[0] = 1;
  for (int i = 0; i < rank; ++i) trn[i + 1] = dig[i];
  int lconj = (conj_bits >> 1) & 1;
  if (conj_bits & 1) lconj = !lconj;
  if (rank == 0) {
    int es = cptal_kind_size(kind);
    memcpy(d, l, (size_t)es);
    if (lconj && is_complex(kind)) {
      if (kind == CPTAL_C4) ((float *)d)[1] = -((float *)d)[1];
      else ((double *)d)[1] = -((double *)d)[1];
    }
    return 0;
  }
  return cptal_copy_dlf(kind, rank, ldims, trn, l, d, lconj);
}

/* Element update of slice/insert: out = out * beta + in (alpha = 1), in the
 * kind's own arithmetic. Complex factors are (beta, 0) and (1, 0), so the
 * complex products reduce to the componentwise forms below. */
static void box_update(int kind, void *out, size_t o, const void *in, size_t i, int beta, int full)
{
  switch (kind) {
  case CPTAL_R4: {
    float *y = (float *)out; const float *x = (const float *)in;
    y[o] = full ? (beta ? y[o] + x[i] : x[i]) : y[o] * (float)beta + x[i] * 1.0f;
    break;
  }
  case CPTAL_R8: {
    double *y = (double *)out; const double *x = (const double *)in;
    y[o] = full ? (beta ? y[o] + x[i] : x[i]) : y[o] * (double)beta + x[i] * 1.0;
    break;
  }
  case CPTAL_C4: {
    float *y = (float *)out; const float *x = (const float *)in;
    for (int c = 0; c < 2; ++c)
      y[2 * o + c] = full ? (beta ? y[2 * o + c] + x[2 * i + c] : x[2 * i + c]) : y[2 * o + c] * (float)beta + x[2 * i + c];
    break;
  }
  case CPTAL_C8: {
    double *y = (double *)out; const double *x = (const double *)in;
    for (int c = 0; c < 2; ++c)
      y[2 * o + c] = full ? (beta ? y[2 * o + c] + x[2 * i + c] : x[2 * i + c]) : y[2 * o + c] * (double)beta + x[2 * i + c];
    break;
  }
  }
}

/* Walks the slice box in DLF order; insert = 0: slice <- tens, 1: tens <- slice. */
static int box_walk(int kind, int rank, const int *tdims, void *tens, const int *sdims, void *slice, const int *offsets,
                    int accumulative, int insert)
{
  if (!cptal_kind_size(kind) || rank <= 0 || rank > CPTAL_MAX_RANK) return 1;
  int full = 1;
  for (int i = 0; i < rank; ++i) {
    if (offsets[i] < 0 || offsets[i] >= tdims[i] || sdims[i] <= 0 || offsets[i] + sdims[i] - 1 >= tdims[i]) return 17;
    if (sdims[i] != tdims[i]) full = 0;
  }
  size_t tbase[CPTAL_MAX_RANK], svol = 1, st = 1;
  for (int i = 0; i < rank; ++i) { tbase[i] = st; st *= (size_t)tdims[i]; svol *= (size_t)sdims[i]; }
  int idx[CPTAL_MAX_RANK] = {0};
  for (size_t j = 0; j < svol; ++j) {
    size_t t = 0;
    for (int i = 0; i < rank; ++i) t += (size_t)(offsets[i] + idx[i]) * tbase[i];
    if (insert) box_update(kind, tens, t, slice, j, accumulative != 0, full);
    else box_update(kind, slice, j, tens, t, accumulative != 0, full);
    for (int i = 0; i < rank; ++i) {
      if (++idx[i] < sdims[i]) break;
      idx[i] = 0;
    }
  }
  return 0;
}

int cptal_slice(int kind, int rank, const int *tdims, const void *tens, const int *sdims, void *slice,
                const int *offsets, int accumulative)
{
  return box_walk(kind, rank, tdims, (void *)tens, sdims, slice, offsets, accumulative, 0);
}

int cptal_insert(int kind, int rank, const int *tdims, void *tens, const int *sdims, const void *slice,
                 const int *offsets, int accumulative)
{
  return box_walk(kind, rank, tdims, tens, sdims, (void *)slice, offsets, accumulative, 1);
}

int cptal_init(int kind, size_t vol, void *d, double re, double im)
{
  switch (kind) {
  case CPTAL_R4: { float v = (float)c2r(re, im); float *p = d;
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) p[i] = v; break; }
  case CPTAL_R8: { double *p = d; double v = c2r(re, im);
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) p[i] = v; break; }
  case CPTAL_C4: { float *p = d; float a = (float)re, b = (float)im;
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) { p[2 * i] = a; p[2 * i + 1] = b; } break; }
  case CPTAL_C8: { double *p = d;
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) { p[2 * i] = re; p[2 * i + 1] = im; } break; }
  default: return 1;
  }
  return 0;
}

int cptal_scale(int kind, size_t vol, void *d, double re, double im)
{
  switch (kind) {
  case CPTAL_R4: { float v = (float)c2r(re, im); float *p = d;
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) p[i] = p[i] * v; break; }
  case CPTAL_R8: { double *p = d; double v = c2r(re, im);
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) p[i] = p[i] * v; break; }
  case CPTAL_C4: { float *p = d; float a = (float)re, b = (float)im;
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) {
      float x = p[2 * i], y = p[2 * i + 1];
      p[2 * i] = x * a - y * b; p[2 * i + 1] = x * b + y * a;
    } break; }
  case CPTAL_C8: { double *p = d;
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) {
      double x = p[2 * i], y = p[2 * i + 1];
      p[2 * i] = x * re - y * im; p[2 * i + 1] = x * im + y * re;
    } break; }
  default: return 1;
  }
  return 0;
}

/* d (+)= s * (conj?) l, element-wise (tensor_block_add's dense loops). */
static void axpy(int kind, size_t vol, void *d, const void *l, double sre, double sim, int lconj, int accum)
{
  switch (kind) {
  case CPTAL_R4: { float *p = d; const float *q = l; float s = (float)c2r(sre, sim);
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) { float t = q[i] * s; p[i] = accum ? p[i] + t : t; } break; }
  case CPTAL_R8: { double *p = d; const double *q = l; double s = c2r(sre, sim);
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) { double t = q[i] * s; p[i] = accum ? p[i] + t : t; } break; }
  case CPTAL_C4: { float *p = d; const float *q = l; float a = (float)sre, b = (float)sim;
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) {
      float x = q[2 * i], y = lconj ? -q[2 * i + 1] : q[2 * i + 1];
      float tr = x * a - y * b, ti = x * b + y * a;
      if (accum) { p[2 * i] += tr; p[2 * i + 1] += ti; } else { p[2 * i] = tr; p[2 * i + 1] = ti; }
    } break; }
  case CPTAL_C8: { double *p = d; const double *q = l;
#pragma omp parallel for
    for (size_t i = 0; i < vol; ++i) {
      double x = q[2 * i], y = lconj ? -q[2 * i + 1] : q[2 * i + 1];
      double tr = x * sre - y * sim, ti = x * sim + y * sre;
      if (accum) { p[2 * i] += tr; p[2 * i + 1] += ti; } else { p[2 * i] = tr; p[2 * i + 1] = ti; }
    } break; }
  }
}

int cptal_tensor_add(int kind, int rank, const int *ldims, const int *dig, const void *l, void *d,
                     double sre, double sim, int conj_bits)
{
  int es = cptal_kind_size(kind);
  if (es == 0) return 1;
  size_t vol = volume(rank, ldims);
  int lconj = (conj_bits >> 1) & 1;
  if (conj_bits & 1) lconj = !lconj;
  if (!is_complex(kind)) lconj = 0;
  int permute = 0;
  for (int i = 0; i < rank; ++i) if (dig[i] != i + 1) { permute = 1; break; }
  if (permute) {
    void *tmp = malloc(vol * (size_t)es);
    if (!tmp) return 2;
    int trn[MAXR + 1];
    trn[0] = 1;
    for (int i = 0; i < rank; ++i) trn[i + 1] = dig[i];
    int e = cptal_copy_dlf(kind, rank, ldims, trn, l, tmp, 0);
    if (e == 0) axpy(kind, vol, d, tmp, sre, sim, lconj, 1);
    free(tmp);
    return e;
  }
  axpy(kind, vol, d, l, sre, sim, lconj, 1);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Contraction                                                               */
/* ------------------------------------------------------------------------- */

/* D(ll,lr) = beta*D + alpha * sum_c L(c,ll) * R(c,lr), column-major, the
 * in-house GEMM of tensor_block_pcontract_dlf_* (8 partial sums per dot
 * product over the contracted index, as in :9351-9388). */
static void pcontract(int kind, size_t dl, size_t dr, size_t dc, const void *L, const void *R, void *D,
                      double are, double aim, int beta_zero)
{
  if (kind == CPTAL_R8 || kind == CPTAL_R4) {
#pragma omp parallel for collapse(2) schedule(static)
    for (size_t r = 0; r < dr; ++r) {
      for (size_t l = 0; l < dl; ++l) {
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (kind == CPTAL_R8) {
          const double *a = (const double *)L + l * dc, *b = (const double *)R + r * dc;
          size_t c = 0;
          for (; c + 8 <= dc; c += 8)
            for (int q = 0; q < 8; ++q) v[q] += a[c + q] * b[c + q] * are;
          for (size_t q = 0; c + q < dc; ++q) v[q] += a[c + q] * b[c + q] * are;
          double *dd = (double *)D + r * dl + l;
          double base = beta_zero ? 0.0 : *dd;
          *dd = base + (v[0] + v[4]) + (v[1] + v[5]) + (v[2] + v[6]) + (v[3] + v[7]);
        } else {
          float w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          const float *a = (const float *)L + l * dc, *b = (const float *)R + r * dc;
          float al = (float)are;
          size_t c = 0;
          for (; c + 8 <= dc; c += 8)
            for (int q = 0; q < 8; ++q) w[q] += a[c + q] * b[c + q] * al;
          for (size_t q = 0; c + q < dc; ++q) w[q] += a[c + q] * b[c + q] * al;
          float *dd = (float *)D + r * dl + l;
          float base = beta_zero ? 0.0f : *dd;
          *dd = base + (w[0] + w[4]) + (w[1] + w[5]) + (w[2] + w[6]) + (w[3] + w[7]);
        }
        (void)v;
      }
    }
  } else {
#pragma omp parallel for collapse(2) schedule(static)
    for (size_t r = 0; r < dr; ++r) {
      for (size_t l = 0; l < dl; ++l) {
        if (kind == CPTAL_C8) {
          const double *a = (const double *)L + 2 * l * dc, *b = (const double *)R + 2 * r * dc;
          double sr = 0, si = 0;
          for (size_t c = 0; c < dc; ++c) {
            double xr = a[2 * c], xi = a[2 * c + 1], yr = b[2 * c], yi = b[2 * c + 1];
            double pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;
            sr += pr * are - pi * aim; si += pr * aim + pi * are;
          }
          double *dd = (double *)D + 2 * (r * dl + l);
          if (beta_zero) { dd[0] = sr; dd[1] = si; } else { dd[0] += sr; dd[1] += si; }
        } else {
          const float *a = (const float *)L + 2 * l * dc, *b = (const float *)R + 2 * r * dc;
          float sr = 0, si = 0, ar = (float)are, ai = (float)aim;
          for (size_t c = 0; c < dc; ++c) {
            float xr = a[2 * c], xi = a[2 * c + 1], yr = b[2 * c], yi = b[2 * c + 1];
            float pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;
            sr += pr * ar - pi * ai; si += pr * ai + pi * ar;
          }
          float *dd = (float *)D + 2 * (r * dl + l);
          if (beta_zero) { dd[0] = sr; dd[1] = si; } else { dd[0] += sr; dd[1] += si; }
        }
      }
    }
  }
}

/* Multiply a scalar element x by y (complex-aware), result into x. */
static void cmul_elem(int kind, double *xr, double *xi, double yr, double yi)
{
  double a = *xr * yr - *xi * yi, b = *xr * yi + *xi * yr;
  if (!is_complex(kind)) { *xr = *xr * yr; *xi = 0; return; }
  *xr = a; *xi = b;
}

static void load_elem(int kind, const void *p, size_t i, double *re, double *im)
{
  switch (kind) {
  case CPTAL_R4: *re = ((const float *)p)[i]; *im = 0; break;
  case CPTAL_R8: *re = ((const double *)p)[i]; *im = 0; break;
  case CPTAL_C4: *re = ((const float *)p)[2 * i]; *im = ((const float *)p)[2 * i + 1]; break;
  default: *re = ((const double *)p)[2 * i]; *im = ((const double *)p)[2 * i + 1]; break;
  }
}

static void store_elem(int kind, void *p, size_t i, double re, double im)
{
  switch (kind) {
  case CPTAL_R4: ((float *)p)[i] = (float)re; break;
  case CPTAL_R8: ((double *)p)[i] = re; break;
  case CPTAL_C4: ((float *)p)[2 * i] = (float)re; ((float *)p)[2 * i + 1] = (float)im; break;
  default: ((double *)p)[2 * i] = re; ((double *)p)[2 * i + 1] = im; break;
  }
}

int cptal_contract(const int *dig, int kind, int drank, const int *ddims, void *d, int lrank,
                   const int *ldims, const void *l, int rrank, const int *rdims, const void *r,
                   double are, double aim, int conj_bits, int accumulative)
{
  int es = cptal_kind_size(kind);
  if (es == 0) return 6;
  if (lrank < 0 || rrank < 0 || drank < 0 || lrank > MAXR || rrank > MAXR || drank > MAXR) return 5;
  int cplx = is_complex(kind);
  int lconj = 0, rconj = 0, conj = cplx ? conj_bits : 0;
  if (cplx) {
    lconj = (conj >> 1) & 1; rconj = (conj >> 2) & 1;
    if (conj & 1) { lconj = !lconj; rconj = !rconj; }
  }
  const double aim_in = aim; /* real kinds: GEMM paths use Re(alpha) only */
  if (!cplx) aim = 0;
  size_t vd = volume(drank, ddims);
  int beta_zero = !accumulative;
  /* Case classification (:3779-3792) */
  int ls = (lrank == 0), rs = (rrank == 0), ds = (drank == 0);
  if (ls && rs && ds) { /* MULTIPLY_SCALARS */
    double xr, xi, yr, yi, zr, zi;
    load_elem(kind, l, 0, &xr, &xi); load_elem(kind, r, 0, &yr, &yi);
    if (lconj) xi = -xi;
    if (rconj) yi = -yi;
    load_elem(kind, d, 0, &zr, &zi);
    if (beta_zero) { zr = 0; zi = 0; }
    cmul_elem(kind, &xr, &xi, yr, yi); cmul_elem(kind, &xr, &xi, are, aim);
    store_elem(kind, d, 0, zr + xr, zi + xi);
    return 0;
  }
  if ((!ls && rs && !ds) || (ls && !rs && !ds)) { /* ADD_TENSOR */
    const void *t = ls ? r : l;
    const void *s = ls ? l : r;
    int trank = ls ? rrank : lrank;
    const int *tdims = ls ? rdims : ldims;
    const int *tdig = ls ? dig + lrank : dig;
    int tconj = ls ? rconj : lconj, sconj = ls ? lconj : rconj;
    double sr, si;
    load_elem(kind, s, 0, &sr, &si);
    if (sconj) si = -si;
    if (cplx) {
      cmul_elem(kind, &sr, &si, are, aim);
    } else { /* tensor_block_add receives the complex factor s*alpha (:4064-4075) */
      double pr = sr * are, pi = sr * aim_in;
      sr = c2r(pr, pi); si = 0;
    }
    if (trank != drank) return 8;
    /* transpose the tensor operand into D's order, then D (+)= s * t */
    int trn[MAXR + 1];
    trn[0] = 1;
    for (int i = 0; i < trank; ++i) trn[i + 1] = tdig[i];
    void *tmp = malloc(vd * (size_t)es);
    if (!tmp) return 2;
    int e = cptal_copy_dlf(kind, trank, tdims, trn, t, tmp, 0);
    if (e == 0) axpy(kind, vd, d, tmp, sr, si, tconj, !beta_zero);
    free(tmp);
    return e ? 10 : 0;
  }
  if (ls || rs) return 5;
  /* PARTIAL or FULL contraction: matricize with get_contraction_permutations(1,0) */
  int dprm[MAXR + 1], lprm[MAXR + 1], rprm[MAXR + 1], ncd, nlu, nru, nhu, ierr;
  cptal_get_contraction_permutations(1, 0, lrank, rrank, dig, conj, dprm, lprm, rprm, &ncd, &nlu, &nru,
                                     &nhu, &ierr);
  if (ierr) return 9;
  if (nhu > 0) return 39; /* hyper-contractions need BLAS in the reference (:3858) */
  /* validate extents */
  for (int i = 0; i < lrank; ++i) {
    int v = dig[i];
    if (v > 0) { if (v > drank || ddims[v - 1] != ldims[i]) return 8; }
    else if (rdims[-v - 1] != ldims[i]) return 8;
  }
  for (int i = 0; i < rrank; ++i) {
    int v = dig[lrank + i];
    if (v > 0 && (v > drank || ddims[v - 1] != rdims[i])) return 8;
  }
  size_t vl = volume(lrank, ldims), vr = volume(rrank, rdims);
  size_t lcd = 1;
  for (int i = 0; i < lrank; ++i) if (dig[i] < 0) lcd *= (size_t)ldims[i];
  size_t lld = lcd ? vl / lcd : 0, lrd = lcd ? vr / lcd : 0;
  /* The matricized right operand is R(c, r) in the TN scheme; a conjugated
   * right operand in the reference's BLAS build is stored (r, c) and passed
   * as 'C' to ?gemm. Here the conjugated copy is re-ordered to (c, r) so the
   * mathematics is the same as the BLAS build (see the oracle's header). */
  if (rconj) {
    for (int i = 1; i <= rrank; ++i) {
      if (rprm[i] <= nru) rprm[i] += ncd; else rprm[i] -= nru;
    }
  }
  int ltr = !perm_trivial(lrank, lprm) || lconj;
  int rtr = !perm_trivial(rrank, rprm) || rconj;
  const void *lp = l, *rp = r;
  void *lt = NULL, *rt = NULL, *dt = NULL;
  int e = 0;
  if (ltr) {
    lt = malloc(vl * (size_t)es);
    e = cptal_copy_dlf(kind, lrank, ldims, lprm, l, lt, lconj);
    lp = lt;
  }
  if (!e && rtr) {
    rt = malloc(vr * (size_t)es);
    e = cptal_copy_dlf(kind, rrank, rdims, rprm, r, rt, rconj);
    rp = rt;
  }
  if (e) { free(lt); free(rt); return 10; }
  if (ds) { /* FULL_CONTRACTION */
    double sr = 0, si = 0;
    if (kind == CPTAL_R8 || kind == CPTAL_C8 || kind == CPTAL_R4 || kind == CPTAL_C4) {
      for (size_t c = 0; c < lcd; ++c) {
        double xr, xi, yr, yi;
        load_elem(kind, lp, c, &xr, &xi); load_elem(kind, rp, c, &yr, &yi);
        double pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;
        sr += pr; si += pi;
      }
    }
    double zr, zi;
    load_elem(kind, d, 0, &zr, &zi);
    if (beta_zero) { zr = 0; zi = 0; }
    cmul_elem(kind, &sr, &si, are, aim);
    store_elem(kind, d, 0, zr + sr, zi + si);
    free(lt); free(rt);
    return 0;
  }
  /* PARTIAL_CONTRACTION */
  int dtr = !perm_trivial(drank, dprm);
  void *dp = d;
  int dn2o_inv[MAXR + 1];
  if (dtr) {
    /* dprm is N2O for the matricized D; the copy D -> Dmat uses its inverse as O2N */
    dn2o_inv[0] = 1;
    for (int k = 1; k <= drank; ++k) dn2o_inv[dprm[k]] = k;
    dt = malloc(vd * (size_t)es);
    if (!beta_zero) e = cptal_copy_dlf(kind, drank, ddims, dn2o_inv, d, dt, 0);
    dp = dt;
  }
  if (!e) pcontract(kind, lld, lrd, lcd, lp, rp, dp, are, aim, beta_zero);
  if (!e && dtr) {
    int mdims[MAXR];
    for (int k = 1; k <= drank; ++k) mdims[k - 1] = ddims[dprm[k] - 1];
    e = cptal_copy_dlf(kind, drank, mdims, dprm, dt, d, 0);
  }
  free(lt); free(rt); free(dt);
  return e ? 35 : 0;
}

/* ------------------------------------------------------------------------- */
/* High-precision direct contraction (error reporting)                       */
/* ------------------------------------------------------------------------- */

typedef struct {
  int nfree, ncon;
  size_t dstr[MAXR];                 /* D stride of each D dim */
  int dext[MAXR];
  size_t lstr_d[MAXR], rstr_d[MAXR]; /* L/R stride contributed by each D dim */
  int cext[MAXR];
  size_t lstr_c[MAXR], rstr_c[MAXR];
} direct_plan_t;

static int make_direct_plan(const int *dig, int drank, const int *ddims, int lrank, const int *ldims,
                            int rrank, const int *rdims, direct_plan_t *p)
{
  size_t ls[MAXR + 1], rs[MAXR + 1], s = 1;
  for (int i = 0; i < lrank; ++i) { ls[i] = s; s *= (size_t)ldims[i]; }
  s = 1;
  for (int i = 0; i < rrank; ++i) { rs[i] = s; s *= (size_t)rdims[i]; }
  p->nfree = drank; p->ncon = 0;
  s = 1;
  for (int q = 0; q < drank; ++q) {
    p->dext[q] = ddims[q]; p->dstr[q] = s; s *= (size_t)ddims[q];
    p->lstr_d[q] = 0; p->rstr_d[q] = 0;
  }
  for (int i = 0; i < lrank; ++i) {
    int v = dig[i];
    if (v > 0) { if (v > drank) return 1; p->lstr_d[v - 1] += ls[i]; }
    else {
      int k = p->ncon++;
      p->cext[k] = ldims[i]; p->lstr_c[k] = ls[i]; p->rstr_c[k] = rs[-v - 1];
    }
  }
  for (int i = 0; i < rrank; ++i) {
    int v = dig[lrank + i];
    if (v > 0) { if (v > drank) return 1; p->rstr_d[v - 1] += rs[i]; }
  }
  return 0;
}

static void direct_elem(const direct_plan_t *p, int kind, const void *l, const void *r, size_t dpos,
                        int lconj, int rconj, long double *outr, long double *outi)
{
  size_t lb = 0, rb = 0, rem = dpos;
  for (int q = 0; q < p->nfree; ++q) {
    size_t c = rem % (size_t)p->dext[q];
    rem /= (size_t)p->dext[q];
    lb += c * p->lstr_d[q]; rb += c * p->rstr_d[q];
  }
  size_t nc = 1;
  for (int k = 0; k < p->ncon; ++k) nc *= (size_t)p->cext[k];
  long double sr = 0, si = 0;
  int idx[MAXR] = {0};
  size_t lo = lb, ro = rb;
  for (size_t t = 0; t < nc; ++t) {
    double xr, xi, yr, yi;
    load_elem(kind, l, lo, &xr, &xi); load_elem(kind, r, ro, &yr, &yi);
    if (lconj) xi = -xi;
    if (rconj) yi = -yi;
    sr += (long double)xr * yr - (long double)xi * yi;
    si += (long double)xr * yi + (long double)xi * yr;
    for (int k = 0; k < p->ncon; ++k) {
      lo += p->lstr_c[k]; ro += p->rstr_c[k];
      if (++idx[k] < p->cext[k]) break;
      lo -= p->lstr_c[k] * (size_t)p->cext[k]; ro -= p->rstr_c[k] * (size_t)p->cext[k];
      idx[k] = 0;
    }
  }
  *outr = sr; *outi = si;
}

int cptal_contract_ld_sample(const int *dig, int kind, int drank, const int *ddims, const void *d,
                             int lrank, const int *ldims, const void *l, int rrank,
                             const int *rdims, const void *r, double are, double aim,
                             int conj_bits, int accumulative, size_t ns, const size_t *pos,
                             long double *out)
{
  direct_plan_t p;
  if (cptal_kind_size(kind) == 0) return 6;
  if (make_direct_plan(dig, drank, ddims, lrank, ldims, rrank, rdims, &p)) return 8;
  int cplx = is_complex(kind), lconj = 0, rconj = 0;
  if (cplx) {
    lconj = (conj_bits >> 1) & 1; rconj = (conj_bits >> 2) & 1;
    if (conj_bits & 1) { lconj = !lconj; rconj = !rconj; }
  } else {
    aim = 0;
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t s = 0; s < ns; ++s) {
    long double vr, vi;
    direct_elem(&p, kind, l, r, pos[s], lconj, rconj, &vr, &vi);
    long double tr = vr * are - vi * aim, ti = vr * aim + vi * are;
    double zr = 0, zi = 0;
    if (accumulative) load_elem(kind, d, pos[s], &zr, &zi);
    out[2 * s] = (long double)zr + tr;
    out[2 * s + 1] = (long double)zi + ti;
  }
  return 0;
}

int cptal_contract_ld(const int *dig, int kind, int drank, const int *ddims, const void *d,
                      int lrank, const int *ldims, const void *l, int rrank, const int *rdims,
                      const void *r, double are, double aim, int conj_bits, int accumulative,
                      long double *out)
{
  size_t vd = volume(drank, ddims);
  size_t *pos = malloc((vd ? vd : 1) * sizeof(size_t));
  if (!pos) return 2;
  for (size_t i = 0; i < vd; ++i) pos[i] = i;
  int e = cptal_contract_ld_sample(dig, kind, drank, ddims, d, lrank, ldims, l, rrank, rdims, r, are, aim,
                                   conj_bits, accumulative, vd, pos, out);
  free(pos);
  return e;
}

/* ------------------------------------------------------------------------- */
/* Seeded inputs                                                             */
/* ------------------------------------------------------------------------- */

static inline uint64_t splitmix64(uint64_t *s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Value number i of the stream with seed `seed` (counter-based, so parallel
 * generation gives the same sequence): splitmix64 of seed + (i+1)*golden. */
static inline double unit_at(uint64_t seed, uint64_t i)
{
  uint64_t s = seed + i * 0x9E3779B97F4A7C15ULL;
  uint64_t z = splitmix64(&s);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void cptal_fill_uniform(int kind, size_t vol, void *d, unsigned long long seed, double lo, double hi)
{
  size_t n = vol * (is_complex(kind) ? 2 : 1);
  double w = hi - lo;
  if (kind == CPTAL_R4 || kind == CPTAL_C4) {
    float *p = d;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) p[i] = (float)(lo + w * unit_at(seed, i));
  } else {
    double *p = d;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) p[i] = lo + w * unit_at(seed, i);
  }
}

void cptal_fill_int(size_t vol, double *d, unsigned long long seed)
{
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < vol; ++i) d[i] = floor(unit_at(seed, i) * 2097153.0) - 1048576.0;
}

void cptal_fill_normal(int kind, size_t vol, void *d, unsigned long long seed)
{
  size_t n = vol * (is_complex(kind) ? 2 : 1);
  size_t np = (n + 1) / 2;
#pragma omp parallel for schedule(static)
  for (size_t k = 0; k < np; ++k) {
    double u1 = unit_at(seed, 2 * k), u2 = unit_at(seed, 2 * k + 1);
    if (u1 < 1e-300) u1 = 1e-300;
    double rad = sqrt(-2.0 * log(u1)), th = 6.283185307179586 * u2;
    double z0 = rad * cos(th), z1 = rad * sin(th);
    if (kind == CPTAL_R4 || kind == CPTAL_C4) {
      float *p = d;
      p[2 * k] = (float)z0;
      if (2 * k + 1 < n) p[2 * k + 1] = (float)z1;
    } else {
      double *p = d;
      p[2 * k] = z0;
      if (2 * k + 1 < n) p[2 * k + 1] = z1;
    }
  }
}
