// Host-side index logic of the engine: the symbolic pattern parser, the GEMM
// matricization planner, coherence-control digits, device-id codec, and the
// small shape/signature/device-resource helpers of tensor_algebra.h.
//
// Semantics (including error codes) follow:
//   parser         TALSH/talshf.F90:619-667 over TALSH/tensor_algebra_cpu.F90:5307-5531
//   printer        TALSH/tensor_algebra_cpu.F90:5533-5611
//   permutations   TALSH/tensor_algebra_cpu.F90:5613-5755 and :5757-5956
//   coherence      TALSH/tensor_algebra_gpu.cpp:196-222
//   device ids     TALSH/tensor_algebra_gpu.cpp:225-276
#include <algorithm>
#include <cctype>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

struct Occurrence {
  int arg;          // 0 destination, 1 left, 2 right
  int pos;          // 1-based slot within the argument
  std::string name; // label without the contravariance '+'
  bool paired = false;
};

bool ascii_alnum(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9'); }

// An index label: alphanumeric, optionally closed by one '+' that is not its
// first character.
bool valid_label(const std::string &s)
{
  if (s.empty()) return false;
  for (size_t i = 0; i < s.size(); ++i) {
    if (ascii_alnum(s[i])) continue;
    if (s[i] == '+' && i + 1 == s.size() && i > 0) continue;
    return false;
  }
  return true;
}

// Parses the argument list and pairs the index occurrences. Returns the
// reference's error code (0 success).
int parse_digital(const std::string &s, int &drank, int &lrank, int &rrank, std::vector<int> &dig, int &conj)
{
  drank = lrank = rrank = -1;
  conj = 0;
  if (s.empty()) return 16;
  std::vector<Occurrence> occ;
  int nargs = 0, rank[3] = {0, 0, 0};
  size_t i = 0;
  while (true) {
    size_t open = s.find('(', i);
    if (open == std::string::npos) break;
    if (open >= 2 && s[open - 1] == '+' && (ascii_alnum(s[open - 2]) || s[open - 2] == '_')) conj |= 1 << nargs;
    if (nargs == 3) return 1;
    const int arg = nargs++;
    size_t start = open + 1, k = start;
    int slot = 0;
    bool closed = false, after_sep = false;
    for (; k < s.size(); ++k) {
      char c = s[k];
      if (c != ',' && c != '|' && c != ')') continue;
      std::string label = s.substr(start, k - start);
      if (c == ')' && label.empty()) {
        if (after_sep) return 3;
      } else {
        if (!valid_label(label)) return c == ')' ? 4 : 2;
        if (label.back() == '+') label.pop_back();
        occ.push_back({arg, ++slot, label});
      }
      if (c == ')') { closed = true; break; }
      start = k + 1;
      after_sep = true;
    }
    if (!closed) return 5;
    rank[arg] = slot;
    i = k + 1;
  }
  const bool binary = (nargs == 3);
  if (nargs == 3) { drank = rank[0]; lrank = rank[1]; rrank = rank[2]; }
  else if (nargs == 2) { drank = rank[0]; lrank = rank[1]; }
  else return 15;
  dig.assign((size_t)std::max(0, lrank) + (size_t)std::max(0, rrank), 0);
  for (size_t a = 0; a < occ.size(); ++a) {
    if (occ[a].paired) continue;
    if (a + 1 == occ.size()) return binary ? 7 : 12;  // nothing can follow the final label
    const int arg1 = occ[a].arg, pos1 = occ[a].pos;
    for (int round = 0; round < (binary ? 2 : 1); ++round) {
      size_t b = a + 1;
      while (b < occ.size() && (occ[b].paired || occ[b].name != occ[a].name)) ++b;
      if (b == occ.size()) {
        if (round == 0) return binary ? 11 : 14;
        break;
      }
      const int arg2 = occ[b].arg, pos2 = occ[b].pos;
      if (!binary) {
        if (arg1 != 0 || arg2 != 1) return 13;
        dig[pos2 - 1] = pos1;
      } else if (arg1 == 0 && arg2 == 1) {
        if (round) return 8;
        dig[pos2 - 1] = pos1;
      } else if (arg1 == 0 && arg2 == 2) {
        dig[lrank + pos2 - 1] = pos1;
      } else if (arg1 == 1 && arg2 == 2) {
        if (round) return 9;
        dig[pos1 - 1] = -pos2;
        dig[lrank + pos2 - 1] = -pos1;
      } else {
        return 10;
      }
      occ[b].paired = true;
    }
    occ[a].paired = true;
  }
  return 0;
}

bool valid_pattern(int lrank, int rrank, const int *c, bool allow_hyper)
{
  const int n = lrank + rrank;
  if (n <= 0) return true;
  std::vector<int> uses(n + 1, 0);
  int nfree = 0;
  for (int p = 1; p <= n; ++p) {
    const int v = c[p - 1];
    if (v < 0) {
      const bool left = p <= lrank;
      const int other = -v;
      if (left ? other > rrank : other > lrank) return false;
      const int back = left ? c[lrank + other - 1] : c[other - 1];
      if (back != -(left ? p : p - lrank)) return false;
    } else if (v > 0 && v <= n) {
      if (uses[v] == 0) { uses[v] = 1; ++nfree; }
      else if (allow_hyper && p > lrank && uses[v] == 1) uses[v] = 2;
      else return false;
    } else {
      return false;
    }
  }
  for (int q = 1; q <= nfree; ++q)
    if (uses[q] != 1 && uses[q] != 2) return false;
  return true;
}

// Stable sort of items trn[1..n] by key[item-1], permutation sign in trn[0].
void stable_sort_by_key(int n, const int *key, int *trn)
{
  std::vector<int> items(trn + 1, trn + 1 + n);
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return key[items[a] - 1] < key[items[b] - 1]; });
  // parity of the sorting permutation
  std::vector<int> seen(n, 0);
  int parity = 0;
  for (int i = 0; i < n; ++i) {
    if (seen[i]) continue;
    int len = 0;
    for (int j = i; !seen[j]; j = order[j]) { seen[j] = 1; ++len; }
    parity += len - 1;
  }
  for (int i = 0; i < n; ++i) trn[i + 1] = items[order[i]];
  if (parity & 1) trn[0] = -trn[0];
}

}  // namespace

extern "C" {

int talsh_get_contr_ptrn_str2dig(const char *c_str, int *dig_ptrn, int *drank, int *lrank, int *rrank, int *conj_bits)
{
  *drank = *lrank = *rrank = -1;
  *conj_bits = 0;
  std::string s;
  size_t len = 0;
  for (; c_str[len] != 0; ++len)
    if (len >= 1024) return -1;
  s.assign(c_str, len);
  size_t star = s.find('*');
  if (star != std::string::npos && s.find('(', star) == std::string::npos) {
    s.resize(star);  // trailing scalar factor, e.g. "...*2.0"
    if (s.empty()) return -2;
    if (s.back() != ')') return -3;
    star = std::string::npos;
  }
  if (star == std::string::npos) s += "*R()";  // unary operation cast as a contraction
  while (!s.empty() && s.back() == ' ') s.pop_back();
  std::vector<int> dig;
  int dr, lr, rr, cj;
  int e = parse_digital(s, dr, lr, rr, dig, cj);
  *drank = dr; *lrank = lr; *rrank = rr;
  *conj_bits = (e == 0 || e == 15) ? cj : 0;  // the flags are only published on these paths
  if (e) return e;
  const int n = lr + rr;
  for (int k = 0; k < n; ++k) dig_ptrn[k] = dig[k];
  return 0;
}

void get_contr_pattern_sym(const int *rank_left, const int *rank_right, const int *conj_bits, const int *cptrn_dig,
                           char *cptrn_sym, int *cpl, int *ierr)
{
  const int lr = *rank_left, rr = *rank_right, cj = *conj_bits;
  *ierr = 0;
  *cpl = 0;
  if (lr < 0 || rr < 0) { *ierr = 1; return; }
  auto args = [](std::string &out, const std::vector<char> &labels) {
    for (size_t k = 0; k < labels.size(); ++k) {
      if (k) out += ',';
      out += labels[k];
    }
    out += ')';
  };
  std::string out = (cj & 1) ? "D+(" : "D(";
  int nfree = 0;
  for (int k = 0; k < lr + rr; ++k) nfree += cptrn_dig[k] > 0;
  std::vector<char> d, l, r, contracted(lr + 1, 0);
  char next = 'a';
  for (int k = 0; k < nfree; ++k) d.push_back(next++);
  args(out, d);
  out += (cj & 2) ? "+=L+(" : "+=L(";
  for (int k = 0; k < lr; ++k) {
    const int v = cptrn_dig[k];
    if (v > 0) l.push_back((char)('a' + v - 1));
    else { contracted[k + 1] = next; l.push_back(next++); }
  }
  args(out, l);
  out += (cj & 4) ? "*R+(" : "*R(";
  for (int k = 0; k < rr; ++k) {
    const int v = cptrn_dig[lr + k];
    r.push_back(v > 0 ? (char)('a' + v - 1) : contracted[-v]);
  }
  args(out, r);
  std::memcpy(cptrn_sym, out.c_str(), out.size() + 1);
  *cpl = (int)out.size();
}

void get_contr_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank, const int *cptrn, int conj_bits, int *dprm,
                            int *lprm, int *rprm, int *ncd, int *nlu, int *nru, int *ierr)
{
  *ierr = 0;
  if (!valid_pattern(lrank, rrank, cptrn, false) || lrank < 0 || rrank < 0) { *ierr = 1; return; }
  bool lconj = conj_bits & 2, rconj = conj_bits & 4;
  if (conj_bits & 1) { lconj = !lconj; rconj = !rconj; }
  if (gemm_tl == 0 && lconj) { *ierr = 2; return; }
  // destination: left-free then right-free positions (N2O)
  int drank = 0;
  dprm[0] = 1;
  for (int p = 0; p < lrank; ++p)
    if (cptrn[p] > 0) dprm[++drank] = cptrn[p];
  *nlu = drank;
  bool simple = drank >= 2;
  for (int p = 0; p < rrank; ++p) {
    const int v = cptrn[lrank + p];
    if (v > 0) {
      dprm[++drank] = v;
      if (v <= *nlu) simple = false;
    }
  }
  if (simple)
    for (int q = 1; q <= drank; ++q) dprm[q] = q;
  // right operand (c, r): contracted slots in R order, then free slots
  rprm[0] = 1;
  *ncd = 0;
  *nru = 0;
  int slot = 0;
  if (rrank > 0) {
    for (int p = 0; p < rrank; ++p)
      if (cptrn[lrank + p] < 0) rprm[p + 1] = ++slot;
    *ncd = slot;
    *nru = rrank - slot;
    for (int p = 0; p < rrank; ++p) {
      const int v = cptrn[lrank + p];
      if (v <= 0) continue;
      rprm[p + 1] = (simple && *nru >= 2) ? *ncd + (v - *nlu) : ++slot;
    }
  }
  // left operand (c, l): contracted slots ordered like R's, then free slots
  lprm[0] = 1;
  if (lrank > 0) {
    std::vector<int> pos, key;
    for (int p = 0; p < lrank; ++p)
      if (cptrn[p] < 0) { pos.push_back(p + 1); key.push_back(-cptrn[p]); }
    const int nc = (int)pos.size();
    std::vector<int> trn(nc + 1);
    trn[0] = 1;
    for (int k = 1; k <= nc; ++k) trn[k] = k;
    if (nc >= 2) stable_sort_by_key(nc, key.data(), trn.data());
    for (int k = 1; k <= nc; ++k) lprm[pos[trn[k] - 1]] = k;
    slot = nc;
    for (int p = 0; p < lrank; ++p) {
      if (cptrn[p] <= 0) continue;
      lprm[p + 1] = (simple && *nlu >= 2) ? *ncd + cptrn[p] : ++slot;
    }
  }
  if (gemm_tl == 0)
    for (int p = 1; p <= lrank; ++p) lprm[p] = lprm[p] <= *ncd ? lprm[p] + *nlu : lprm[p] - *ncd;
  if (gemm_tr != 0 || rconj)
    for (int p = 1; p <= rrank; ++p) rprm[p] = rprm[p] <= *ncd ? rprm[p] + *nru : rprm[p] - *ncd;
}

void get_contraction_permutations(int gemm_tl, int gemm_tr, int lrank, int rrank, const int *cptrn, int conj_bits,
                                  int *dprm, int *lprm, int *rprm, int *ncd, int *nlu, int *nru, int *nhu, int *ierr)
{
  *ierr = 0;
  *ncd = *nlu = *nru = *nhu = 0;
  dprm[0] = lprm[0] = rprm[0] = 1;
  if (!valid_pattern(lrank, rrank, cptrn, true)) { *ierr = 3; return; }
  if (lrank < 0 || rrank < 0) { *ierr = 1; return; }
  bool lconj = conj_bits & 2, rconj = conj_bits & 4;
  if (conj_bits & 1) { lconj = !lconj; rconj = !rconj; }
  if (gemm_tl == 0 && lconj) { *ierr = 2; return; }
  const int n = lrank + rrank;
  if (n > 0) {
    int drank = 0, ncontr = 0;
    std::vector<int> count(n + 1, 0);
    for (int p = 0; p < n; ++p) {
      const int v = cptrn[p];
      if (v > 0) {
        drank = std::max(drank, v);
        count[v]++;
        (p < lrank ? *nlu : *nru) += 1;
      } else if (v < 0) {
        ++ncontr;
      }
    }
    *ncd = ncontr / 2;
    // relabel D indices: simple ones 1..s in D order, hyper ones drank+1..
    std::vector<int> label(drank + 1, 0);
    int simple = 0;
    for (int q = 1; q <= drank; ++q) {
      if (count[q] == 1) label[q] = ++simple;
      else if (count[q] == 2) label[q] = drank + (++*nhu);
    }
    *nlu -= *nhu;
    *nru -= *nhu;
    // keys: contracted indices get negative keys in L order, free ones their D label
    std::vector<int> lkey(lrank + 1), rkey(rrank + 1);
    int seen = 0;
    for (int p = 1; p <= lrank; ++p) lkey[p] = cptrn[p - 1] < 0 ? -(*ncd - seen++) : label[cptrn[p - 1]];
    for (int p = 1; p <= rrank; ++p) {
      const int v = cptrn[lrank + p - 1];
      rkey[p] = v < 0 ? lkey[-v] : label[v];
    }
    std::vector<int> dnew(n + 1, 0), trn(n + 1);
    int k = 0;
    auto order = [&](int rank, const std::vector<int> &key, int *prm, bool right) {
      trn[0] = 1;
      for (int q = 1; q <= rank; ++q) trn[q] = q;
      if (rank > 1) stable_sort_by_key(rank, key.data() + 1, trn.data());
      for (int q = 1; q <= rank; ++q) {
        const int kk = key[trn[q]];
        if (kk >= 1 && kk <= drank) dnew[++k] = kk;
        else if (right && kk > drank) { ++k; dnew[k] = k; }
      }
      prm[0] = trn[0];
      for (int q = 1; q <= rank; ++q) prm[trn[q]] = q;
    };
    if (lrank > 0) order(lrank, lkey, lprm, false);
    if (rrank > 0) order(rrank, rkey, rprm, true);
    if (drank > 0) {
      dprm[0] = 1;
      for (int q = 1; q <= drank; ++q) dprm[q] = dnew[q];
    }
  }
  if (gemm_tl == 0) {
    for (int p = 1; p <= lrank; ++p) {
      if (lprm[p] <= *ncd) lprm[p] += *nlu;
      else if (lprm[p] <= *ncd + *nlu) lprm[p] -= *ncd;
    }
  }
  if (gemm_tr != 0 || rconj) {
    for (int p = 1; p <= rrank; ++p) {
      if (rprm[p] <= *ncd) rprm[p] += *nru;
      else if (rprm[p] <= *ncd + *nru) rprm[p] -= *ncd;
    }
  }
}

int permutation_trivial(const int perm_len, const int *perm, const int base)
{
  for (int i = 0; i < perm_len; ++i)
    if (perm[i] != base + i) return NOPE;
  return YEP;
}

size_t fortran_cptr_int(void *cptr) { return (size_t)cptr; }

int tens_valid_data_kind(int datk, int *datk_size)
{
  const int s = xt::kind_size(datk);
  if (datk_size) *datk_size = datk == NO_TYPE ? 0 : s;
  return (s > 0 || datk == NO_TYPE) ? YEP : NOPE;
}
int tens_valid_data_kind_(int datk, int *datk_size) { return tens_valid_data_kind(datk, datk_size); }
int talshValidDataKind(int datk, int *datk_size) { return tens_valid_data_kind(datk, datk_size); }

size_t tens_elem_offset_f(unsigned int num_dim, const unsigned int *dims, const unsigned int *mlndx)
{
  size_t off = 0;
  for (int q = (int)num_dim - 1; q >= 0; --q) off = off * dims[q] + mlndx[q];
  return off;
}

void tens_elem_mlndx_f(size_t offset, unsigned int num_dim, const unsigned int *dims, unsigned int *mlndx)
{
  for (unsigned q = 0; q < num_dim; ++q) {
    mlndx[q] = (unsigned)(offset % dims[q]);
    offset /= dims[q];
  }
}

unsigned int argument_coherence_get_value(unsigned int coh_ctrl, unsigned int tot_args, unsigned int arg_num)
{
  if (arg_num >= tot_args || tot_args > MAX_TENSOR_OPERANDS) return 0xFFFFFFFFu;
  return (coh_ctrl >> (2 * (tot_args - 1 - arg_num))) & 3u;
}

int argument_coherence_set_value(unsigned int *coh_ctrl, unsigned int tot_args, unsigned int arg_num,
                                 unsigned int coh_val)
{
  if (arg_num >= tot_args || tot_args > MAX_TENSOR_OPERANDS || coh_val > 3) return 1;
  const unsigned sh = 2 * (tot_args - 1 - arg_num);
  *coh_ctrl = (*coh_ctrl & ~(3u << sh)) | (coh_val << sh);
  return 0;
}

int valid_device_kind(int dev_kind)
{
  return (dev_kind == DEV_NULL || dev_kind == DEV_HOST || dev_kind == DEV_NVIDIA_GPU || dev_kind == DEV_INTEL_MIC ||
          dev_kind == DEV_AMD_GPU)
             ? YEP
             : NOPE;
}

int encode_device_id(int dev_kind, int dev_num)
{
  switch (dev_kind) {
  case DEV_HOST: return dev_num == 0 ? 0 : DEV_MAX;
  case DEV_NVIDIA_GPU: return (dev_num >= 0 && dev_num < MAX_GPUS_PER_NODE) ? 1 + dev_num : DEV_MAX;
  case DEV_INTEL_MIC:
    return (dev_num >= 0 && dev_num < MAX_MICS_PER_NODE) ? 1 + MAX_GPUS_PER_NODE + dev_num : DEV_MAX;
  case DEV_AMD_GPU:
    return (dev_num >= 0 && dev_num < MAX_AMDS_PER_NODE) ? 1 + MAX_GPUS_PER_NODE + MAX_MICS_PER_NODE + dev_num
                                                         : DEV_MAX;
  default: return DEV_MAX;
  }
}

int decode_device_id(int dev_id, int *dev_kind)
{
  int kind = DEV_NULL, num = -1;
  if (dev_id == 0) { kind = DEV_HOST; num = 0; }
  else if (dev_id >= 1 && dev_id <= MAX_GPUS_PER_NODE) { kind = DEV_NVIDIA_GPU; num = dev_id - 1; }
  else if (dev_id > MAX_GPUS_PER_NODE && dev_id <= MAX_GPUS_PER_NODE + MAX_MICS_PER_NODE) {
    kind = DEV_INTEL_MIC; num = dev_id - 1 - MAX_GPUS_PER_NODE;
  } else if (dev_id > MAX_GPUS_PER_NODE + MAX_MICS_PER_NODE && dev_id < DEV_MAX) {
    kind = DEV_AMD_GPU; num = dev_id - 1 - MAX_GPUS_PER_NODE - MAX_MICS_PER_NODE;
  }
  if (dev_kind) *dev_kind = kind;
  return num;
}

int talsh_tens_no_init(const talsh_tens_data_t *, const talsh_tens_shape_t *, const talsh_tens_signature_t *)
{
  return 0;
}

// ---- shapes -------------------------------------------------------------
int tensShape_create(talsh_tens_shape_t **tshape)
{
  if (!tshape) return TALSH_INVALID_ARGS;
  *tshape = (talsh_tens_shape_t *)std::malloc(sizeof(talsh_tens_shape_t));
  if (!*tshape) return TRY_LATER;
  return tensShape_clean(*tshape);
}

int tensShape_clean(talsh_tens_shape_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  t->num_dim = -1;
  t->dims = t->divs = t->grps = nullptr;
  return 0;
}

int tensShape_construct(talsh_tens_shape_t *t, int pinned, int rank, const int *dims, const int *divs, const int *grps)
{
  (void)pinned;
  if (!t || rank < 0 || rank > MAX_TENSOR_RANK) return TALSH_INVALID_ARGS;
  if (t->num_dim >= 0) return TALSH_OBJECT_NOT_EMPTY;
  if (rank > 0 && !dims) return TALSH_INVALID_ARGS;
  for (int q = 0; q < rank; ++q)
    if (dims[q] <= 0) return TALSH_INVALID_ARGS;
  t->num_dim = rank;
  if (rank > 0) {
    t->dims = (int *)std::malloc(sizeof(int) * rank * 3);
    if (!t->dims) { t->num_dim = -1; return TRY_LATER; }
    t->divs = t->dims + rank;
    t->grps = t->dims + 2 * rank;
    for (int q = 0; q < rank; ++q) {
      t->dims[q] = dims[q];
      t->divs[q] = divs ? divs[q] : dims[q];
      t->grps[q] = grps ? grps[q] : 0;
    }
  }
  return 0;
}

int tensShape_destruct(talsh_tens_shape_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  std::free(t->dims);
  return tensShape_clean(t);
}

int tensShape_destroy(talsh_tens_shape_t *t)
{
  if (!t) return TALSH_INVALID_ARGS;
  tensShape_destruct(t);
  std::free(t);
  return 0;
}

size_t tensShape_volume(const talsh_tens_shape_t *t)
{
  if (!t || t->num_dim < 0) return 0;
  size_t v = 1;
  for (int q = 0; q < t->num_dim; ++q) v *= (size_t)t->dims[q];
  return v;
}

int tensShape_rank(const talsh_tens_shape_t *t) { return t ? t->num_dim : -1; }

int tensShape_reshape(talsh_tens_shape_t *t, int rank, const int *dims, const int *divs, const int *grps)
{
  if (!t || t->num_dim < 0) return TALSH_INVALID_ARGS;
  talsh_tens_shape_t n;
  tensShape_clean(&n);
  int e = tensShape_construct(&n, NOPE, rank, dims, divs, grps);
  if (e) return e;
  if (tensShape_volume(&n) != tensShape_volume(t)) { tensShape_destruct(&n); return TALSH_INVALID_ARGS; }
  tensShape_destruct(t);
  *t = n;
  return 0;
}

void tensShape_print(const talsh_tens_shape_t *t)
{
  if (!t) return;
  std::printf("#MESSAGE: Tensor shape rank %d:", t->num_dim);
  for (int q = 0; q < t->num_dim; ++q) std::printf(" %d", t->dims[q]);
  std::printf("\n");
}

// ---- signatures -------------------------------------------------------------
int tensSignature_create(talsh_tens_signature_t **s)
{
  if (!s) return TALSH_INVALID_ARGS;
  *s = (talsh_tens_signature_t *)std::malloc(sizeof(talsh_tens_signature_t));
  if (!*s) return TRY_LATER;
  return tensSignature_clean(*s);
}

int tensSignature_clean(talsh_tens_signature_t *s)
{
  if (!s) return TALSH_INVALID_ARGS;
  s->num_dim = -1;
  s->offsets = nullptr;
  return 0;
}

int tensSignature_construct(talsh_tens_signature_t *s, int rank, const size_t *offsets)
{
  if (!s || rank < 0 || rank > MAX_TENSOR_RANK) return TALSH_INVALID_ARGS;
  if (s->num_dim >= 0) return TALSH_OBJECT_NOT_EMPTY;
  s->num_dim = rank;
  if (rank > 0) {
    s->offsets = (size_t *)std::malloc(sizeof(size_t) * rank);
    if (!s->offsets) { s->num_dim = -1; return TRY_LATER; }
    for (int q = 0; q < rank; ++q) s->offsets[q] = offsets ? offsets[q] : 0;
  }
  return 0;
}

int tensSignature_destruct(talsh_tens_signature_t *s)
{
  if (!s) return TALSH_INVALID_ARGS;
  std::free(s->offsets);
  return tensSignature_clean(s);
}

int tensSignature_destroy(talsh_tens_signature_t *s)
{
  if (!s) return TALSH_INVALID_ARGS;
  tensSignature_destruct(s);
  std::free(s);
  return 0;
}

}  // extern "C"
