// Contraction planning (TTGT matricization candidates and fused-gather
// candidates), their features, and the learned plan selector that picks
// among them.
//
// The reference fixes one matricization, D(ll x lr) += L^T(lc x ll) R(lc x lr)
// with every operand transposed into that order
// (TALSH/tensor_algebra_cpu.F90:5613-5755, get_contr_permutations), and fixed
// heuristics elsewhere (TALSH/tensor_algebra_gpu_nvidia.cu:1188-1272,
// TALSH/talshc.cpp:366-488). Here the GEMM reads either operand k-fast or
// m/n-fast and may compute D^T, and a fused-gather plan reads the tensors
// where they lie through offset tables, so several candidates exist; the
// byte-count heuristic ranks the transposing ones and an optional learned
// model (a small MLP predicting the log of kernel time from the features
// below) chooses among all candidates and GEMM tile shapes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "contract_plan.hpp"
#include "kernels.hpp"
#include "mlp.hpp"
#include "selector.hpp"

namespace xt {

// o2n (0-based) that reorders labels `from` into the order `to`.
std::vector<int> reorder(const std::vector<int> &from, const std::vector<int> &to)
{
  std::vector<int> o2n(from.size());
  for (size_t i = 0; i < from.size(); ++i)
    o2n[i] = (int)(std::find(to.begin(), to.end(), from[i]) - to.begin());
  return o2n;
}

bool is_identity(const std::vector<int> &p)
{
  for (size_t i = 0; i < p.size(); ++i)
    if (p[i] != (int)i) return false;
  return true;
}

static std::vector<int> concat(const std::vector<int> &a, const std::vector<int> &b)
{
  std::vector<int> c(a);
  c.insert(c.end(), b.begin(), b.end());
  return c;
}

// Heuristic plan: among the index orders that keep an operand in place
// (contracted indices in L's or R's order, free indices in their own or D's
// order, optionally computing D^T), pick the one moving the fewest bytes.
int plan_contraction(const std::vector<int> &dig, int dr, int lr, int rr, const std::vector<int> &dd,
                     const std::vector<int> &ld, const std::vector<int> &rd, int es, bool acc, int variant,
                     ContractPlan &P, std::vector<ContractPlan> *all)
{
  std::vector<int> llab(lr), rlab(rr), dlab(dr);
  std::vector<int> ext(256, 0);
  for (int q = 0; q < dr; ++q) { dlab[q] = q + 1; ext[q + 1] = dd[q]; }
  for (int i = 0; i < lr; ++i) {
    llab[i] = dig[i] > 0 ? dig[i] : 100 + i + 1;
    if (dig[i] < 0) ext[100 + i + 1] = ld[i];
  }
  for (int j = 0; j < rr; ++j) rlab[j] = dig[lr + j] > 0 ? dig[lr + j] : 100 + (-dig[lr + j]);
  std::vector<int> Kl, Kr, Ml, Nr, Md, Nd;
  for (int i = 0; i < lr; ++i) (dig[i] < 0 ? Kl : Ml).push_back(llab[i]);
  for (int j = 0; j < rr; ++j) (dig[lr + j] < 0 ? Kr : Nr).push_back(rlab[j]);
  for (int lab : Ml)
    if (std::find(Nr.begin(), Nr.end(), lab) != Nr.end()) return TALSH_INVALID_ARGS;  // hyper-index
  for (int q = 0; q < dr; ++q) {
    if (std::find(Ml.begin(), Ml.end(), dlab[q]) != Ml.end()) Md.push_back(dlab[q]);
    else Nd.push_back(dlab[q]);
  }
  auto vol = [&](const std::vector<int> &labs) {
    size_t v = 1;
    for (int l : labs) v *= (size_t)ext[l];
    return v;
  };
  for (int i = 0; i < lr; ++i) ext[llab[i]] = ld[i];
  const size_t vl = vol(llab), vr = vol(rlab), vd = vol(dlab);
  std::vector<ContractPlan> cands;
  for (int ko = 0; ko < 2; ++ko)
    for (int mo = 0; mo < 2; ++mo)
      for (int no = 0; no < 2; ++no)
        for (int sw = 0; sw < 2; ++sw) {
          const std::vector<int> &K = ko ? Kr : Kl, &M = mo ? Md : Ml, &N = no ? Nd : Nr;
          ContractPlan c;
          c.kase = 3;
          c.swap = sw;
          c.m = vol(M); c.n = vol(N); c.k = vol(K);
          double cost = 0;
          const auto KM = concat(K, M), MK = concat(M, K), KN = concat(K, N), NK = concat(N, K);
          bool lk = true, rk = true;
          if (llab == KM) lk = true;
          else if (llab == MK) lk = false;
          else { c.l_o2n = reorder(llab, KM); cost += 2.0 * vl * es; }
          if (rlab == KN) rk = true;
          else if (rlab == NK) rk = false;
          else { c.r_o2n = reorder(rlab, KN); cost += 2.0 * vr * es; }
          const auto DT = sw ? concat(N, M) : concat(M, N);
          if (dlab != DT) {
            c.d_in_o2n = reorder(dlab, DT);
            c.d_out_o2n = reorder(DT, dlab);
            cost += (acc ? 4.0 : 2.0) * vd * es;
            for (int lab : DT) c.dmat_dims.push_back(ext[lab]);
          }
          if (!sw) { c.a_kfast = lk; c.b_kfast = rk; }
          else { c.a_kfast = rk; c.b_kfast = lk; }
          c.bytes_moved = cost;
          cands.push_back(c);
        }
  // deterministic preference on ties: fewer bytes, then k-fast operands, then no swap
  auto score = [](const ContractPlan &c) {
    return c.bytes_moved * 4.0 + (c.a_kfast ? 0 : 1) + (c.b_kfast ? 0 : 1) + (c.swap ? 0.5 : 0);
  };
  std::stable_sort(cands.begin(), cands.end(),
                   [&](const ContractPlan &a, const ContractPlan &b) { return score(a) < score(b); });
  // orders that coincide (e.g. L's free indices already in D's order) give
  // the same launches: keep the first of each
  std::vector<ContractPlan> uniq;
  for (const ContractPlan &c : cands) {
    bool dup = false;
    for (const ContractPlan &u : uniq)
      if (u.swap == c.swap && u.a_kfast == c.a_kfast && u.b_kfast == c.b_kfast && u.l_o2n == c.l_o2n &&
          u.r_o2n == c.r_o2n && u.d_in_o2n == c.d_in_o2n && u.m == c.m && u.n == c.n && u.k == c.k) {
        dup = true;
        break;
      }
    if (!dup) uniq.push_back(c);
  }
  cands.swap(uniq);
  if (all) *all = cands;
  if (variant >= kFusedVariantBase) {
    std::vector<ContractPlan> fz;
    const int e = plan_fused(dig, dr, lr, rr, dd, ld, rd, fz);
    if (e) return e;
    if (variant - kFusedVariantBase < (int)fz.size()) {
      P = fz[variant - kFusedVariantBase];
      return 0;
    }
  }
  if (variant >= 0 && variant < (int)cands.size()) P = cands[variant];
  else P = cands.front();
  return 0;
}

namespace {

// Index group of one tensor for the fused gather: the labels' extents and
// strides in that tensor, unit extents dropped and runs that stay contiguous
// (stride[j + 1] == stride[j] * ext[j]) merged. false if more than
// kGatherMaxDims dimensions remain.
bool gather_group(const std::vector<int> &labs, const std::vector<int> &ext, const std::vector<int64_t> &str,
                  GatherGroup &g)
{
  g = GatherGroup{};
  std::vector<int64_t> e, s;
  for (int lab : labs) {
    if (ext[lab] == 1) continue;
    if (!e.empty() && str[lab] == s.back() * e.back()) {
      e.back() *= ext[lab];
      continue;
    }
    e.push_back(ext[lab]);
    s.push_back(str[lab]);
  }
  if ((int)e.size() > kGatherMaxDims) return false;
  for (size_t j = 0; j < e.size(); ++j) {
    if (e[j] > INT32_MAX) return false;
    g.ext[j] = (int)e[j];
    g.str[j] = s[j];
  }
  g.nd = (int)e.size();
  return true;
}

}  // namespace

int plan_fused(const std::vector<int> &dig, int dr, int lr, int rr, const std::vector<int> &dd,
               const std::vector<int> &ld, const std::vector<int> &rd, std::vector<ContractPlan> &out)
{
  out.clear();
  // labels: D index q -> q + 1; contracted index i of L -> 101 + i
  std::vector<int> ext(256, 1);
  std::vector<int64_t> sl(256, 0), sr(256, 0), sd(256, 0);
  std::vector<int> llab(lr), rlab(rr), Kl, Kr, M, N;
  int64_t st = 1;
  for (int q = 0; q < dr; ++q) { ext[q + 1] = dd[q]; sd[q + 1] = st; st *= dd[q]; }
  st = 1;
  for (int i = 0; i < lr; ++i) {
    llab[i] = dig[i] > 0 ? dig[i] : 100 + i + 1;
    ext[llab[i]] = ld[i];
    sl[llab[i]] = st;
    st *= ld[i];
    (dig[i] < 0 ? Kl : M).push_back(llab[i]);
  }
  st = 1;
  for (int j = 0; j < rr; ++j) {
    rlab[j] = dig[lr + j] > 0 ? dig[lr + j] : 100 + (-dig[lr + j]);
    sr[rlab[j]] = st;
    st *= rd[j];
    (dig[lr + j] < 0 ? Kr : N).push_back(rlab[j]);
  }
  for (int lab : M)
    if (std::find(N.begin(), N.end(), lab) != N.end()) return TALSH_INVALID_ARGS;  // hyper-index
  auto vol = [&](const std::vector<int> &labs) {
    size_t v = 1;
    for (int l : labs) v *= (size_t)ext[l];
    return v;
  };
  // the tensor's stride-1 label (first non-unit dimension), 0 if none
  auto unit_label = [&](const std::vector<int> &labs) {
    for (int l : labs)
      if (ext[l] > 1) return l;
    return 0;
  };
  const int l1 = unit_label(llab), r1 = unit_label(rlab);
  // staging along k for an operand whose stride-1 index is contracted, else
  // along its rows
  const bool lkf = l1 != 0 && std::find(Kl.begin(), Kl.end(), l1) != Kl.end();
  const bool rkf = r1 != 0 && std::find(Kr.begin(), Kr.end(), r1) != Kr.end();
  // row order: the operand's own order when it is staged along rows (its
  // stride-1 index first), else D's order (the result's rows and the split-K
  // sum then write along D's stride-1 index)
  auto by_d = [&](std::vector<int> &labs) {
    std::stable_sort(labs.begin(), labs.end(), [&](int a, int b) { return sd[a] < sd[b]; });
  };
  if (lkf) by_d(M);
  if (rkf) by_d(N);
  for (int ko = 0; ko < 2; ++ko) {
    const std::vector<int> &K = ko ? Kr : Kl;
    if (ko && Kr == Kl) break;
    for (int sw = 0; sw < 2; ++sw) {
      ContractPlan c;
      c.kase = 3;
      c.fused = true;
      c.swap = sw;
      c.m = vol(M); c.n = vol(N); c.k = vol(K);
      GatherGroup lrow, lk, rrow, rk, dm, dn;
      if (!gather_group(M, ext, sl, lrow) || !gather_group(K, ext, sl, lk) || !gather_group(N, ext, sr, rrow) ||
          !gather_group(K, ext, sr, rk) || !gather_group(M, ext, sd, dm) || !gather_group(N, ext, sd, dn))
        return 0;
      if (!sw) {
        c.maps = GatherMaps{lrow, lk, rrow, rk, dm, dn};
        c.a_kfast = lkf; c.b_kfast = rkf;
      } else {
        c.maps = GatherMaps{rrow, rk, lrow, lk, dn, dm};
        c.a_kfast = rkf; c.b_kfast = lkf;
      }
      out.push_back(c);
    }
  }
  return 0;
}


// ---- features ------------------------------------------------------------------
namespace {

inline float lg2(double v) { return (float)std::log2(v < 1.0 ? 1.0 : v); }

// Shape of a DLF permutation after dropping unit extents and fusing index runs
// that stay adjacent: fused rank, contiguous input run and contiguous output
// run (elements).
void perm_shape(const std::vector<int> &dims, const std::vector<int> &o2n, double *in_run, double *out_run,
                int *frank)
{
  std::vector<int> d, p;
  for (size_t i = 0; i < dims.size(); ++i)
    if (dims[i] > 1) { d.push_back(dims[i]); p.push_back(o2n[i]); }
  const int r = (int)d.size();
  std::vector<int> idx(r), rank(r);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return p[a] < p[b]; });
  for (int q = 0; q < r; ++q) rank[idx[q]] = q;
  std::vector<std::pair<double, int>> groups;  // (extent, first output rank)
  for (int i = 0; i < r; ++i) {
    if (i > 0 && rank[i] == rank[i - 1] + 1) groups.back().first *= d[i];
    else groups.push_back({(double)d[i], rank[i]});
  }
  *frank = (int)groups.size();
  *in_run = groups.empty() ? 1.0 : groups[0].first;
  *out_run = 1.0;
  for (auto &g : groups)
    if (g.second == 0) *out_run = g.first;
}

void put_perm(float *f, const std::vector<int> &dims, const std::vector<int> &o2n, double bytes)
{
  if (o2n.empty()) return;  // no transpose: all zero
  double ir, orr;
  int fr;
  perm_shape(dims, o2n, &ir, &orr, &fr);
  f[0] = 1.f;
  f[1] = lg2(bytes);
  f[2] = lg2(ir);
  f[3] = lg2(orr);
  f[4] = (float)fr;
}

void put_plan(float *f, int kind, const std::vector<int> &dims, const std::vector<int> &o2n)
{
  if (o2n.empty()) return;
  PermuteInfo pi{};
  if (plan_permute_info(kind, (int)dims.size(), dims.data(), o2n.data(), &pi, false)) return;
  f[0] = (float)std::log2((double)std::max(pi.vec, 1));
  f[1] = pi.path == 2;
}

}  // namespace

int plan_candidates(const std::vector<int> &dig, int dr, int lr, int rr, const std::vector<int> &dd,
                    const std::vector<int> &ld, const std::vector<int> &rd, int kind, R4Mode r4mode, bool acc,
                    int num_cus, std::vector<PlanCandidate> &out)
{
  out.clear();
  const int es = kind_size(kind);
  if (!es) return TALSH_INVALID_ARGS;
  ContractPlan P;
  std::vector<ContractPlan> all;
  const int e = plan_contraction(dig, dr, lr, rr, dd, ld, rd, es, acc, -1, P, &all);
  if (e) return e;
  size_t vl = 1, vr = 1, vd = 1;
  for (int x : ld) vl *= (size_t)x;
  for (int x : rd) vr *= (size_t)x;
  for (int x : dd) vd *= (size_t)x;
  const bool cplx = kind_complex(kind);
  const int ntiles = gemm_tile_choices(kind, r4mode);
  for (int v = 0; v < (int)all.size(); ++v) {
    const ContractPlan &c = all[v];
    const double gm = (double)(c.swap ? c.n : c.m), gn = (double)(c.swap ? c.m : c.n), gk = (double)c.k;
    for (int t = 0; t < ntiles; ++t) {
      PlanCandidate pc;
      pc.variant = v;
      pc.tile = t;
      float *f = pc.f;
      f[0] = kind == R4; f[1] = kind == R8; f[2] = kind == C4; f[3] = kind == C8;
      f[4] = kind == R4 && r4mode == R4Mode::kBF16x3;
      f[5] = acc;
      f[6] = lg2(gm); f[7] = lg2(gn); f[8] = lg2(gk);
      f[9] = lg2((cplx ? 8.0 : 2.0) * gm * gn * gk);
      f[10] = c.a_kfast; f[11] = c.b_kfast; f[12] = c.swap;
      f[13 + t] = 1.f;  // tile one-hot (3 slots)
      f[16] = lg2(1.0 + c.bytes_moved);
      f[17] = lg2((double)es * (gm * gk + gn * gk + gm * gn * (acc ? 2.0 : 1.0)));
      put_perm(f + 18, ld, c.l_o2n, 2.0 * vl * es);
      put_perm(f + 23, rd, c.r_o2n, 2.0 * vr * es);
      if (!c.d_out_o2n.empty()) put_perm(f + 28, c.dmat_dims, c.d_out_o2n, (acc ? 4.0 : 2.0) * vd * es);
      int bm, bn, occ;
      gemm_tile_shape(kind, r4mode, t, &bm, &bn, &occ);
      const double tm = std::ceil(gm / bm), tn = std::ceil(gn / bn);
      f[33] = lg2(tm * bm * tn * bn / (gm * gn));
      f[34] = (float)std::log2(tm * tn / (double)(num_cus > 0 ? num_cus : 256));
      f[35] = (float)occ;
      // 16-byte vector loads of the GEMM operands (k-fast rows of 16-byte multiples;
      // a_kfast / b_kfast already describe the GEMM's A and B after a swap)
      const bool kal = ((size_t)gk * es) % 16 == 0;
      f[36] = c.a_kfast && kal && !cplx;
      f[37] = c.b_kfast && kal && !cplx;
      // transpose kernel plans: vector width and the scatter fallback
      f[53] = lg2(gemm_splitk_count(kind, r4mode, (size_t)gm, (size_t)gn, (size_t)gk, t, num_cus > 0 ? num_cus : 256));
      put_plan(f + 38, kind, ld, c.l_o2n);
      put_plan(f + 40, kind, rd, c.r_o2n);
      if (!c.d_out_o2n.empty()) put_plan(f + 42, kind, c.dmat_dims, c.d_out_o2n);
      out.push_back(pc);
    }
  }
  const bool split3 = (kind == R4 || kind == C4) && r4mode == R4Mode::kBF16x3;
  std::vector<ContractPlan> fz;
  if (split3 || plan_fused(dig, dr, lr, rr, dd, ld, rd, fz) != 0) return 0;
  for (int v = 0; v < (int)fz.size(); ++v) {
    const ContractPlan &c = fz[v];
    const double gm = (double)(c.swap ? c.n : c.m), gn = (double)(c.swap ? c.m : c.n), gk = (double)c.k;
    const GatherGroup &sa = c.a_kfast ? c.maps.a_k : c.maps.a_row, &sb = c.b_kfast ? c.maps.b_k : c.maps.b_row;
    auto unit_run = [](const GatherGroup &g) { return (g.nd > 0 && g.str[0] == 1) ? (double)g.ext[0] : 1.0; };
    auto first_ext = [](const GatherGroup &g) { return g.nd > 0 ? g.ext[0] : 1; };
    for (int t = 0; t < ntiles; ++t) {
      PlanCandidate pc;
      pc.variant = kFusedVariantBase + v;
      pc.tile = t;
      float *f = pc.f;
      f[0] = kind == R4; f[1] = kind == R8; f[2] = kind == C4; f[3] = kind == C8;
      f[5] = acc;
      f[6] = lg2(gm); f[7] = lg2(gn); f[8] = lg2(gk);
      f[9] = lg2((cplx ? 8.0 : 2.0) * gm * gn * gk);
      f[10] = c.a_kfast; f[11] = c.b_kfast; f[12] = c.swap;
      f[13 + t] = 1.f;
      f[17] = lg2((double)es * (gm * gk + gn * gk + gm * gn * (acc ? 2.0 : 1.0)));
      int bm, bn, occ;
      gemm_tile_shape(kind, r4mode, t, &bm, &bn, &occ);
      const double tm = std::ceil(gm / bm), tn = std::ceil(gn / bn);
      f[33] = lg2(tm * bm * tn * bn / (gm * gn));
      f[34] = (float)std::log2(tm * tn / (double)(num_cus > 0 ? num_cus : 256));
      f[35] = (float)occ;
      // 16-byte groups along each operand's staging direction
      f[36] = !cplx && unit_run(sa) * es >= 16 && ((int)unit_run(sa) * es) % 16 == 0;
      f[37] = !cplx && unit_run(sb) * es >= 16 && ((int)unit_run(sb) * es) % 16 == 0;
      f[44] = 1.f;  // fused gather
      const int run = bm / 16;  // elements one thread stages along its direction
      f[45] = first_ext(sa) % run == 0 && first_ext(sb) % run == 0;
      f[46] = c.maps.c_row.nd > 0 && c.maps.c_row.str[0] == 1;  // result rows along D's stride-1 index
      // contiguous run along each operand's staging direction, and the
      // stride between its staged elements when that is not 1
      f[47] = lg2(unit_run(sa));
      f[48] = lg2(unit_run(sb));
      f[49] = lg2(sa.nd > 0 ? (double)sa.str[0] : 1.0);
      f[50] = lg2(sb.nd > 0 ? (double)sb.str[0] : 1.0);
      // runs of the result's rows and columns along D's stride-1 index
      f[51] = lg2(unit_run(c.maps.c_row));
      f[52] = lg2(unit_run(c.maps.c_col));
      f[53] = lg2(gemm_splitk_count(kind, r4mode, (size_t)gm, (size_t)gn, (size_t)gk, t, num_cus > 0 ? num_cus : 256));
      out.push_back(pc);
    }
  }
  return 0;
}

// ---- host inference and the installed selector --------------------------------
float mlp_forward_host(const MlpShape &s, const float *params, const float *x)
{
  float a[kMlpMaxWidth], b[kMlpMaxWidth];
  std::memcpy(a, x, sizeof(float) * s.nin);
  for (int l = 0; l <= s.hidden; ++l) {
    size_t wo, bo;
    int ni, no;
    mlp_layer_offsets(s, l, &wo, &bo, &ni, &no);
    const float *W = params + wo, *bias = params + bo;
    const bool last = l == s.hidden;
    for (int o = 0; o < no; ++o) {
      float acc = bias[o];
      for (int i = 0; i < ni; ++i) acc = std::fmaf(a[i], W[i * no + o], acc);
      b[o] = last ? acc : std::max(acc, 0.f);
    }
    std::memcpy(a, b, sizeof(float) * no);
  }
  return a[0];
}

float MlpModel::predict_host(const float *x) const
{
  float xn[kMlpMaxWidth];
  normalize(x, xn);
  return mlp_forward_host(s, params.data(), xn) * ystd + ymean;
}

std::string plan_key(const std::vector<int> &dig, int lr, int rr, const std::vector<int> &dd,
                     const std::vector<int> &ld, const std::vector<int> &rd, int kind, R4Mode r4mode, bool acc,
                     int num_cus)
{
  std::string key;
  key.reserve(64);
  auto put = [&](int v) { key.append(std::to_string(v)).push_back(','); };
  put(kind); put((int)r4mode); put(acc); put(lr); put(rr); put(num_cus);
  for (int v : dig) put(v);
  key.push_back('|');
  for (int v : dd) put(v);
  key.push_back('|');
  for (int v : ld) put(v);
  key.push_back('|');
  for (int v : rd) put(v);
  return key;
}

namespace {
std::mutex g_sel_mu;
MlpModel *g_sel = nullptr;
std::unordered_map<std::string, std::pair<int, int>> g_sel_memo;
}  // namespace

void selector_install(const MlpModel *m)
{
  std::lock_guard<std::mutex> g(g_sel_mu);
  delete g_sel;
  g_sel = m ? new MlpModel(*m) : nullptr;
  g_sel_memo.clear();
}

bool selector_installed()
{
  std::lock_guard<std::mutex> g(g_sel_mu);
  return g_sel != nullptr;
}

bool selector_choose(const std::vector<int> &dig, int dr, int lr, int rr, const std::vector<int> &dd,
                     const std::vector<int> &ld, const std::vector<int> &rd, int kind, R4Mode r4mode, bool acc,
                     int num_cus, int *variant, int *tile)
{
  std::lock_guard<std::mutex> g(g_sel_mu);
  if (!g_sel) return false;
  if (kind != R8 && kind != R4 && kind != C8) return false;  // the data kinds it was trained on
  {  // outside the model's training domain: no choice
    double fl = 2.0, vl = 1, vr = 1, vd = 1;  // real flop count, as the domain is defined
    for (size_t i = 0; i < ld.size(); ++i) { fl *= ld[i]; vl *= ld[i]; }
    for (size_t j = 0; j < rd.size(); ++j) {
      vr *= rd[j];
      if (dig[lr + j] > 0) fl *= rd[j];
    }
    for (int x : dd) vd *= x;
    if (std::log2(fl) > kSelectorMaxLog2Flops || std::log2(std::max(vl, std::max(vr, vd))) > kSelectorMaxLog2Volume)
      return false;
  }
  const std::string key = plan_key(dig, lr, rr, dd, ld, rd, kind, r4mode, acc, num_cus);
  auto it = g_sel_memo.find(key);
  if (it != g_sel_memo.end()) {
    *variant = it->second.first;
    *tile = it->second.second;
    return true;
  }
  std::vector<PlanCandidate> cands;
  if (plan_candidates(dig, dr, lr, rr, dd, ld, rd, kind, r4mode, acc, num_cus, cands) || cands.empty() ||
      g_sel->s.nin != kPlanFeatures)
    return false;
  float best = 0.f;
  int bi = -1;
  for (int i = 0; i < (int)cands.size(); ++i) {
    const float y = g_sel->predict_host(cands[i].f);
    if (bi < 0 || y < best) { best = y; bi = i; }
  }
  *variant = cands[bi].variant;
  *tile = cands[bi].tile;
  if (g_sel_memo.size() >= 65536) g_sel_memo.clear();
  g_sel_memo[key] = {*variant, *tile};
  return true;
}

}  // namespace xt
