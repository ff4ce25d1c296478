// Hierarchical block-sparse tensors: subspace aggregation trees and the
// contraction split that turns one contraction of recursive tensors into the
// list of block (subtensor) contractions the batched GPU path runs.
//
// Follows ExaTENSOR's recursive tensors:
//  * a space of dimension N with its subspace tree (HSpaceCtorSimple,
//    INTRAVIRT/subspaces.F90:2098-2372): subspaces 0..N-1 are the single
//    basis vectors, N is the full space (the root); the tree grows level by
//    level, left to right, every aggregate cut into min(length, branching
//    factor) contiguous pieces by SegIntSplit's unaligned rule (:866-979):
//    pieces of length len / m, the first len % m of them one longer; pieces
//    longer than one get the next free id, single vectors keep theirs;
//  * the contraction split (TensContractionSplitFunc,
//    INTRAVIRT/tensor_recursive.F90:9202-9784): the tree levels of paired
//    dimensions are aligned (the finer side is promoted to its ancestor at
//    the coarser level, decided on each argument's first subtensor), every
//    subtensor is described by its (promoted) subspace ids, L and R are
//    joined on the contracted ids, and the products whose open ids name an
//    existing D subtensor are kept. Output: D subtensors in ascending order
//    of their id multi-index (dimension 0 most significant), within one D
//    subtensor ascending contracted ids, then R subtensor (open ids, index),
//    then L subtensor (open ids, index).
// Every subtensor of an argument must sit at one tree level per dimension
// (the reference decides the alignment on the first subtensor and assumes
// it); a mix is rejected. Symmetric dimensions are not supported.
#include <algorithm>
#include <map>
#include <vector>

#include "talsh.h"
#include "talsh_amd.h"

namespace xt {
namespace {

struct HSpace {
  int dim = 0, root = 0;
  std::vector<int> lo, hi, parent, lev;
  std::vector<std::vector<int>> kids;

  int add(int l, int h, int p, int level)
  {
    lo.push_back(l);
    hi.push_back(h);
    parent.push_back(p);
    lev.push_back(level);
    kids.emplace_back();
    return (int)lo.size() - 1;
  }
  int size() const { return (int)lo.size(); }
  int ancestor(int sid, int distance) const
  {
    for (int i = 0; i < distance && sid >= 0; ++i) sid = parent[sid];
    return sid;
  }
};

int build(int dim, int nlevels, const int *branch_fac, HSpace &h)
{
  if (dim <= 0 || nlevels < 0 || (nlevels > 0 && !branch_fac)) return TALSH_INVALID_ARGS;
  for (int i = 0; i < nlevels; ++i)
    if (branch_fac[i] < 2) return TALSH_INVALID_ARGS;
  h.dim = dim;
  for (int i = 0; i < dim; ++i) h.add(i, i + 1, -1, 0);
  if (dim == 1) {
    h.root = 0;
    return 0;
  }
  h.root = h.add(0, dim, -1, 0);
  std::vector<int> frontier{h.root};
  for (int level = 0; !frontier.empty(); ++level) {
    const int brf = level < nlevels ? branch_fac[level] : 2;
    std::vector<int> next;
    for (int sid : frontier) {
      const int length = h.hi[sid] - h.lo[sid];
      const int m = std::min(length, brf);
      if (m <= 1) continue;
      const int base = length / m, extra = length % m;
      int lb = h.lo[sid];
      for (int p = 0; p < m; ++p) {
        const int ub = lb + base + (p < extra ? 1 : 0);
        int child;
        if (ub - lb > 1) {
          child = h.add(lb, ub, sid, level + 1);
          next.push_back(child);
        } else {
          child = lb;  // the single vector keeps its id
          h.parent[child] = sid;
          h.lev[child] = level + 1;
        }
        h.kids[sid].push_back(child);
        lb = ub;
      }
    }
    frontier.swap(next);
  }
  return 0;
}

struct Arg {
  int rank, n;
  const HSpace *const *sp;  // per dimension
  const int *sub;           // n x rank subspace ids (row per subtensor)
  int id(int t, int d) const { return sub[(size_t)t * rank + d]; }
};

}  // namespace
}  // namespace xt

using namespace xt;

extern "C" {

int talshAmdHSpaceCreate(int dim, int nlevels, const int *branch_fac, void **hspace)
{
  if (!hspace) return TALSH_INVALID_ARGS;
  *hspace = nullptr;
  HSpace *h = new HSpace();
  const int e = build(dim, nlevels, branch_fac, *h);
  if (e) {
    delete h;
    return e;
  }
  *hspace = h;
  return 0;
}

int talshAmdHSpaceDestroy(void *hspace)
{
  if (!hspace) return TALSH_INVALID_ARGS;
  delete static_cast<HSpace *>(hspace);
  return 0;
}

int talshAmdHSpaceInfo(void *hspace, int *num_subspaces, int *root)
{
  if (!hspace || !num_subspaces || !root) return TALSH_INVALID_ARGS;
  const HSpace *h = static_cast<const HSpace *>(hspace);
  *num_subspaces = h->size();
  *root = h->root;
  return 0;
}

int talshAmdHSpaceSubspace(void *hspace, int sid, int *offset, int *extent, int *level, int *parent, int *nchildren)
{
  if (!hspace) return TALSH_INVALID_ARGS;
  const HSpace *h = static_cast<const HSpace *>(hspace);
  if (sid < 0 || sid >= h->size()) return TALSH_INVALID_ARGS;
  if (offset) *offset = h->lo[sid];
  if (extent) *extent = h->hi[sid] - h->lo[sid];
  if (level) *level = h->lev[sid];
  if (parent) *parent = h->parent[sid];
  if (nchildren) *nchildren = (int)h->kids[sid].size();
  return 0;
}

int talshAmdHSpaceChildren(void *hspace, int sid, int cap, int *children)
{
  if (!hspace) return TALSH_INVALID_ARGS;
  const HSpace *h = static_cast<const HSpace *>(hspace);
  if (sid < 0 || sid >= h->size() || (cap > 0 && !children)) return TALSH_INVALID_ARGS;
  const std::vector<int> &k = h->kids[sid];
  if ((int)k.size() > cap) return TALSH_LIMIT_EXCEEDED;
  std::copy(k.begin(), k.end(), children);
  return 0;
}

int talshAmdContractionSplit(const char *cptrn, int drank, void *const *dspaces, int nd, const int *dsub, int lrank,
                             void *const *lspaces, int nl, const int *lsub, int rrank, void *const *rspaces, int nr,
                             const int *rsub, long long cap, int *triples, long long *ntriples)
{
  if (!cptrn || !ntriples || nd < 0 || nl < 0 || nr < 0 || (cap > 0 && !triples)) return TALSH_INVALID_ARGS;
  *ntriples = 0;
  int dig[2 * MAX_TENSOR_RANK], dr, lr, rr, cj;
  if (talsh_get_contr_ptrn_str2dig(cptrn, dig, &dr, &lr, &rr, &cj)) return TALSH_INVALID_ARGS;
  if (dr != drank || lr != lrank || rr != rrank || rr < 0 || cj != 0) return TALSH_INVALID_ARGS;
  if (nd == 0 || nl == 0 || nr == 0) return 0;
  Arg A[3] = {{dr, nd, reinterpret_cast<const HSpace *const *>(dspaces), dsub},
              {lr, nl, reinterpret_cast<const HSpace *const *>(lspaces), lsub},
              {rr, nr, reinterpret_cast<const HSpace *const *>(rspaces), rsub}};
  // valid ids, one tree level per dimension within every argument
  for (const Arg &a : A) {
    if (a.rank > 0 && (!a.sp || !a.sub)) return TALSH_INVALID_ARGS;
    for (int d = 0; d < a.rank; ++d) {
      if (!a.sp[d]) return TALSH_INVALID_ARGS;
      const HSpace &h = *a.sp[d];
      for (int t = 0; t < a.n; ++t) {
        const int s = a.id(t, d);
        if (s < 0 || s >= h.size()) return TALSH_INVALID_ARGS;
        if (h.lev[s] != h.lev[a.id(0, d)]) return TALSH_INVALID_ARGS;
      }
    }
  }
  // level alignment of paired dimensions, decided on the first subtensors
  std::vector<int> adj[3] = {std::vector<int>(dr, 0), std::vector<int>(lr, 0), std::vector<int>(rr, 0)};
  auto align = [&](int a, int i, int b, int j) {
    const int la = A[a].sp[i]->lev[A[a].id(0, i)], lb = A[b].sp[j]->lev[A[b].id(0, j)];
    if (la < lb) adj[b][j] = lb - la;
    else if (la > lb) adj[a][i] = la - lb;
  };
  for (int i = 0; i < lr; ++i) {
    const int j = dig[i];
    if (j < 0) align(1, i, 2, -j - 1);
    else align(1, i, 0, j - 1);
  }
  for (int i = 0; i < rr; ++i)
    if (dig[lr + i] > 0) align(2, i, 0, dig[lr + i] - 1);
  auto descr = [&](int a, int t) {
    std::vector<int> v(A[a].rank);
    for (int d = 0; d < A[a].rank; ++d) v[d] = A[a].sp[d]->ancestor(A[a].id(t, d), adj[a][d]);
    return v;
  };
  std::vector<std::vector<int>> desc[3];
  for (int a = 0; a < 3; ++a)
    for (int t = 0; t < A[a].n; ++t) desc[a].push_back(descr(a, t));
  std::vector<int> lc, rc, lo, ro, dpos;  // contracted dims (L order) and open dims
  for (int i = 0; i < lr; ++i)
    if (dig[i] < 0) { lc.push_back(i); rc.push_back(-dig[i] - 1); }
  for (int i = 0; i < lr; ++i)
    if (dig[i] > 0) { lo.push_back(i); dpos.push_back(dig[i] - 1); }
  for (int i = 0; i < rr; ++i)
    if (dig[lr + i] > 0) { ro.push_back(i); dpos.push_back(dig[lr + i] - 1); }
  auto pick = [](const std::vector<int> &v, const std::vector<int> &idx) {
    std::vector<int> o(idx.size());
    for (size_t q = 0; q < idx.size(); ++q) o[q] = v[idx[q]];
    return o;
  };
  std::map<std::vector<int>, std::vector<int>> lgroups, rgroups;
  for (int t = 0; t < nl; ++t) lgroups[pick(desc[1][t], lc)].push_back(t);
  for (int t = 0; t < nr; ++t) rgroups[pick(desc[2][t], rc)].push_back(t);
  auto by_open = [&](std::vector<int> &ts, int a, const std::vector<int> &open) {
    std::sort(ts.begin(), ts.end(), [&](int x, int y) {
      const std::vector<int> ox = pick(desc[a][x], open), oy = pick(desc[a][y], open);
      return ox != oy ? ox < oy : x < y;
    });
  };
  // D multi-index -> (r, l) pairs in (contracted ids, R, L) order
  std::map<std::vector<int>, std::vector<std::pair<int, int>>> products;
  for (auto &kv : lgroups) {  // std::map iterates keys in ascending order
    auto rit = rgroups.find(kv.first);
    if (rit == rgroups.end()) continue;
    std::vector<int> rs = rit->second, ls = kv.second;
    by_open(rs, 2, ro);
    by_open(ls, 1, lo);
    for (int r : rs)
      for (int l : ls) {
        std::vector<int> dkey(dr, 0);
        size_t q = 0;
        for (int i : lo) dkey[dpos[q++]] = desc[1][l][i];
        for (int i : ro) dkey[dpos[q++]] = desc[2][r][i];
        products[dkey].emplace_back(r, l);
      }
  }
  std::vector<int> dorder(nd);
  for (int t = 0; t < nd; ++t) dorder[t] = t;
  std::sort(dorder.begin(), dorder.end(), [&](int x, int y) {
    return desc[0][x] != desc[0][y] ? desc[0][x] < desc[0][y] : x < y;
  });
  long long n = 0;
  for (int d : dorder) {
    auto it = products.find(desc[0][d]);
    if (it == products.end()) continue;
    for (const auto &rl : it->second) {
      if (n < cap) {
        triples[3 * n] = d;
        triples[3 * n + 1] = rl.second;
        triples[3 * n + 2] = rl.first;
      }
      ++n;
    }
  }
  *ntriples = n;
  return n > cap ? TALSH_LIMIT_EXCEEDED : 0;
}

int talshAmdBatchCreateFromSplit(const char *cptrn, long long ntriples, const int *triples, talsh_tens_t *const *dblocks,
                                 int nd, talsh_tens_t *const *lblocks, int nl, talsh_tens_t *const *rblocks, int nr,
                                 double alpha_re, int accumulative, int dev_num, void **batch)
{
  if (!triples || !dblocks || !lblocks || !rblocks || ntriples <= 0 || ntriples > (1ll << 30)) return TALSH_INVALID_ARGS;
  std::vector<talsh_tens_t *> D(ntriples), L(ntriples), R(ntriples);
  for (long long i = 0; i < ntriples; ++i) {
    const int d = triples[3 * i], l = triples[3 * i + 1], r = triples[3 * i + 2];
    if (d < 0 || d >= nd || l < 0 || l >= nl || r < 0 || r >= nr) return TALSH_INVALID_ARGS;
    D[i] = dblocks[d];
    L[i] = lblocks[l];
    R[i] = rblocks[r];
  }
  return talshAmdBatchCreate(cptrn, (int)ntriples, D.data(), L.data(), R.data(), alpha_re, accumulative, dev_num,
                             batch);
}

}  // extern "C"
