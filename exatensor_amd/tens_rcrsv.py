"""Hierarchical block-sparse tensors and the contraction split that turns one
tensor contraction into the list of block (subtensor) contractions the
batched GPU path executes.

Follows ExaTENSOR's recursive tensors:
  * HSpace: a vector space of dimension N with its subspace aggregation tree
    (INTRAVIRT/subspaces.F90:2098-2372, HSpaceCtorSimple): subspaces 0..N-1
    are the single basis vectors, subspace N is the full space (the root),
    and the tree is grown level by level, left to right, every aggregate
    being cut into min(length, branching factor) contiguous pieces by the
    unaligned rule of SegIntSplit (:866-979): pieces of length len // m, the
    first len % m of them one longer; pieces longer than one get the next
    free id, single vectors keep theirs.
  * RcrsvTensor: a tensor whose every dimension is a subspace of an HSpace,
    composed of subtensors (one child subspace per dimension; by default
    the full Cartesian product of the children, dimension 0 fastest, and a
    caller may keep only some of them: block sparsity).
  * contraction_split: TensContractionSplitFunc
    (INTRAVIRT/tensor_recursive.F90:9202-9784) -- align the tree levels of
    dimensions that the pattern pairs (the finer side is promoted to its
    ancestor at the coarser level, align_levels :9350-9440, decided on each
    argument's first subtensor), describe every subtensor by its (promoted)
    subspace ids, join L and R on the contracted ids, form the Cartesian
    products, and keep those whose open ids name an existing D subtensor.
    Output order: D subtensors in ascending order of their id multi-index
    (dimension 1 most significant), and within one D subtensor ascending
    contracted multi-index, then R subtensor, then L subtensor -- the order
    the reference's sorted merge produces.
Every subtensor of an argument must sit at one tree level per dimension
(the reference decides the alignment on the first subtensor and assumes
this; here a mix raises ValueError). Symmetric-dimension restrictions
(get_dim_symmetry) are not supported: all dimensions are independent.

to_batch() makes engine tensors for the subtensors of D, L and R and a
talshAmdBatch of the split, i.e. the block-level loop of
INTERVIRT/tavp_worker.F90:9440-9480 as one grouped launch.
"""
import itertools

from . import talsh as T

CMP_EQ, CMP_LT, CMP_GT, CMP_OV = 0, -1, 1, 2


class HSpace:
    """Subspace aggregation tree of an N-dimensional space (see module doc)."""

    def __init__(self, dim, branch_fac=()):
        if dim <= 0:
            raise ValueError("space dimension must be positive")
        if any(b < 2 for b in branch_fac):
            raise ValueError("branching factors must be >= 2")
        self.dim = dim
        self.lo = list(range(dim))
        self.hi = [i + 1 for i in range(dim)]
        self.parent = [-1] * dim
        self.kids = [[] for _ in range(dim)]
        self.lev = [0] * dim
        if dim == 1:
            self.root = 0
            return
        self.root = self._new(0, dim, -1, 0)
        level, frontier = 0, [self.root]
        while frontier:
            brf = branch_fac[level] if level < len(branch_fac) else 2
            nxt = []
            for sid in frontier:
                length = self.hi[sid] - self.lo[sid]
                m = min(length, brf)
                if m <= 1:
                    continue
                base, extra = divmod(length, m)
                lb = self.lo[sid]
                for p in range(m):
                    ub = lb + base + (1 if p < extra else 0)
                    if ub - lb > 1:
                        child = self._new(lb, ub, sid, level + 1)
                        nxt.append(child)
                    else:
                        child = lb  # the single vector keeps its id
                        self.parent[child] = sid
                        self.lev[child] = level + 1
                    self.kids[sid].append(child)
                    lb = ub
            frontier = nxt
            level += 1

    def _new(self, lo, hi, parent, level):
        sid = len(self.lo)
        self.lo.append(lo)
        self.hi.append(hi)
        self.parent.append(parent)
        self.kids.append([])
        self.lev.append(level)
        return sid

    @property
    def num_subspaces(self):
        return len(self.lo)

    def extent(self, sid):
        return self.hi[sid] - self.lo[sid]

    def offset(self, sid):
        return self.lo[sid]

    def level(self, sid):
        return self.lev[sid]

    def children(self, sid):
        return list(self.kids[sid])

    def ancestor(self, sid, distance):
        for _ in range(distance):
            if self.parent[sid] < 0:
                raise ValueError("no ancestor at that distance")
            sid = self.parent[sid]
        return sid

    def level_composition(self, level):
        """Ids forming a level of the tree (a direct-sum decomposition once
        shallower leaves are carried down), in basis order."""
        out, stack = [], [self.root]
        while stack:
            sid = stack.pop()
            if self.lev[sid] == level or not self.kids[sid]:
                out.append(sid)
            else:
                stack.extend(reversed(self.kids[sid]))
        return out

    def compare_subranges(self, a, b):
        if self.lo[a] == self.lo[b] and self.hi[a] == self.hi[b]:
            return CMP_EQ
        if self.hi[a] <= self.lo[b]:
            return CMP_LT
        if self.hi[b] <= self.lo[a]:
            return CMP_GT
        return CMP_OV


class RcrsvTensor:
    """A tensor over subspaces of HSpaces, composed of subtensors."""

    def __init__(self, name, spaces, subspaces=None):
        self.name = name
        self.spaces = list(spaces)
        self.subspaces = list(subspaces) if subspaces is not None else [s.root for s in self.spaces]
        if len(self.subspaces) != len(self.spaces):
            raise ValueError("one subspace per dimension")
        self.subtensors = None  # list of subspace-id tuples; None: not split yet

    @property
    def rank(self):
        return len(self.spaces)

    def dims(self, ids=None):
        ids = self.subspaces if ids is None else ids
        return [sp.extent(i) for sp, i in zip(self.spaces, ids)]

    def split(self, keep=None):
        """Composes the tensor of the children of its subspaces (a dimension
        with no children keeps its subspace), dimension 0 fastest; keep(ids)
        may drop subtensors. Returns the subtensor list."""
        per = [sp.children(i) or [i] for sp, i in zip(self.spaces, self.subspaces)]
        subs = [tuple(reversed(c)) for c in itertools.product(*reversed(per))]
        self.subtensors = [s for s in subs if keep is None or keep(s)]
        return self.subtensors

    def subtensor(self, ids):
        return RcrsvTensor(self.name, self.spaces, ids)


def _pattern_map(pattern, drank, lrank, rrank):
    e, dig, dr, lr, rr, cj = T.str2dig(T.load_library(), pattern)
    if e != 0:
        raise ValueError(f"bad contraction pattern {pattern!r}: {e}")
    if (dr, lr, rr) != (drank, lrank, rrank):
        raise ValueError(f"pattern ranks {dr},{lr},{rr} do not match the tensors")
    if cj:
        raise ValueError("conjugated arguments are not supported by the split")
    return dig


def contraction_split(pattern, dten, lten, rten):
    """Block contractions (d, l, r) -- indices into dten/lten/rten.subtensors
    -- of D += L * R (see module doc for the rule and the order)."""
    for t in (dten, lten, rten):
        if not t.subtensors:
            raise ValueError(f"tensor {t.name} has no subtensors (call split())")
    dr, lr, rr = dten.rank, lten.rank, rten.rank
    cp = _pattern_map(pattern, dr, lr, rr)
    # the alignment is decided on the first subtensor, so every subtensor of
    # an argument must sit at one tree level per dimension
    for t in (dten, lten, rten):
        for i, sp in enumerate(t.spaces):
            if len({sp.level(s[i]) for s in t.subtensors}) > 1:
                raise ValueError(f"subtensors of {t.name} mix tree levels in dimension {i}")
    # level alignment, decided on the first subtensor of each argument
    adj = [[0] * dr, [0] * lr, [0] * rr]
    first = (dten.subtensors[0], lten.subtensors[0], rten.subtensors[0])
    tens = (dten, lten, rten)

    def align(a, i, b, j):
        la = tens[a].spaces[i].level(first[a][i])
        lb = tens[b].spaces[j].level(first[b][j])
        if la < lb:
            adj[b][j] = lb - la
        elif la > lb:
            adj[a][i] = la - lb

    for i in range(lr):
        j = cp[i]
        if j < 0:
            align(1, i, 2, -j - 1)
        else:
            align(1, i, 0, j - 1)
    for i in range(rr):
        j = cp[lr + i]
        if j > 0:
            align(2, i, 0, j - 1)

    def descr(a, ids):
        return tuple(tens[a].spaces[d].ancestor(s, adj[a][d]) if adj[a][d] else s for d, s in enumerate(ids))

    # dimension orders: contracted dims (L order) then open dims
    lc = [i for i in range(lr) if cp[i] < 0]
    rc = [-cp[i] - 1 for i in lc]
    lo = [i for i in range(lr) if cp[i] > 0]
    ro = [i for i in range(rr) if cp[lr + i] > 0]
    # D position of each open dim of L, then of R
    dpos = [cp[i] - 1 for i in lo] + [cp[lr + i] - 1 for i in ro]
    ldesc = [descr(1, s) for s in lten.subtensors]
    rdesc = [descr(2, s) for s in rten.subtensors]
    ddesc = [descr(0, s) for s in dten.subtensors]
    lgroups, rgroups = {}, {}
    for n, d in enumerate(ldesc):
        lgroups.setdefault(tuple(d[i] for i in lc), []).append(n)
    for n, d in enumerate(rdesc):
        rgroups.setdefault(tuple(d[i] for i in rc), []).append(n)
    products = {}  # D multi-index -> [(contracted key, r, l)]
    for key in sorted(set(lgroups) & set(rgroups)):
        for r in sorted(rgroups[key], key=lambda n: (tuple(rdesc[n][i] for i in ro), n)):
            for l in sorted(lgroups[key], key=lambda n: (tuple(ldesc[n][i] for i in lo), n)):
                open_ids = [ldesc[l][i] for i in lo] + [rdesc[r][i] for i in ro]
                dkey = [0] * dr
                for p, v in zip(dpos, open_ids):
                    dkey[p] = v
                products.setdefault(tuple(dkey), []).append((key, r, l))
    out = []
    for d in sorted(range(len(ddesc)), key=lambda n: (ddesc[n], n)):
        for _key, r, l in products.get(ddesc[d], ()):
            out.append((d, l, r))
    return out


def to_batch(lib, pattern, dten, lten, rten, split, gpu, dbody, lbody, rbody, alpha=1.0, accumulate=True):
    """Engine batch (talshAmdBatchCreate on GPU `gpu`) of the block
    contractions `split`; xbody(ids) -> T.Tensor returns the engine tensor
    (GPU-resident) of a subtensor of the respective argument. Returns the
    batch handle."""
    import ctypes as C
    cp = _pattern_map(pattern, dten.rank, lten.rank, rten.rank)
    lr = lten.rank
    for d, l, r in split:
        dd, ld, rd = dten.dims(dten.subtensors[d]), lten.dims(lten.subtensors[l]), rten.dims(rten.subtensors[r])
        ok = all(ld[i] == (rd[-cp[i] - 1] if cp[i] < 0 else dd[cp[i] - 1]) for i in range(lr))
        ok = ok and all(cp[lr + i] < 0 or rd[i] == dd[cp[lr + i] - 1] for i in range(rten.rank))
        if not ok:
            raise ValueError("block contraction with subtensors of different tree levels: whole-block batches "
                             "need D, L and R split to matching levels")
    n = len(split)
    arr = [(T._TP * n)() for _ in range(3)]
    for k, (d, l, r) in enumerate(split):
        arr[0][k] = C.pointer(dbody(dten.subtensors[d]).t)
        arr[1][k] = C.pointer(lbody(lten.subtensors[l]).t)
        arr[2][k] = C.pointer(rbody(rten.subtensors[r]).t)
    h = C.c_void_p()
    e = lib.talshAmdBatchCreate(pattern.encode(), n, arr[0], arr[1], arr[2], alpha, T.YEP if accumulate else T.NOPE,
                                gpu, C.byref(h))
    if e != 0:
        raise RuntimeError(f"talshAmdBatchCreate: {e}")
    return h
