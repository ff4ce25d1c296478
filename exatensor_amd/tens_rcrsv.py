"""Hierarchical block-sparse tensors and the contraction split that turns one
tensor contraction into the list of block (subtensor) contractions the
batched GPU path executes: a Python view of the engine's C-ABI
(talshAmdHSpace*, talshAmdContractionSplit, talshAmdBatchCreateFromSplit in
include/talsh_amd.h, implemented in exatensor_amd/csrc/hsplit.cpp).

Follows ExaTENSOR's recursive tensors:
  * HSpace: a vector space of dimension N with its subspace aggregation tree
    (INTRAVIRT/subspaces.F90:2098-2372, HSpaceCtorSimple): subspaces 0..N-1
    are the single basis vectors, subspace N is the full space (the root),
    and the tree is grown level by level, left to right, every aggregate
    being cut into min(length, branching factor) contiguous pieces by the
    unaligned rule of SegIntSplit (:866-979): pieces of length len // m, the
    first len % m of them one longer; pieces longer than one get the next
    free id, single vectors keep theirs. The tree is built by the engine;
    this class mirrors its ranges, parents, levels and children.
  * RcrsvTensor: a tensor whose every dimension is a subspace of an HSpace,
    composed of subtensors (one child subspace per dimension; by default
    the full Cartesian product of the children, dimension 0 fastest, and a
    caller may keep only some of them: block sparsity).
  * contraction_split: TensContractionSplitFunc
    (INTRAVIRT/tensor_recursive.F90:9202-9784), computed by the engine --
    align the tree levels of dimensions that the pattern pairs (the finer
    side is promoted to its ancestor at the coarser level, align_levels
    :9350-9440, decided on each argument's first subtensor), describe every
    subtensor by its (promoted) subspace ids, join L and R on the contracted
    ids, form the Cartesian products, and keep those whose open ids name an
    existing D subtensor. Output order: D subtensors in ascending order of
    their id multi-index (dimension 1 most significant), and within one D
    subtensor ascending contracted multi-index, then R subtensor, then L
    subtensor -- the order the reference's sorted merge produces.
Every subtensor of an argument must sit at one tree level per dimension
(the reference decides the alignment on the first subtensor and assumes
this; here a mix raises ValueError). Symmetric-dimension restrictions
(get_dim_symmetry) are not supported: all dimensions are independent.

to_batch() makes a talshAmdBatch of the split from engine tensors of the
subtensors of D, L and R, i.e. the block-level loop of
INTERVIRT/tavp_worker.F90:9440-9480 as one grouped launch.
"""
import ctypes as C
import itertools

from . import talsh as T

CMP_EQ, CMP_LT, CMP_GT, CMP_OV = 0, -1, 1, 2

_IP = C.POINTER(C.c_int)
_VPP = C.POINTER(C.c_void_p)


def _lib():
    L = T.load_library()
    if not getattr(L, "_hsplit_bound", False):
        sig = {
            "talshAmdHSpaceCreate": (C.c_int, [C.c_int, C.c_int, _IP, C.POINTER(C.c_void_p)]),
            "talshAmdHSpaceDestroy": (C.c_int, [C.c_void_p]),
            "talshAmdHSpaceInfo": (C.c_int, [C.c_void_p, _IP, _IP]),
            "talshAmdHSpaceSubspace": (C.c_int, [C.c_void_p, C.c_int, _IP, _IP, _IP, _IP, _IP]),
            "talshAmdHSpaceChildren": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _IP]),
            "talshAmdContractionSplit": (C.c_int, [C.c_char_p, C.c_int, _VPP, C.c_int, _IP, C.c_int, _VPP, C.c_int,
                                                   _IP, C.c_int, _VPP, C.c_int, _IP, C.c_longlong, _IP,
                                                   C.POINTER(C.c_longlong)]),
            "talshAmdBatchCreateFromSplit": (C.c_int, [C.c_char_p, C.c_longlong, _IP, C.POINTER(T._TP), C.c_int,
                                                       C.POINTER(T._TP), C.c_int, C.POINTER(T._TP), C.c_int,
                                                       C.c_double, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._hsplit_bound = True
    return L


class HSpace:
    """Subspace aggregation tree of an N-dimensional space (see module doc),
    built by the engine (talshAmdHSpaceCreate)."""

    def __init__(self, dim, branch_fac=()):
        if dim <= 0:
            raise ValueError("space dimension must be positive")
        if any(b < 2 for b in branch_fac):
            raise ValueError("branching factors must be >= 2")
        L = _lib()
        self._L = L
        self.handle = C.c_void_p()
        bf = T.ints(branch_fac)
        e = L.talshAmdHSpaceCreate(dim, len(branch_fac), bf, C.byref(self.handle))
        if e:
            raise ValueError(f"talshAmdHSpaceCreate: {e}")
        n, root = C.c_int(), C.c_int()
        L.talshAmdHSpaceInfo(self.handle, C.byref(n), C.byref(root))
        self.dim = dim
        self.root = root.value
        self.lo, self.hi, self.parent, self.lev, self.kids = [], [], [], [], []
        off, ext, lev, par, nk = (C.c_int() for _ in range(5))
        for sid in range(n.value):
            L.talshAmdHSpaceSubspace(self.handle, sid, C.byref(off), C.byref(ext), C.byref(lev), C.byref(par),
                                     C.byref(nk))
            kids = (C.c_int * max(nk.value, 1))()
            L.talshAmdHSpaceChildren(self.handle, sid, nk.value, kids)
            self.lo.append(off.value)
            self.hi.append(off.value + ext.value)
            self.lev.append(lev.value)
            self.parent.append(par.value)
            self.kids.append(list(kids[:nk.value]))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and h.value:
            self._L.talshAmdHSpaceDestroy(h)
            self.handle = C.c_void_p()

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


def _arg(t):
    spaces = (C.c_void_p * max(t.rank, 1))(*[sp.handle.value for sp in t.spaces])
    flat = [i for s in t.subtensors for i in s]
    return spaces, len(t.subtensors), T.ints(flat)


def contraction_split(pattern, dten, lten, rten):
    """Block contractions (d, l, r) -- indices into dten/lten/rten.subtensors
    -- of D += L * R (see module doc for the rule and the order), computed by
    talshAmdContractionSplit."""
    for t in (dten, lten, rten):
        if not t.subtensors:
            raise ValueError(f"tensor {t.name} has no subtensors (call split())")
    _pattern_map(pattern, dten.rank, lten.rank, rten.rank)
    for t in (dten, lten, rten):
        for i, sp in enumerate(t.spaces):
            if len({sp.level(s[i]) for s in t.subtensors}) > 1:
                raise ValueError(f"subtensors of {t.name} mix tree levels in dimension {i}")
    L = _lib()
    args = []
    for t in (dten, lten, rten):
        args.append((t.rank,) + _arg(t))
    flat = []
    for rank, spaces, n, sub in args:
        flat += [rank, spaces, n, sub]
    n = C.c_longlong()
    cap = 1 << 12
    while True:
        out = (C.c_int * (3 * cap))()
        e = L.talshAmdContractionSplit(pattern.encode(), *flat, cap, out, C.byref(n))
        if e == T.TALSH_LIMIT_EXCEEDED and n.value > cap:
            cap = n.value
            continue
        if e:
            raise ValueError(f"talshAmdContractionSplit: {e}")
        return [(out[3 * i], out[3 * i + 1], out[3 * i + 2]) for i in range(n.value)]


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
    L = _lib()
    blocks = []
    for t, body in ((dten, dbody), (lten, lbody), (rten, rbody)):
        arr = (T._TP * len(t.subtensors))()
        for k, ids in enumerate(t.subtensors):
            arr[k] = C.pointer(body(ids).t)
        blocks.append(arr)
    n = len(split)
    tri = T.ints([x for dlr in split for x in dlr])
    h = C.c_void_p()
    e = L.talshAmdBatchCreateFromSplit(pattern.encode(), n, tri, blocks[0], len(dten.subtensors), blocks[1],
                                       len(lten.subtensors), blocks[2], len(rten.subtensors), alpha,
                                       T.YEP if accumulate else T.NOPE, gpu, C.byref(h))
    if e != 0:
        raise RuntimeError(f"talshAmdBatchCreateFromSplit: {e}")
    return h
