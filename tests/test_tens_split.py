"""Hierarchical spaces and the contraction split (exatensor_amd/tens_rcrsv.py).

Pinned by: the subspace-tree rule of HSpaceCtorSimple / SegIntSplit on a
hand-derived case; range geometry (children tile their parent, ids, levels);
a brute-force enumeration of block triples whose paired dimensions have
nested ranges (tree subspaces either nest or are disjoint, which is what
the level alignment + id equality of TensContractionSplitFunc decides); the
documented output order; and the sum of the block contractions against one
dense contraction (numpy, float64), dense and block-sparse, on the CPU and
through the engine's batched GPU path."""
import random

import numpy as np
import pytest

from exatensor_amd import tens_rcrsv as TR

PAT = "D(a,b,c,d)+=L(a,k,c,l)*R(l,b,k,d)"


def test_hspace_known_tree():
    h = TR.HSpace(10)
    # singletons 0..9, root 10 = [0,10); level by level, first pieces longer
    assert h.root == 10 and h.num_subspaces == 19
    assert [(h.lo[c], h.hi[c]) for c in h.children(10)] == [(0, 5), (5, 10)]
    assert h.children(10) == [11, 12]
    assert [(h.lo[c], h.hi[c]) for c in h.children(11)] == [(0, 3), (3, 5)]
    assert h.children(11) == [13, 14] and h.children(12) == [15, 16]
    assert h.children(13) == [17, 2] and h.children(14) == [3, 4]
    assert h.children(17) == [0, 1]
    assert [h.level(i) for i in range(10)] == [4, 4, 3, 3, 3, 4, 4, 3, 3, 3]
    assert h.level_composition(2) == [13, 14, 15, 16]
    h3 = TR.HSpace(20, (3, 2))
    assert [h3.extent(c) for c in h3.children(h3.root)] == [7, 7, 6]
    assert [h3.extent(c) for c in h3.children(h3.children(h3.root)[0])] == [4, 3]


@pytest.mark.parametrize("n,bf", [(1, ()), (2, ()), (7, ()), (33, ()), (48, (4,)), (100, (3, 5, 2))])
def test_hspace_geometry(n, bf):
    h = TR.HSpace(n, bf)
    for i in range(n):
        assert (h.lo[i], h.hi[i]) == (i, i + 1) and not h.children(i)
    assert (h.lo[h.root], h.hi[h.root]) == (0, n)
    for s in range(h.num_subspaces):
        kids = h.children(s)
        if kids:
            assert h.lo[kids[0]] == h.lo[s] and h.hi[kids[-1]] == h.hi[s]
            assert all(h.hi[a] == h.lo[b] for a, b in zip(kids, kids[1:]))
            assert all(h.level(k) == h.level(s) + 1 and h.ancestor(k, 1) == s for k in kids)
        if s != h.root:
            assert h.ancestor(s, h.level(s)) == h.root
    for lev in range(4):
        comp = h.level_composition(lev)
        assert h.lo[comp[0]] == 0 and h.hi[comp[-1]] == n
        assert all(h.hi[a] == h.lo[b] for a, b in zip(comp, comp[1:]))


def _tensors(seed, depth=(1, 1, 1), fill=1.0):
    """D, L, R over shared spaces a,b,c,d,k,l, split to the given depths."""
    rng = random.Random(seed)
    # extents whose trees keep one level per depth (up to 2)
    sp = {m: TR.HSpace(*rng.choice(((4, ()), (8, ()), (6, (3,)), (12, (3,)), (16, (4,))))) for m in "abcdkl"}

    def make(name, modes, lev):
        t = TR.RcrsvTensor(name, [sp[m] for m in modes])
        subs = [tuple(t.subspaces)]
        for _ in range(lev):  # split every subtensor one level further
            nxt = []
            for s in subs:
                c = t.subtensor(s)
                nxt += c.split()
            subs = nxt
        t.subtensors = [s for s in subs if rng.random() < fill] or subs[:1]
        return t

    return sp, make("D", "abcd", depth[0]), make("L", "akcl", depth[1]), make("R", "lbkd", depth[2])


def _nested(h, a, b):
    return not (h.hi[a] <= h.lo[b] or h.hi[b] <= h.lo[a])


def _brute(sp, D, L, R):
    """Every (d, l, r) whose paired dimensions have overlapping ranges."""
    def compat(t1, t2, pairs):
        m = np.ones((len(t1.subtensors), len(t2.subtensors)), bool)
        for mode, i, j in pairs:
            h = sp[mode]
            lo1 = np.array([h.lo[s[i]] for s in t1.subtensors])[:, None]
            hi1 = np.array([h.hi[s[i]] for s in t1.subtensors])[:, None]
            lo2 = np.array([h.lo[s[j]] for s in t2.subtensors])[None, :]
            hi2 = np.array([h.hi[s[j]] for s in t2.subtensors])[None, :]
            m &= (lo1 < hi2) & (lo2 < hi1)
        return m

    dl = compat(D, L, [("a", 0, 0), ("c", 2, 2)])
    dr = compat(D, R, [("b", 1, 1), ("d", 3, 3)])
    lr = compat(L, R, [("k", 1, 2), ("l", 3, 0)])
    out = set()
    for d in range(len(D.subtensors)):
        ls, rs = np.nonzero(dl[d])[0], np.nonzero(dr[d])[0]
        for i, j in zip(*np.nonzero(lr[np.ix_(ls, rs)])):
            out.add((d, int(ls[i]), int(rs[j])))
    return out


@pytest.mark.parametrize("seed,depth,fill", [(1, (1, 1, 1), 1.0), (2, (1, 1, 1), 0.5), (3, (2, 1, 2), 0.6),
                                             (4, (1, 2, 2), 0.4), (5, (2, 2, 2), 0.3), (6, (0, 1, 1), 1.0)])
def test_split_matches_brute_force_and_order(seed, depth, fill):
    sp, D, L, R = _tensors(seed, depth, fill)
    got = TR.contraction_split(PAT, D, L, R)
    assert len(got) == len(set(got))
    assert set(got) == _brute(sp, D, L, R)
    # every D subtensor's block contractions are contiguous; with equal
    # levels the D subtensors come in ascending order of their ids
    seen, last = set(), None
    for d, _, _ in got:
        if d != last:
            assert d not in seen
            seen.add(d)
            last = d
    if len(set(depth)) == 1:
        dkeys = [D.subtensors[d] for d, _, _ in got]
        assert dkeys == sorted(dkeys)


def _dense(sp, seed):
    rng = np.random.default_rng(seed)
    la = rng.uniform(-1, 1, [sp[m].dim for m in "akcl"])
    ra = rng.uniform(-1, 1, [sp[m].dim for m in "lbkd"])
    return la, ra


def _box(t, ids):
    return tuple(slice(sp.lo[i], sp.hi[i]) for sp, i in zip(t.spaces, ids))


def _boxes(sp, D, L, R, d, l, r):
    """Index ranges a block contraction covers: for every index, the
    intersection of the ranges of the subtensors that carry it (the finer
    one when levels differ)."""
    ds, ls, rs = D.subtensors[d], L.subtensors[l], R.subtensors[r]
    rng = {}
    for m, a, b in (("a", ds[0], ls[0]), ("b", ds[1], rs[1]), ("c", ds[2], ls[2]), ("d", ds[3], rs[3]),
                    ("k", ls[1], rs[2]), ("l", ls[3], rs[0])):
        h = sp[m]
        rng[m] = slice(max(h.lo[a], h.lo[b]), min(h.hi[a], h.hi[b]))
    return tuple(rng[m] for m in "abcd"), tuple(rng[m] for m in "akcl"), tuple(rng[m] for m in "lbkd")


@pytest.mark.parametrize("seed,depth,fill", [(11, (1, 1, 1), 1.0), (12, (2, 1, 2), 1.0), (13, (1, 2, 1), 0.5),
                                             (14, (2, 2, 2), 0.4), (15, (0, 2, 1), 1.0)])
def test_block_sum_equals_dense_contraction(seed, depth, fill):
    sp, D, L, R = _tensors(seed, depth, fill)
    la, ra = _dense(sp, seed)
    # zero the blocks the block-sparse L and R do not hold
    lm, rm = np.zeros_like(la), np.zeros_like(ra)
    for s in L.subtensors:
        lm[_box(L, s)] = la[_box(L, s)]
    for s in R.subtensors:
        rm[_box(R, s)] = ra[_box(R, s)]
    want = np.einsum("akcl,lbkd->abcd", lm, rm)
    got = np.zeros_like(want)
    for d, l, r in TR.contraction_split(PAT, D, L, R):
        db, lb, rb = _boxes(sp, D, L, R, d, l, r)
        got[db] += np.einsum("akcl,lbkd->abcd", lm[lb], rm[rb])
    held = np.zeros(want.shape, bool)
    for s in D.subtensors:
        held[_box(D, s)] = True
    assert np.allclose(got[held], want[held], rtol=1e-13, atol=1e-13)
    assert not got[~held].any()


def test_bad_pattern_rank():
    sp, D, L, R = _tensors(1)
    with pytest.raises(ValueError):
        TR.contraction_split("D(a,b,c)+=L(a,k,c,l)*R(l,b,k,d)", D, L, R)


def test_mixed_levels_rejected():
    h = TR.HSpace(3)  # [0,3) -> [0,2) (splits again), 2 (a single vector)
    t = TR.RcrsvTensor("L", [h, h])
    t.subtensors = [(h.children(h.root)[0], h.root), (2, h.root), (0, h.root)]
    u = TR.RcrsvTensor("R", [h, h]); u.split()
    d = TR.RcrsvTensor("D", [h, h]); d.split()
    with pytest.raises(ValueError):
        TR.contraction_split("D(a,b)+=L(a,k)*R(k,b)", d, t, u)


@pytest.mark.gpu
@pytest.mark.parametrize("seed,depth,fill", [(21, (1, 1, 1), 1.0), (22, (2, 2, 2), 0.5)])
def test_split_batch_on_gpu(seed, depth, fill):
    from exatensor_amd import talsh as T
    from gpu_util import engine, gpu_tensor

    eng = engine()
    lib = eng.lib
    sp, D, L, R = _tensors(seed, depth, fill)
    la, ra = _dense(sp, seed)
    bodies = {}

    def body(tag, t, arr):
        def get(ids):
            key = (tag, ids)
            if key not in bodies:
                blk = arr[_box(t, ids)] if arr is not None else np.zeros(t.dims(ids))
                bodies[key] = gpu_tensor(np.asfortranarray(blk).ravel(order="F"), list(blk.shape))
            return bodies[key]
        return get

    split = TR.contraction_split(PAT, D, L, R)
    h = TR.to_batch(lib, PAT, D, L, R, split, 0, body("D", D, None), body("L", L, la), body("R", R, ra))
    assert lib.talshAmdBatchExecute(h, 1) == 0
    assert lib.talshAmdBatchDestroy(h) == 0
    want = np.zeros([sp[m].dim for m in "abcd"])
    for d, l, r in split:
        db, lb, rb = _box(D, D.subtensors[d]), _box(L, L.subtensors[l]), _box(R, R.subtensors[r])
        want[db] += np.einsum("akcl,lbkd->abcd", la[lb], ra[rb])
    worst = 0.0
    for (tag, ids), t in bodies.items():
        if tag != "D":
            continue
        got = t.download().reshape(D.dims(ids), order="F")
        ref = want[_box(D, ids)]
        worst = max(worst, float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)))
    for t in bodies.values():
        t.destruct()
    assert worst <= 1e-12


def test_to_batch_rejects_mixed_levels():
    sp, D, L, R = _tensors(12, (2, 1, 2), 1.0)
    split = TR.contraction_split(PAT, D, L, R)
    with pytest.raises(ValueError):
        TR.to_batch(None, PAT, D, L, R, split, 0, None, None, None)


def test_split_c_abi_capacity_and_rejections():
    """talshAmdContractionSplit directly: a capacity below the number of
    block contractions reports TALSH_LIMIT_EXCEEDED with the count (and the
    first cap triples in order), conjugated or rank-mismatched patterns and
    out-of-range subspace ids are rejected."""
    import ctypes as C
    from exatensor_amd import talsh as T
    sp, D, L, R = _tensors(2, (1, 1, 1), 0.5)
    full = TR.contraction_split(PAT, D, L, R)
    assert len(full) > 8
    lib = TR._lib()
    args = []
    for t in (D, L, R):
        args += [t.rank, *TR._arg(t)]
    out = (C.c_int * 24)()
    n = C.c_longlong()
    assert lib.talshAmdContractionSplit(PAT.encode(), *args, 8, out, C.byref(n)) == T.TALSH_LIMIT_EXCEEDED
    assert n.value == len(full)
    assert [tuple(out[3 * i:3 * i + 3]) for i in range(8)] == full[:8]
    assert lib.talshAmdContractionSplit(b"D(a,b,c,d)+=L(a,k,c,l)*R+(l,b,k,d)", *args, 8, out, C.byref(n)) == \
        T.TALSH_INVALID_ARGS
    bad = list(args)
    bad[7] = T.ints([x if i else 10 ** 6 for i, x in enumerate(sum((list(s) for s in L.subtensors), []))])
    assert lib.talshAmdContractionSplit(PAT.encode(), *bad, 8, out, C.byref(n)) == T.TALSH_INVALID_ARGS
