"""Tensor slice / insert: the oracle against the reference's fixtures and a
numpy restatement (CPU), the GPU path against the oracle (bit-exact)."""
import json
import os

import numpy as np
import pytest

from exatensor_amd import talsh as T

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _cases():
    meta = json.load(open(os.path.join(GOLD, "slice_insert.json")))["cases"]
    arr = np.load(os.path.join(GOLD, "slice_insert.npz"), allow_pickle=False)
    return meta, arr


def _np_box(tdims, sdims, offs):
    """Index (DLF) of every slice element inside the tensor, by numpy."""
    idx = np.arange(int(np.prod(tdims))).reshape(tdims[::-1])
    sl = tuple(slice(o, o + s) for o, s in zip(offs[::-1], sdims[::-1]))
    return idx[sl].ravel()


def test_oracle_matches_reference_fixtures(oracle):
    meta, arr = _cases()
    for c in meta:
        k = c["key"]
        t, s = arr[k + "_t"], arr[k + "_s"]
        e1, o1 = oracle.slice(t, c["tdims"], s, c["sdims"], c["offsets"], c["accumulative"])
        e2, o2 = oracle.insert(t, c["tdims"], s, c["sdims"], c["offsets"], c["accumulative"])
        assert e1 == 0 and e2 == 0
        assert np.array_equal(o1.view(np.uint8), arr[k + "_slice"].view(np.uint8)), k
        assert np.array_equal(o2.view(np.uint8), arr[k + "_insert"].view(np.uint8)), k


def test_oracle_against_numpy_indexing(oracle):
    rng = np.random.default_rng(5)
    for _ in range(30):
        rank = int(rng.integers(1, 6))
        tdims = [int(x) for x in rng.integers(1, 8, rank)]
        sdims = [int(rng.integers(1, t + 1)) for t in tdims]
        offs = [int(rng.integers(0, t - s + 1)) for t, s in zip(tdims, sdims)]
        t = rng.uniform(-1, 1, int(np.prod(tdims)))
        s = rng.uniform(-1, 1, int(np.prod(sdims)))
        pos = _np_box(tdims, sdims, offs)
        e, o = oracle.slice(t, tdims, s, sdims, offs, False)
        assert e == 0 and np.array_equal(o, t[pos])
        e, o = oracle.slice(t, tdims, s, sdims, offs, True)
        assert e == 0 and np.array_equal(o, s + t[pos])
        want = t.copy(); want[pos] = s
        e, o = oracle.insert(t, tdims, s, sdims, offs, False)
        assert e == 0 and np.array_equal(o, want)
        want = t.copy(); want[pos] += s
        e, o = oracle.insert(t, tdims, s, sdims, offs, True)
        assert e == 0 and np.array_equal(o, want)
    # out-of-range boxes are rejected with the reference's code
    assert oracle.slice(np.zeros(4), [4], np.zeros(2), [2], [3], False)[0] == 17


# ---------------------------------------------------------------------------- GPU
def _gpu_box(op, t, tdims, s, sdims, offs, acc, task=False):
    from gpu_util import engine, gpu_tensor
    eng = engine()
    L = eng.lib
    tt, ts = gpu_tensor(t, tdims), gpu_tensor(s, sdims)
    fn = L.talshTensorSlice if op == "slice" else L.talshTensorInsert
    d, l = (ts, tt) if op == "slice" else (tt, ts)
    import ctypes as C
    tk = T.TalshTask()
    L.talshTaskClean(C.byref(tk))
    e = fn(d.ptr, l.ptr, T.ints(offs), eng.gpu(0), T.DEV_NULL, T.COPY_MT, T.YEP if acc else T.NOPE,
           C.byref(tk) if task else None)
    if task and e == 0:
        st = C.c_int()
        e = L.talshTaskWait(C.byref(tk), C.byref(st))
        L.talshTaskDestruct(C.byref(tk))
    out = d.download() if e == 0 else None
    for x in (tt, ts):
        x.destruct()
    return e, out


@pytest.mark.gpu
def test_gpu_matches_reference_fixtures():
    meta, arr = _cases()
    for i, c in enumerate(meta):
        k = c["key"]
        t, s = arr[k + "_t"], arr[k + "_s"]
        for op in ("slice", "insert"):
            e, o = _gpu_box(op, t, c["tdims"], s, c["sdims"], c["offsets"], c["accumulative"], task=bool(i % 2))
            assert e == 0, (k, op, e)
            assert np.array_equal(o.view(np.uint8), arr[k + "_" + op].view(np.uint8)), (k, op)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [T.R4, T.R8, T.C4, T.C8])
def test_gpu_large_boxes_bit_exact(oracle, kind):
    rng = np.random.default_rng(kind)
    for tdims, sdims, offs in (([300, 200, 9], [17, 199, 5], [283, 1, 4]),
                               ([64, 64, 64, 8], [64, 64, 30, 8], [0, 0, 34, 0]),
                               ([1, 500, 3], [1, 7, 2], [0, 493, 1]),
                               ([7, 5, 3, 4, 6, 2], [3, 5, 2, 1, 4, 2], [4, 0, 1, 3, 2, 0])):
        t = oracle.uniform(kind, [int(np.prod(tdims))], 21)
        s = oracle.uniform(kind, [int(np.prod(sdims))], 22)
        for op in ("slice", "insert"):
            for acc in (False, True):
                e, got = _gpu_box(op, t, tdims, s, sdims, offs, acc)
                fn = oracle.slice if op == "slice" else oracle.insert
                e2, want = fn(t, tdims, s, sdims, offs, acc)
                assert e == 0 and e2 == 0
                assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (tdims, op, acc)


@pytest.mark.gpu
def test_gpu_rejects_out_of_range():
    e, _ = _gpu_box("slice", np.zeros(10), [10], np.zeros(4), [4], [7], False)
    assert e == T.TALSH_INVALID_ARGS


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [T.R8, T.C4])
def test_gpu_high_rank_boxes(oracle, kind):
    """Boxes with more box dimensions than the compact parameter block holds:
    rank 12 with every dimension partial (nothing fuses), rank 20, and rank
    56 (the maximum) with 16 partial dimensions among unit extents; slice and
    insert, plain and accumulating, bit-exact against the oracle."""
    rng = np.random.default_rng(12 + kind)
    cases = []
    t12 = [3, 4, 3, 2, 5, 3, 2, 4, 3, 2, 3, 2]
    cases.append((t12, [x - 1 for x in t12], [1] * 12))
    t20 = [int(x) for x in rng.integers(2, 4, 20)]
    s20 = [int(rng.integers(1, x + 1)) for x in t20]
    cases.append((t20, s20, [int(rng.integers(0, t - s + 1)) for t, s in zip(t20, s20)]))
    t56 = [3] * 16 + [1] * 40
    rng.shuffle(t56)
    cases.append((t56, [2 if x == 3 else 1 for x in t56], [int(rng.integers(0, 2)) if x == 3 else 0 for x in t56]))
    for tdims, sdims, offs in cases:
        t = oracle.uniform(kind, [int(np.prod(tdims))], 41)
        s = oracle.uniform(kind, [int(np.prod(sdims))], 42)
        for op in ("slice", "insert"):
            for acc in (False, True):
                e, got = _gpu_box(op, t, tdims, s, sdims, offs, acc)
                fn = oracle.slice if op == "slice" else oracle.insert
                e2, want = fn(t, tdims, s, sdims, offs, acc)
                assert e == 0 and e2 == 0, (len(tdims), op, e)
                assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (len(tdims), op, acc)


def _hash_u32(idx):
    """Bit pattern of element idx: a multiplicative hash kept below the
    float32 NaN range (0x3FFFFFFF at most)."""
    return ((idx.astype(np.uint64) * np.uint64(2654435761)) & np.uint64(0x3FFFFFFF)).astype(np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("tdims,sdims,offs", [
    ([40000, 2, 30000], [39999, 2, 27000], [1, 0, 3000]),   # 2.16e9 elements in slabs of rows
    ([1100000000, 3], [1100000000, 2], [0, 1]),             # one contiguous run of 2.2e9 elements
])
def test_gpu_slice_beyond_2_31_elements(tdims, sdims, offs):
    """R4 slices of more than 2^31 elements (the box is cut into pieces
    below 2^31 on the host): every tensor element carries a hash of its index
    as its bit pattern; 4M sampled slice positions plus both ends are checked
    against the index formula."""
    from gpu_util import engine, gpu_tensor, pinned
    import ctypes as C
    eng = engine()
    L = eng.lib
    nt = int(np.prod(tdims))
    ns = int(np.prod(sdims))
    assert ns >= 1 << 31
    x = np.empty(nt, dtype=np.uint32)
    step = 1 << 28
    for s0 in range(0, nt, step):
        s1 = min(nt, s0 + step)
        x[s0:s1] = _hash_u32(np.arange(s0, s1, dtype=np.uint64))
    tt = gpu_tensor(x.view(np.float32), tdims)
    del x
    ts = T.Tensor(L, T.R4, sdims, eng.gpu(0))
    assert ts.err == 0
    e = L.talshTensorSlice(ts.ptr, tt.ptr, T.ints(offs), eng.gpu(0), T.DEV_NULL, T.COPY_MT, T.NOPE, None)
    assert e == 0, e
    got = ts.download(pinned(ns * 4, np.float32)).view(np.uint32)
    rng = np.random.default_rng(31)
    o = np.concatenate([rng.integers(0, ns, 1 << 22, dtype=np.int64), np.arange(1 << 16, dtype=np.int64),
                        np.arange(ns - (1 << 16), ns, dtype=np.int64)])
    src = np.zeros(o.size, dtype=np.int64)
    rem, tstride = o.copy(), 1
    for sd, td, off in zip(sdims, tdims, offs):
        src += (rem % sd + off) * tstride
        rem //= sd
        tstride *= td
    assert np.array_equal(got[o], _hash_u32(src))
    ts.destruct(); tt.destruct()
