"""Batched block-sparse contraction (one persistent grouped launch, optionally
replayed from a hipGraph) against the oracle, block by block."""
import ctypes as C

import numpy as np
import pytest

from exatensor_amd import blocksparse as BS
from exatensor_amd import talsh as T
from gpu_util import engine, gpu_tensor, rel_err

pytestmark = pytest.mark.gpu


def _build(kind, segments, fill, max_ops, seed, ext_set=(3, 5, 8, 13, 16, 24)):
    eng = engine()
    ext, ops = BS.generate(seed, segments, fill, max_ops, ext_set)
    rng = np.random.default_rng(seed)
    dt = T.NP_DTYPE[kind]
    host = {}
    tens = {}

    def get(tag, key, modes):
        if (tag, key) not in tens:
            dims = BS.dims(ext, key, modes)
            n = int(np.prod(dims))
            arr = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if kind in (T.C4, T.C8) else 0)
            arr = arr.astype(dt)
            host[(tag, key)] = (arr, dims)
            tens[(tag, key)] = gpu_tensor(arr, dims)
        return tens[(tag, key)]

    D, L, R = [], [], []
    for dk, lk, rk in ops:
        D.append(get("D", dk, "abcd")); L.append(get("L", lk, "akcl")); R.append(get("R", rk, "lbkd"))
    return eng, ext, ops, host, tens, D, L, R


@pytest.mark.parametrize("kind,use_graph,fast", [(T.R8, 0, 0), (T.R8, 1, 0), (T.R4, 1, 0), (T.R4, 1, 1)])
def test_batched_blocks_match_oracle(oracle, kind, use_graph, fast):
    eng, ext, ops, host, tens, D, L, R = _build(kind, 4, 0.35, 300, 11 + kind + fast)
    lib = eng.lib
    n = len(ops)
    arrD = (T._TP * n)(*[C.pointer(t.t) for t in D])
    arrL = (T._TP * n)(*[C.pointer(t.t) for t in L])
    arrR = (T._TP * n)(*[C.pointer(t.t) for t in R])
    b = C.c_void_p()
    if fast:  # R4 three-way BF16 split
        assert lib.talshEnableFastMath(T.DEV_AMD_GPU, 0) == 0
    try:
        assert lib.talshAmdBatchCreate(BS.PATTERN.encode(), n, arrD, arrL, arrR, 0.75, T.YEP, 0, C.byref(b)) == 0
    finally:
        if fast:
            lib.talshDisableFastMath(T.DEV_AMD_GPU, 0)
    for rep in range(2):  # second replay accumulates again
        assert lib.talshAmdBatchExecute(b, use_graph) == 0
    assert lib.talshAmdBatchDestroy(b) == 0
    # expected: each D block += 2 * sum over its ops of 0.75 * L*R (two
    # executions); the split form is held to FP64 sums
    cdt = np.float64 if fast else T.NP_DTYPE[kind]
    exp = {k: v[0].astype(cdt) for k, v in host.items() if k[0] == "D"}
    for dk, lk, rk in ops:
        dd = BS.dims(ext, dk, "abcd")
        tmp = np.zeros(int(np.prod(dd)), dtype=cdt)
        l, ld = host[("L", lk)]; r, rd = host[("R", rk)]
        assert oracle.contract(BS.PATTERN, tmp, dd, l.astype(cdt), ld, r.astype(cdt), rd, 0.75, True) == 0
        exp[("D", dk)] = exp[("D", dk)] + 2 * tmp
    tol = 1e-12 if kind == T.R8 else (1e-6 if fast else 1e-5)
    for key, val in exp.items():
        got = tens[key].download()
        assert rel_err(got, val) <= tol, key
    for t in tens.values():
        t.destruct()


@pytest.mark.parametrize("kind,fast", [(T.R8, 0), (T.R4, 0), (T.R4, 1)])
def test_configured_batch_sampled(oracle, kind, fast):
    """The benchmark's batch at full size: 4096 block contractions of
    D(a,b,c,d)+=L(a,k,c,l)*R(l,b,k,d), 7 segments per mode with extents from
    {8..64}, 20% block fill (multi-tile D blocks, padded K, chains of up to 8
    blocks on one D), seeded U[-1,1] blocks, one grouped launch from a
    hipGraph. 72 D blocks (every block of the deepest chains, the largest
    blocks, and a seeded sample) are checked at up to 512 sampled elements
    against long-double sums over their chains: 1e-12 (R8), 1e-5 (R4),
    1e-6 (R4 three-way BF16 split)."""
    eng = engine()
    lib = eng.lib
    ext, ops = BS.generate()
    assert len(ops) == 4096
    rng = np.random.default_rng(404 + kind + fast)
    chains = {}
    for dk, lk, rk in ops:
        chains.setdefault(dk, []).append((lk, rk))
    keys = sorted(chains)
    by_depth = sorted(keys, key=lambda k: -len(chains[k]))
    by_size = sorted(keys, key=lambda k: -int(np.prod(BS.dims(ext, k, "abcd"))))
    pick = list(dict.fromkeys(by_depth[:24] + by_size[:16]))
    rest = [k for k in keys if k not in pick]
    pick += [rest[i] for i in rng.choice(len(rest), 72 - len(pick), replace=False)]
    assert len(pick) == 72 and len(chains[pick[0]]) == max(len(c) for c in chains.values())
    # blocks that feed a checked D block carry seeded U[-1,1] data; the
    # others (which only feed unchecked D blocks) are constant-initialised on
    # the GPU
    seeded = {("D", k) for k in pick}
    for k in pick:
        for lk, rk in chains[k]:
            seeded.add(("L", lk)); seeded.add(("R", rk))
    dt = T.NP_DTYPE[kind]
    host, tens = {}, {}

    def get(tag, key, modes):
        if (tag, key) not in tens:
            dims = BS.dims(ext, key, modes)
            if (tag, key) in seeded:
                arr = rng.uniform(-1, 1, int(np.prod(dims))).astype(dt)
                host[(tag, key)] = (arr, dims)
                tens[(tag, key)] = gpu_tensor(arr, dims)
            else:
                t = T.Tensor(lib, kind, dims, eng.gpu(0), {"D": 0.0, "L": 0.01, "R": 0.02}[tag])
                assert t.err == 0
                tens[(tag, key)] = t
        return tens[(tag, key)]

    D = [get("D", o[0], "abcd") for o in ops]
    L = [get("L", o[1], "akcl") for o in ops]
    R = [get("R", o[2], "lbkd") for o in ops]
    n = len(ops)
    arr = lambda ts: (T._TP * n)(*[C.pointer(t.t) for t in ts])
    b = C.c_void_p()
    if fast:
        assert lib.talshEnableFastMath(T.DEV_AMD_GPU, 0) == 0
    try:
        assert lib.talshAmdBatchCreate(BS.PATTERN.encode(), n, arr(D), arr(L), arr(R), 0.75, T.YEP, 0,
                                       C.byref(b)) == 0
    finally:
        if fast:
            lib.talshDisableFastMath(T.DEV_AMD_GPU, 0)
    assert lib.talshAmdBatchExecute(b, 1) == 0
    assert lib.talshAmdBatchDestroy(b) == 0
    tol = 1e-12 if kind == T.R8 else (1e-6 if fast else 1e-5)
    for dk in pick:
        d0, dd = host[("D", dk)]
        nd = d0.size
        pos = np.arange(nd) if nd <= 512 else np.concatenate([rng.choice(nd, 510, replace=False), [0, nd - 1]])
        pos = pos.astype(np.uint64)
        want = d0[pos.astype(np.int64)].astype(np.complex128)
        zero = np.zeros(nd, dtype=dt)
        for lk, rk in chains[dk]:
            l, ld = host[("L", lk)]
            r, rd = host[("R", rk)]
            want = want + oracle.contract_exact(BS.PATTERN, zero, dd, l, ld, r, rd, 0.75, True, pos)
        got = tens[("D", dk)].download()[pos.astype(np.int64)]
        assert rel_err(got, want) <= tol, (dk, len(chains[dk]))
    for t in tens.values():
        t.destruct()


@pytest.mark.parametrize("kind,fast,pattern", [
    (T.C8, 0, BS.PATTERN),
    (T.C8, 0, "D(a,b,c,d)+=L+(a,k,c,l)*R(l,b,k,d)"),
    (T.C8, 0, "D+(a,b,c,d)+=L(a,k,c,l)*R+(l,b,k,d)"),
    (T.C4, 0, "D(a,b,c,d)+=L(a,k,c,l)*R+(l,b,k,d)"),
    (T.C4, 1, BS.PATTERN),
    (T.C4, 1, "D+(a,b,c,d)+=L(a,k,c,l)*R(l,b,k,d)"),
])
def test_complex_batches_match_oracle(oracle, kind, fast, pattern):
    """Complex block batches: operands packed realified (L as 2m x 2k, R as
    n x 2k), the real grouped kernels (FP64, FP32, or the three-way BF16
    split for C4 under fast math), conjugation flags of the pattern applied
    in the pack; every D block against the oracle: 1e-12 (C8), 1e-5 (C4),
    1e-6 (C4 split, against complex128 sums)."""
    eng, ext, ops, host, tens, D, L, R = _build(kind, 4, 0.35, 250, 31 + kind + fast)
    lib = eng.lib
    n = len(ops)
    arrD = (T._TP * n)(*[C.pointer(t.t) for t in D])
    arrL = (T._TP * n)(*[C.pointer(t.t) for t in L])
    arrR = (T._TP * n)(*[C.pointer(t.t) for t in R])
    b = C.c_void_p()
    if fast:
        assert lib.talshEnableFastMath(T.DEV_AMD_GPU, 0) == 0
    try:
        assert lib.talshAmdBatchCreate(pattern.encode(), n, arrD, arrL, arrR, 0.75, T.YEP, 0, C.byref(b)) == 0
    finally:
        if fast:
            lib.talshDisableFastMath(T.DEV_AMD_GPU, 0)
    nops, nitems, flops = C.c_int(), C.c_int(), C.c_double()
    assert lib.talshAmdBatchInfo(b, C.byref(nops), C.byref(nitems), C.byref(flops)) == 0
    want_flops = 0.0
    for dk, lk, rk in ops:
        a_, b_, c_, d_ = BS.dims(ext, dk, "abcd")
        want_flops += 8.0 * a_ * b_ * c_ * d_ * ext["k"][lk[1]] * ext["l"][lk[3]]
    assert nops.value == n and flops.value == want_flops
    assert lib.talshAmdBatchExecute(b, 1) == 0
    assert lib.talshAmdBatchDestroy(b) == 0
    cdt = np.complex128 if fast else T.NP_DTYPE[kind]
    exp = {k: v[0].astype(cdt) for k, v in host.items() if k[0] == "D"}
    for dk, lk, rk in ops:
        dd = BS.dims(ext, dk, "abcd")
        l, ld = host[("L", lk)]
        r, rd = host[("R", rk)]
        tmp = exp[("D", dk)]
        assert oracle.contract(pattern, tmp, dd, l.astype(cdt), ld, r.astype(cdt), rd, 0.75, True) == 0
    tol = 1e-12 if kind == T.C8 else (1e-6 if fast else 1e-5)
    for key, val in exp.items():
        assert rel_err(tens[key].download(), val) <= tol, key
    for t in tens.values():
        t.destruct()
