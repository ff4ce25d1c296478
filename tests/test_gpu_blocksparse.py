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
            arr = rng.uniform(-1, 1, int(np.prod(dims))).astype(dt)
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
