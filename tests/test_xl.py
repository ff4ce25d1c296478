"""Out-of-device-memory contractions: talshTensorOpDecompose2 against the
reference's decompositions (CPU), and talshTensorContractXL split into pieces
by a small budget against the oracle (GPU), with operands on the host or on
the GPU."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import tensop_ctypes as TC
from exatensor_amd import talsh as T

HERE = os.path.dirname(os.path.abspath(__file__))


def test_decompose2_matches_reference_fixtures():
    lib = C.CDLL(T.LIB_PATH)  # own handle: the helper sets its own argtypes
    lib.talshInit.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_int), C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                              C.c_int, C.c_void_p]
    sz, mx = C.c_size_t(1 << 24), C.c_int()
    started = lib.talshInit(C.byref(sz), C.byref(mx), 0, None, 0, None, 0, None) == 0
    try:
        cases = json.load(open(os.path.join(HERE, "golden", "decompose.json")))["cases"]
        nleaf = 0
        for c in cases:
            got = TC.decompose(lib, c["pattern"], T.R8, c["ddims"], c["ldims"], c["rdims"], c["limit"])
            want = [[tuple(p) for p in leaf] for leaf in c["leaves"]]
            got = [[(list(o), list(d)) for o, d in leaf] for leaf in got]
            want = [[(list(o), list(d)) for o, d in leaf] for leaf in want]
            assert got == want, c["pattern"]
            nleaf += len(got)
        assert nleaf > 1000
    finally:
        if started:
            lib.talshShutdown()


@pytest.mark.gpu
@pytest.mark.parametrize("where", ["host", "gpu"])
@pytest.mark.parametrize("kind,acc", [(T.R8, True), (T.R8, False), (T.C8, True)])
def test_contract_xl_in_pieces(oracle, where, kind, acc):
    from gpu_util import engine, rel_err
    eng = engine()
    L = eng.lib
    pat = "D(a,b,c,d)+=L(a,k,c,l)*R(l,b,k,d)"
    ext = dict(a=13, b=11, c=9, d=12, k=10, l=7)
    dd = [ext[x] for x in "abcd"]; ld = [ext[x] for x in "akcl"]; rd = [ext[x] for x in "lbkd"]
    vol = lambda v: int(np.prod(v))
    if kind == T.R8:
        d, l, r = (oracle.uniform(T.R8, [vol(v)], 60 + i) for i, v in enumerate((dd, ld, rd)))
        alpha = 0.75
    else:
        d, l, r = (oracle.normal(T.C8, vol(v), 60 + i) for i, v in enumerate((dd, ld, rd)))
        alpha = complex(0.5, -0.25)
    want = oracle.contract_exact(pat, d, dd, l, ld, r, rd, alpha, acc)
    dev = 0 if where == "host" else eng.gpu(0)
    ts = []
    for arr, dims in ((d, dd), (l, ld), (r, rd)):
        t = T.Tensor(L, kind, dims, dev, 0.0)
        assert t.err == 0
        t.upload(arr)
        ts.append(t)
    es = 8 if kind == T.R8 else 16
    total = (vol(dd) + vol(ld) + vol(rd)) * es
    a = complex(alpha)
    try:
        assert L.talshAmdSetXLBudget(3 * total // 7) == 0  # forces several levels of halving
        e = L.talshTensorContractXL(pat.encode(), ts[0].ptr, ts[1].ptr, ts[2].ptr, a.real, a.imag, eng.gpu(0),
                                    T.DEV_NULL, T.YEP if acc else T.NOPE)
    finally:
        L.talshAmdSetXLBudget(0)
    assert e == 0
    got = ts[0].download()
    assert ts[0].images() == [dev]
    assert rel_err(got, want) <= 1e-12
    for t in ts:
        t.destruct()


@pytest.mark.gpu
def test_contract_xl_whole_when_it_fits(oracle):
    from gpu_util import engine, rel_err
    eng = engine()
    L = eng.lib
    pat = "D(a,b)+=L(a,i)*R(i,b)"
    d, l, r = (oracle.uniform(T.R8, [n], 70 + i) for i, n in enumerate((30 * 40, 30 * 50, 50 * 40)))
    ts = []
    for arr, dims in ((d, [30, 40]), (l, [30, 50]), (r, [50, 40])):
        t = T.Tensor(L, T.R8, dims, 0, 0.0)
        t.upload(arr)
        ts.append(t)
    e = L.talshTensorContractXL(pat.encode(), ts[0].ptr, ts[1].ptr, ts[2].ptr, 1.0, 0.0, eng.gpu(0), T.DEV_NULL, T.YEP)
    assert e == 0
    assert 0 in ts[0].images()  # the result is back on the host
    want = oracle.contract_exact(pat, d, [30, 40], l, [30, 50], r, [50, 40], 1.0, True)
    assert rel_err(ts[0].download(), want) <= 1e-12
    for t in ts:
        t.destruct()
