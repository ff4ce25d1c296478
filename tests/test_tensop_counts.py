"""Flop and byte accounting of tensor operations (talshTensorOpGetFlopCount,
GetByteCount, GetIntensity; TALSH/talshc.cpp:3068-3148) on host tensors:
against the closed forms (flops = 2 or 8 times the product of the distinct
index extents; bytes = element size times the summed argument volumes),
against the children of talshTensorOpDecompose2 (the halves' flops add up to
the parent's), and against the reference's own library when it is built."""
import ctypes as C
import os
import random

import numpy as np
import pytest

import tensop_ctypes as TC
from exatensor_amd import talsh as T

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libtalsh_ref.so")


def _bind(lib):
    P = C.POINTER
    lib.talshInit.argtypes = [P(C.c_size_t), P(C.c_int), C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                              C.c_void_p]
    lib.talshTensorClean.argtypes = [P(TC.Tens)]
    lib.talshTensorConstruct.argtypes = [P(TC.Tens), C.c_int, C.c_int, P(C.c_int), C.c_int, C.c_void_p, C.c_int,
                                         C.c_void_p, C.c_double, C.c_double]
    lib.talshTensorDestruct.argtypes = [P(TC.Tens)]
    lib.talshTensorOpClean.argtypes = [P(TC.TensOp)]
    lib.talshTensorOpSetArgument.argtypes = [P(TC.TensOp), P(TC.Tens), P(C.c_size_t), P(C.c_int)]
    lib.talshTensorOpSpecify.argtypes = [P(TC.TensOp), C.c_int, C.c_int, C.c_char_p, C.c_double, C.c_double]
    lib.talshTensorOpDecompose2.argtypes = [P(TC.TensOp), P(TC.TensOp), P(TC.TensOp)]
    lib.talshTensorOpDestruct.argtypes = [P(TC.TensOp)]
    lib.talshTensorOpGetFlopCount.argtypes = [P(TC.TensOp)]
    lib.talshTensorOpGetFlopCount.restype = C.c_double
    lib.talshTensorOpGetByteCount.argtypes = [P(TC.TensOp), C.c_uint]
    lib.talshTensorOpGetByteCount.restype = C.c_double
    lib.talshTensorOpGetIntensity.argtypes = [P(TC.TensOp)]
    lib.talshTensorOpGetIntensity.restype = C.c_double
    return lib


def _start(lib):
    sz, mx = C.c_size_t(1 << 24), C.c_int()
    return lib.talshInit(C.byref(sz), C.byref(mx), 0, None, 0, None, 0, None) == 0


def _op(lib, pattern, kind, dims3):
    """A contraction operation over three new host tensors; returns (op, tensors)."""
    tens = []
    for dims in dims3:
        t = TC.Tens()
        lib.talshTensorClean(C.byref(t))
        assert lib.talshTensorConstruct(C.byref(t), kind, len(dims), TC._ia(dims), 0, None, -1, None, 0.0, 0.0) == 0
        tens.append(t)
    op = TC.TensOp()
    lib.talshTensorOpClean(C.byref(op))
    for t, dims in zip(tens, dims3):
        z = (C.c_size_t * max(len(dims), 1))()
        assert lib.talshTensorOpSetArgument(C.byref(op), C.byref(t), z, TC._ia(dims)) == 0
    pat = pattern.encode()  # referenced by the operation: kept alive with it
    assert lib.talshTensorOpSpecify(C.byref(op), TC.TENSOR_CONTRACT, kind, pat, 1.0, 0.0) == 0
    tens.append(pat)
    return op, tens


def _free(lib, op, tens):
    lib.talshTensorOpDestruct(C.byref(op))
    for t in tens[:3]:
        lib.talshTensorDestruct(C.byref(t))


def _cases(n, seed):
    """Random contractions (free, contracted and hyper indices) with labels."""
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        nl, nr, nc, nh = rng.randint(0, 3), rng.randint(0, 3), rng.randint(0, 3), rng.randint(0, 1)
        names = [f"i{q}" for q in range(nl + nr + nc + nh)]
        ext = {x: rng.randint(1, 9) for x in names}
        fl, fr = names[:nl], names[nl:nl + nr]
        cc, hh = names[nl + nr:nl + nr + nc], names[nl + nr + nc:]
        dl = fl + fr + hh
        rng.shuffle(dl)
        ll = fl + cc + hh
        rng.shuffle(ll)
        rl = cc + fr + hh
        rng.shuffle(rl)
        pat = f"D({','.join(dl)})+=L({','.join(ll)})*R({','.join(rl)})"
        kind = rng.choice([T.R4, T.R8, T.C4, T.C8])
        out.append((pat, kind, [[ext[x] for x in lab] for lab in (dl, ll, rl)], ext))
    return out


ES = {T.R4: 4, T.R8: 8, T.C4: 8, T.C8: 16}


def test_flop_byte_counts_closed_form_and_halving():
    lib = _bind(C.CDLL(T.LIB_PATH))
    started = _start(lib)
    try:
        for pat, kind, dims3, ext in _cases(200, 7):
            op, tens = _op(lib, pat, kind, dims3)
            flops = (8.0 if kind in (T.C4, T.C8) else 2.0) * float(np.prod(list(ext.values())))
            nbytes = ES[kind] * sum(float(np.prod(d)) if d else 1.0 for d in dims3)
            assert lib.talshTensorOpGetFlopCount(C.byref(op)) == flops, pat
            assert lib.talshTensorOpGetByteCount(C.byref(op), 1) == nbytes, pat
            assert lib.talshTensorOpGetIntensity(C.byref(op)) == pytest.approx(flops / nbytes, rel=1e-15), pat
            if sum(float(np.prod(d)) if d else 1.0 for d in dims3) > 3:
                c1, c2 = TC.TensOp(), TC.TensOp()
                lib.talshTensorOpClean(C.byref(c1)); lib.talshTensorOpClean(C.byref(c2))
                if lib.talshTensorOpDecompose2(C.byref(op), C.byref(c1), C.byref(c2)) == 0:
                    f1 = lib.talshTensorOpGetFlopCount(C.byref(c1))
                    f2 = lib.talshTensorOpGetFlopCount(C.byref(c2))
                    assert f1 + f2 == flops, pat
                lib.talshTensorOpDestruct(C.byref(c1)); lib.talshTensorOpDestruct(C.byref(c2))
            _free(lib, op, tens)
    finally:
        if started:
            lib.talshShutdown()


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference CPU library not built")
def test_flop_byte_counts_match_reference():
    ours = _bind(C.CDLL(T.LIB_PATH))
    ref = _bind(C.CDLL(REF_SO))
    s1, s2 = _start(ours), _start(ref)
    try:
        for pat, kind, dims3, ext in _cases(150, 11):
            a, ta = _op(ours, pat, kind, dims3)
            b, tb = _op(ref, pat, kind, dims3)
            for es in (1, 4):
                assert ours.talshTensorOpGetByteCount(C.byref(a), es) == ref.talshTensorOpGetByteCount(C.byref(b), es)
            assert ours.talshTensorOpGetFlopCount(C.byref(a)) == ref.talshTensorOpGetFlopCount(C.byref(b)), pat
            assert ours.talshTensorOpGetIntensity(C.byref(a)) == ref.talshTensorOpGetIntensity(C.byref(b)), pat
            _free(ours, a, ta)
            _free(ref, b, tb)
    finally:
        if s1:
            ours.talshShutdown()
        if s2:
            ref.talshShutdown()
