"""Fused index gather on the GPU: contractions whose matrix product reads L
and R and updates D through index offset tables, with no transposes. Every
fused candidate and GEMM tile of each case is checked against long-double
direct evaluation (relative Frobenius 1e-12 for R8/C8, 1e-5 for R4/C4), and
the kernel profile shows that no transpose ran."""
import ctypes as C

import numpy as np
import pytest

from exatensor_amd import talsh as T
from gpu_util import contract, engine, rel_err

pytestmark = pytest.mark.gpu

KT_TRANSPOSE, KT_GEMM = 0, 1

CASES = [
    # (pattern, D, L, R extents)
    ("D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)", [12, 10, 9, 11], [11, 7, 10, 6], [6, 9, 7, 12]),
    ("D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)", [16, 16, 16, 16], [16, 16, 16, 16], [16, 16, 16, 16]),
    ("D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)", [24, 20, 16, 12], [24, 8, 16, 10], [10, 20, 8, 12]),
    ("D(a,b)+=L(i,a)*R(b,i)", [130, 70], [300, 130], [70, 300]),
    ("D(a,b)+=L(a,i)*R(i,b)", [1, 1], [1, 1], [1, 1]),
    ("D(a,b,c)+=L(c,a)*R(b)", [4, 5, 6], [6, 4], [5]),
    ("D(a)+=L(a,i)*R(i)", [77], [77, 33], [33]),
    ("D(a,b,c,d)+=L(a,i,c,j)*R(j,d,i,b)", [6, 1, 5, 7], [6, 3, 5, 1], [1, 7, 3, 1]),
    ("D(c,a,b)+=L(k,a,c)*R(b,k)", [12, 100, 300], [700, 100, 12], [300, 700]),  # split K on 64x64 tiles
]


def _fused_count(L, pat, dd, ld, rd):
    n = C.c_int()
    assert L.talshAmdPlanFusedCandidates(pat.encode(), len(dd), T.ints(dd), len(ld), T.ints(ld), len(rd), T.ints(rd),
                                         C.byref(n)) == 0
    return n.value


def _launches(L, kclass):
    ms, n = C.c_double(), C.c_longlong()
    assert L.talshAmdProfileQuery(kclass, C.byref(ms), C.byref(n)) == 0
    return n.value


def _rand(rng, kind, dims):
    n = int(np.prod(dims)) if dims else 1
    x = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if kind in (T.C4, T.C8) else 0)
    return x.astype(T.NP_DTYPE[kind])


@pytest.mark.parametrize("kind", [T.R8, T.R4, T.C8, T.C4])
def test_every_fused_candidate(oracle, kind):
    eng = engine()
    L = eng.lib
    rng = np.random.default_rng(900 + kind)
    tol = 1e-12 if kind in (T.R8, T.C8) else 1e-5
    tiles = [0, 1, 2] if kind in (T.R8, T.R4) else [0]
    alpha = (0.75 - 0.5j) if kind in (T.C4, T.C8) else -1.25
    ran = 0
    assert L.talshAmdProfile(1) == 0
    try:
        for ci, (pat, dd, ld, rd) in enumerate(CASES):
            big = np.prod(dd) * np.prod(ld) > 1e9
            if big and kind in (T.C4, T.C8):
                continue
            nf = _fused_count(L, pat, dd, ld, rd)
            assert 1 <= nf <= 4, pat
            d, l, r = _rand(rng, kind, dd), _rand(rng, kind, ld), _rand(rng, kind, rd)
            for acc in (True, False):
                ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, alpha, acc)
                for v in range(nf):
                    for t in tiles:
                        L.talshAmdSetPlanVariant(T.FUSED_VARIANT + v, t)
                        _launches(L, KT_TRANSPOSE)
                        _launches(L, KT_GEMM)
                        e, got = contract(pat, d, dd, l, ld, r, rd, alpha, acc)
                        assert e == 0, (pat, v, t, acc, e)
                        assert rel_err(got, ex) <= tol, (pat, v, t, acc, rel_err(got, ex))
                        assert _launches(L, KT_TRANSPOSE) == 0, (pat, v, t)
                        if np.prod(dd) > 1:
                            assert _launches(L, KT_GEMM) >= 1, (pat, v, t)
                        ran += 1
    finally:
        L.talshAmdSetPlanVariant(-1, -1)
        L.talshAmdProfile(0)
    assert ran >= (100 if len(tiles) > 1 else 24), ran


def test_fused_conjugated_operands(oracle):
    """Conjugation flags of L, R and D+ patterns go through the gathered
    complex staging."""
    eng = engine()
    L = eng.lib
    rng = np.random.default_rng(77)
    cases = [("D(a,b)+=L+(i,a)*R(b,i)", [40, 30], [50, 40], [30, 50]),
             ("D(a,b)+=L(i,a)*R+(b,i)", [40, 30], [50, 40], [30, 50]),
             ("D+(a,b,c)+=L(c,i,a)*R+(i,b)", [9, 8, 7], [7, 11, 9], [11, 8])]
    try:
        for kind in (T.C8, T.C4):
            tol = 1e-12 if kind == T.C8 else 1e-5
            for pat, dd, ld, rd in cases:
                d, l, r = _rand(rng, kind, dd), _rand(rng, kind, ld), _rand(rng, kind, rd)
                ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, 0.5 + 0.25j)
                for v in range(_fused_count(L, pat, dd, ld, rd)):
                    L.talshAmdSetPlanVariant(T.FUSED_VARIANT + v, -1)
                    e, got = contract(pat, d, dd, l, ld, r, rd, 0.5 + 0.25j)
                    assert e == 0 and rel_err(got, ex) <= tol, (pat, kind, v)
    finally:
        L.talshAmdSetPlanVariant(-1, -1)


def test_fused_extent32_contraction_full_size(oracle):
    """The extent-32 rank-4 R8 contraction through every fused candidate,
    against the oracle (the reference's CPU algorithm) at full size."""
    eng = engine()
    L = eng.lib
    pat = "D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)"
    dims = [32] * 4
    n = 32 ** 4
    l = oracle.uniform(T.R8, [n], 1)
    r = oracle.uniform(T.R8, [n], 2)
    d = np.zeros(n)
    ref = d.copy()
    assert oracle.contract(pat, ref, dims, l, dims, r, dims, 1.0) == 0
    try:
        for v in range(_fused_count(L, pat, dims, dims, dims)):
            for t in (-1, 0, 2):
                L.talshAmdSetPlanVariant(T.FUSED_VARIANT + v, t)
                e, got = contract(pat, d, dims, l, dims, r, dims, 1.0)
                assert e == 0 and rel_err(got, ref) <= 1e-12, (v, t)
    finally:
        L.talshAmdSetPlanVariant(-1, -1)
