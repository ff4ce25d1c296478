"""Fused-gather plan candidates on the host (no GPU): how many the planner
offers for a contraction, and when it offers none."""
import ctypes as C

import pytest

from exatensor_amd import talsh as T


@pytest.fixture(scope="module")
def lib():
    return T.load_library()


def _count(lib, pat, dd, ld, rd):
    n = C.c_int(-1)
    e = lib.talshAmdPlanFusedCandidates(pat.encode(), len(dd), T.ints(dd), len(ld), T.ints(ld), len(rd), T.ints(rd),
                                        C.byref(n))
    return e, n.value


def test_candidate_counts(lib):
    # contracted order from L or from R (two when they differ), times D or its transpose
    assert _count(lib, "D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)", [8] * 4, [8] * 4, [8] * 4) == (0, 4)
    # one contracted index: its order is the same from either side
    assert _count(lib, "D(a,b)+=L(a,k)*R(k,b)", [5, 6], [5, 7], [7, 6]) == (0, 2)
    # unit extents everywhere still plan (every index group is empty)
    assert _count(lib, "D(a,b)+=L(a,k)*R(k,b)", [1, 1], [1, 1], [1, 1]) == (0, 2)


def test_group_too_wide_has_no_fused_candidate(lib):
    """17 free indices of L interleaved with contracted ones: no two of them
    are adjacent in L, so the row group keeps 17 dimensions, one more than
    the offset-table kernel describes; the planner then offers only the
    transposing plans."""
    f = [f"a{i}" for i in range(17)]
    k = [f"k{i}" for i in range(17)]
    lab_l = [x for pair in zip(f, k) for x in pair]
    pat = f"D({','.join(f + ['b'])})+=L({','.join(lab_l)})*R({','.join(k + ['b'])})"
    assert _count(lib, pat, [2] * 18, [2] * 34, [2] * 18) == (0, 0)
    # the same indices with L's free ones adjacent merge into one run: candidates again
    lab_l2 = f + k
    pat2 = f"D({','.join(f + ['b'])})+=L({','.join(lab_l2)})*R({','.join(k + ['b'])})"
    e, n = _count(lib, pat2, [2] * 18, [2] * 34, [2] * 18)
    assert e == 0 and n >= 1


def test_hyper_index_is_rejected(lib):
    e, _ = _count(lib, "D(a,b)+=L(a,b)*R(b,a)", [3, 4], [3, 4], [4, 3])
    assert e != 0
