"""Fused-gather plan candidates on the host (no GPU): how many the planner
offers for a contraction, and when it offers none."""
import ctypes as C

import numpy as np
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


def _table(ext, stri, nd):
    off = np.zeros(1, dtype=np.int64)
    for j in range(nd):  # first dimension fastest
        off = (off[None, :] + np.arange(ext[j], dtype=np.int64)[:, None] * stri[j]).ravel()
    return off


def test_maps_address_every_element_once_and_reproduce_the_product(lib):
    """For random contractions of ranks up to 9 with unit and small extents,
    every fused candidate's index groups address each element of A, B and C
    exactly once, stage each operand along the direction of its stride-1
    index, and the product assembled through them in numpy equals einsum."""
    rng = np.random.default_rng(41)
    checked = 0
    for case in range(60):
        nl, nr, nc = (int(v) for v in rng.integers(1, 4, 3))
        nc = int(rng.integers(0, 4))
        ext = {f"x{i}": int(rng.choice([1, 2, 3, 4])) for i in range(nl + nr + nc)}
        names = list(ext)
        fl, fr, cc = names[:nl], names[nl:nl + nr], names[nl + nr:]
        dlab = list(rng.permutation(fl + fr))
        llab = list(rng.permutation(fl + cc))
        rlab = list(rng.permutation(cc + fr))
        pat = f"D({','.join(dlab)})+=L({','.join(llab)})*R({','.join(rlab)})"
        dd, ld, rd = ([ext[x] for x in lab] for lab in (dlab, llab, rlab))
        e, nf = _count(lib, pat, dd, ld, rd)
        assert e == 0 and 1 <= nf <= 4, pat
        L = rng.uniform(-1, 1, int(np.prod(ld)))
        R = rng.uniform(-1, 1, int(np.prod(rd)))
        letters = {x: chr(97 + i) for i, x in enumerate(names)}
        spec = (''.join(letters[x] for x in llab) + ',' + ''.join(letters[x] for x in rlab) + '->' +
                ''.join(letters[x] for x in dlab))
        want = np.einsum(spec, L.reshape(ld, order="F"), R.reshape(rd, order="F")).ravel(order="F")

        def unit_label(labs):
            for x in labs:
                if ext[x] > 1:
                    return x
            return None

        for v in range(nf):
            info = (C.c_int * 9)()
            ex = (C.c_int * 96)()
            st = (C.c_longlong * 96)()
            assert lib.talshAmdPlanFusedMaps(pat.encode(), len(dd), T.ints(dd), len(ld), T.ints(ld), len(rd),
                                             T.ints(rd), v, info, ex, st) == 0
            tabs = [_table(ex[16 * g:16 * g + 16], st[16 * g:16 * g + 16], info[3 + g]) for g in range(6)]
            a_row, a_k, b_row, b_k, c_row, c_col = tabs
            swap = info[0] == 1
            A, B = (R, L) if swap else (L, R)
            alab, blab = (rlab, llab) if swap else (llab, rlab)
            assert len(a_k) == len(b_k) and len(c_row) == len(a_row) and len(c_col) == len(b_row), pat
            for rows, ks, vol in ((a_row, a_k, A.size), (b_row, b_k, B.size), (c_row, c_col, len(want))):
                addr = np.sort((rows[:, None] + ks[None, :]).ravel())
                assert np.array_equal(addr, np.arange(vol)), (pat, v)
            # staged along k exactly when the operand's stride-1 index is contracted
            for kf, labs in ((info[1], alab), (info[2], blab)):
                u = unit_label(labs)
                assert kf == (1 if u is not None and u in cc else 0), (pat, v)
            Am = A[a_row[:, None] + a_k[None, :]]
            Bm = B[b_row[:, None] + b_k[None, :]]
            got = np.zeros(len(want))
            got[c_row[:, None] + c_col[None, :]] = Am @ Bm.T
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12, err_msg=f"{pat} candidate {v}")
            checked += 1
    assert checked >= 120
