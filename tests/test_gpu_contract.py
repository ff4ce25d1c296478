"""Tensor contractions, additions, initialization and scaling on the GPU
through the talsh.h C-ABI, against the reference's golden outputs, the
oracle and long-double direct evaluation. Tolerances (relative Frobenius):
1e-12 for R8/C8, 1e-5 for R4/C4."""
import json
import os

import numpy as np
import pytest

from exatensor_amd import talsh as T
from gpu_util import contract, engine, gpu_tensor, rel_err

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _tol(kind):
    return 1e-12 if kind in (T.R8, T.C8) else 1e-5


@pytest.fixture(scope="module")
def golden():
    meta = json.load(open(os.path.join(GOLD, "tensor_ops.json")))["cases"]
    arrs = np.load(os.path.join(GOLD, "tensor_ops.npz"))
    return meta, arrs


def test_golden_contractions(golden, oracle):
    meta, arrs = golden
    n = ndirect = 0
    for c in meta:
        if c["op"] != "contract" or c["ierr"] != 0:
            continue
        k = c["key"]
        alpha = complex(*c["alpha"]) if c["kind"] in (T.C4, T.C8) else c["alpha"][0]
        d0 = arrs[k + "_d0"]
        e, got = contract(c["pattern"], d0, c["ddims"], arrs[k + "_l"], c["ldims"], arrs[k + "_r"], c["rdims"], alpha,
                          c["accumulative"])
        assert e == 0, c
        exact = oracle.contract_exact(c["pattern"], d0, c["ddims"], arrs[k + "_l"], c["ldims"], arrs[k + "_r"],
                                      c["rdims"], alpha, c["accumulative"])
        assert rel_err(got, exact) <= _tol(c["kind"]), c
        n += 1
        # and directly against the output the reference's CPU path wrote,
        # except where its -DNO_BLAS build mis-orders a conjugated R operand
        e0, dig, dr, lr, rr, cj = oracle.str2dig(c["pattern"])
        lconj, rconj = (cj >> 1) & 1, (cj >> 2) & 1
        if cj & 1:
            lconj, rconj = 1 - lconj, 1 - rconj
        if c["kind"] in (T.C4, T.C8) and rconj and lr > 0 and rr > 0:
            continue
        assert rel_err(got, arrs[k + "_d"]) <= _tol(c["kind"]), c
        ndirect += 1
    assert n > 150 and ndirect > 100


def test_golden_copies_adds_init_scale(golden):
    meta, arrs = golden
    eng = engine()
    L = eng.lib
    for c in meta:
        k = c["key"]
        if c["op"] == "copy":
            tl = gpu_tensor(arrs[k + "_l"], c["ldims"])
            td = T.Tensor(L, c["kind"], c["ddims"], eng.gpu(0))
            assert L.talshTensorCopy(c["pattern"].encode(), td.ptr, tl.ptr, eng.gpu(0), T.DEV_NULL, T.COPY_MT,
                                     None) == 0
            assert np.array_equal(td.download().view(np.uint8), arrs[k + "_d"].view(np.uint8)), c
            td.destruct(); tl.destruct()
        elif c["op"] == "add":
            tl = gpu_tensor(arrs[k + "_l"], c["ldims"])
            td = gpu_tensor(arrs[k + "_d0"], c["ddims"])
            s = complex(*c["scale"])
            assert L.talshTensorAdd(c["pattern"].encode(), td.ptr, tl.ptr, s.real, s.imag, eng.gpu(0), T.DEV_NULL,
                                    T.COPY_MT, None) == 0
            got, exp = td.download(), arrs[k + "_d"]
            if c["kind"] in (T.R4, T.R8):
                assert np.array_equal(got, exp), c  # separate multiply and add, as on the CPU path
            else:
                assert rel_err(got, exp) <= 1e-6 if c["kind"] == T.C4 else rel_err(got, exp) <= 1e-15, c
            td.destruct(); tl.destruct()
        elif c["op"] == "init_scale":
            v = complex(*c["value"])
            td = gpu_tensor(arrs[k + "_d0"], c["ddims"])
            assert L.talshTensorInit(td.ptr, v.real, v.imag, eng.gpu(0), T.DEV_NULL, T.COPY_M, None) == 0
            assert np.array_equal(td.download(), arrs[k + "_init"]), c
            td.upload(arrs[k + "_d0"])
            assert L.talshTensorScale(td.ptr, v.real, v.imag, eng.gpu(0), T.DEV_NULL, T.COPY_M, None) == 0
            got, exp = td.download(), arrs[k + "_scale"]
            if c["kind"] in (T.R4, T.R8):
                assert np.array_equal(got, exp), c
            else:
                assert rel_err(got, exp) <= 1e-6, c
            td.destruct()


def test_closed_form_reference_test_cpp():
    """TALSH/test.cpp:92-147: D(a,b,i,j)+=L(c,b,d,a)*R(j,d,i,c)*2 with constant
    L=0.001, R=0.01: every element of D is 2*0.001*0.01*(c*d extent)."""
    V = 24
    dims = [V] * 4
    e, got = contract("D(a,b,i,j)+=L(c,b,d,a)*R(j,d,i,c)", np.zeros(V ** 4), dims, np.full(V ** 4, 0.001), dims,
                      np.full(V ** 4, 0.01), dims, 2.0)
    assert e == 0
    np.testing.assert_allclose(got, 2 * 0.001 * 0.01 * V * V, rtol=1e-13)


@pytest.mark.parametrize("kind", [T.R8, T.R4, T.C8, T.C4])
def test_config1_pattern_full_size(oracle, kind):
    """D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a), every extent 32 (the CPU-runnable
    case), against the oracle and long-double direct evaluation (sampled)."""
    pat = "D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)"
    dims = [32] * 4
    n = 32 ** 4
    l = oracle.uniform(kind, [n], 1)
    r = oracle.uniform(kind, [n], 2)
    d = np.zeros(n, dtype=T.NP_DTYPE[kind])
    e, got = contract(pat, d, dims, l, dims, r, dims, 1.0)
    assert e == 0
    ref = d.copy()
    assert oracle.contract(pat, ref, dims, l, dims, r, dims, 1.0) == 0
    assert rel_err(got, ref) <= _tol(kind)
    pos = np.random.default_rng(0).choice(n, 2000, replace=False).astype(np.uint64)
    ex = oracle.contract_exact(pat, d, dims, l, dims, r, dims, 1.0, True, pos)
    assert rel_err(got[pos.astype(np.int64)], ex) <= _tol(kind)


def test_config3_r8_full_size_sampled(oracle):
    """D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j), a=b=i=j=96, k=l=128 in R8 on FP64
    matrix cores; 4096 sampled D elements against long-double sums."""
    pat = "D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)"
    dd, ld, rd = [96] * 4, [96, 128, 96, 128], [128, 96, 128, 96]
    l = oracle.uniform(T.R8, [int(np.prod(ld))], 3)
    r = oracle.uniform(T.R8, [int(np.prod(rd))], 4)
    d = np.zeros(96 ** 4)
    e, got = contract(pat, d, dd, l, ld, r, rd, 1.0)
    assert e == 0
    pos = np.random.default_rng(1).choice(d.size, 4096, replace=False).astype(np.uint64)
    ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, 1.0, True, pos)
    assert rel_err(got[pos.astype(np.int64)], ex) <= 1e-12


@pytest.mark.parametrize("fast", [False, True])
def test_config3_r4_modes_sampled(oracle, fast):
    """R4 as native FP32 (<=1e-5) and as the three-way BF16 split (<=1e-6 vs the
    FP64 evaluation of the same FP32 inputs)."""
    eng = engine()
    L = eng.lib
    pat = "D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)"
    dd, ld, rd = [96] * 4, [96, 128, 96, 128], [128, 96, 128, 96]
    l = oracle.uniform(T.R4, [int(np.prod(ld))], 3)
    r = oracle.uniform(T.R4, [int(np.prod(rd))], 4)
    d = np.zeros(96 ** 4, dtype=np.float32)
    if fast:
        assert L.talshEnableFastMath(T.DEV_AMD_GPU, 0) == 0
    try:
        e, got = contract(pat, d, dd, l, ld, r, rd, 1.0)
    finally:
        L.talshDisableFastMath(T.DEV_AMD_GPU, 0)
    assert e == 0
    pos = np.random.default_rng(2).choice(d.size, 4096, replace=False).astype(np.uint64)
    ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, 1.0, True, pos)
    err = rel_err(got[pos.astype(np.int64)], ex)
    assert err <= (1e-6 if fast else 1e-5), err


def test_config3_c8_sampled(oracle):
    pat = "D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)"
    dd, ld, rd = [48] * 4, [48, 64, 48, 64], [64, 48, 64, 48]
    l = oracle.normal(T.C8, int(np.prod(ld)), 5)
    r = oracle.normal(T.C8, int(np.prod(rd)), 6)
    d = np.zeros(48 ** 4, dtype=np.complex128)
    e, got = contract(pat, d, dd, l, ld, r, rd, 1.0)
    assert e == 0
    pos = np.random.default_rng(3).choice(d.size, 2048, replace=False).astype(np.uint64)
    ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, 1.0, True, pos)
    assert rel_err(got[pos.astype(np.int64)], ex) <= 1e-12


def test_conjugation_and_alpha_complex(oracle):
    rng = np.random.default_rng(9)
    for pat in ["D(a,b)+=L+(a,i)*R(i,b)", "D(a,b)+=L(a,i)*R+(i,b)", "D+(a,b)+=L(i,a)*R(b,i)",
                "D(a,b,c)+=L+(c,i,a)*R+(b,i)"]:
        for kind in (T.C8, T.C4):
            e0, dig, dr, lr, rr, cj = oracle.str2dig(pat)
            ext = {}
            lab_l = pat[pat.index("(", pat.index("L")) + 1:pat.index(")", pat.index("L"))].split(",")
            lab_r = pat[pat.index("(", pat.index("R")) + 1:pat.index(")", pat.index("R"))].split(",")
            lab_d = pat[pat.index("(") + 1:pat.index(")")].split(",")
            for x in set(lab_l + lab_r + lab_d):
                ext[x] = int(rng.integers(3, 40))
            ld = [ext[x] for x in lab_l]; rd = [ext[x] for x in lab_r]; dd = [ext[x] for x in lab_d]
            mk = lambda n: (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(T.NP_DTYPE[kind])
            l, r, d = mk(int(np.prod(ld))), mk(int(np.prod(rd))), mk(int(np.prod(dd)))
            alpha = complex(0.7, -0.4)
            e, got = contract(pat, d, dd, l, ld, r, rd, alpha)
            assert e == 0, pat
            ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, alpha)
            assert rel_err(got, ex) <= _tol(kind), (pat, kind)


def test_every_plan_variant_gives_the_same_result(oracle):
    eng = engine()
    L = eng.lib
    pat = "D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)"
    dims = [12, 10, 9, 11]
    ld = [11, 7, 10, 6]; rd = [6, 9, 7, 12]
    l = oracle.uniform(T.R8, [int(np.prod(ld))], 7)
    r = oracle.uniform(T.R8, [int(np.prod(rd))], 8)
    d = oracle.uniform(T.R8, [int(np.prod(dims))], 9)
    ex = oracle.contract_exact(pat, d, dims, l, ld, r, rd, 0.5)
    import ctypes as C
    moved = (C.c_double * 64)()
    nc = C.c_int()
    assert L.talshAmdPlanCandidates(pat.encode(), 4, T.ints(dims), 4, T.ints(ld), 4, T.ints(rd), T.R8, 1, 64, moved,
                                    C.byref(nc)) == 0
    assert nc.value >= 8
    try:
        for v in range(nc.value):
            L.talshAmdSetPlanVariant(v, -1)
            e, got = contract(pat, d, dims, l, ld, r, rd, 0.5)
            assert e == 0
            assert rel_err(got, ex) <= 1e-12, v
    finally:
        L.talshAmdSetPlanVariant(-1, -1)


def test_non_accumulative_and_edge_shapes(oracle):
    rng = np.random.default_rng(4)
    cases = [("D(a,b)+=L(a,i)*R(i,b)", [1, 1], [1, 1], [1, 1]),
             ("D(a,b)+=L(a,i)*R(i,b)", [129, 67], [129, 257], [257, 67]),
             ("D(a,b)+=L(i,a)*R(b,i)", [3, 500], [1000, 3], [500, 1000]),
             ("D(a)+=L(a,i)*R(i)", [77], [77, 33], [33]),
             ("D()+=L(a,b)*R(b,a)", [], [70, 60], [60, 70]),
             ("D(a,b)+=L(b,a)*R()", [13, 17], [17, 13], []),
             ("D(a,b)+=L()*R(a,b)", [13, 17], [], [13, 17]),
             ("D()+=L()*R()", [], [], []),
             ("D(a,b,c)+=L(c,a)*R(b)", [4, 5, 6], [6, 4], [5])]
    for pat, dd, ld, rd in cases:
        for acc in (True, False):
            n = lambda dims: int(np.prod(dims)) if dims else 1
            l = rng.uniform(-1, 1, n(ld)); r = rng.uniform(-1, 1, n(rd)); d = rng.uniform(-1, 1, n(dd))
            e, got = contract(pat, d, dd, l, ld, r, rd, 1.5, acc)
            assert e == 0, pat
            ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, 1.5, acc)
            assert rel_err(got, ex) <= 1e-12, (pat, acc)


def test_async_task_and_coherence(oracle):
    """COPY_MTT with host operands: D moves to the GPU, L and R stay on the
    host only; the task handle reports completion and timings."""
    import ctypes as C
    eng = engine()
    L = eng.lib
    dims = [16, 16]
    l = oracle.uniform(T.R8, [256], 1); r = oracle.uniform(T.R8, [256], 2)
    td = T.Tensor(L, T.R8, dims, 0, 0.0); tl = T.Tensor(L, T.R8, dims, 0); tr = T.Tensor(L, T.R8, dims, 0)
    tl.upload(l); tr.upload(r)
    task = T.TalshTask()
    L.talshTaskClean(C.byref(task))
    e = L.talshTensorContract(b"D(a,b)+=L(a,i)*R(i,b)", td.ptr, tl.ptr, tr.ptr, 1.0, 0.0, eng.gpu(0), T.DEV_NULL,
                              T.COPY_MTT, T.YEP, C.byref(task))
    assert e == 0
    st = C.c_int()
    assert L.talshTaskWait(C.byref(task), C.byref(st)) == 0 and st.value == T.TALSH_TASK_COMPLETED
    tot, comp, inp, out, mm = (C.c_double() for _ in range(5))
    assert L.talshTaskTime(C.byref(task), C.byref(tot), C.byref(comp), C.byref(inp), C.byref(out), C.byref(mm)) == 0
    assert tot.value > 0
    assert L.talshTaskDestruct(C.byref(task)) == 0
    assert td.images() == [eng.gpu(0)]   # M: moved to the executing GPU
    assert tl.images() == [0] and tr.images() == [0]  # T: device copies discarded
    got = td.download()
    ex = (l.reshape(16, 16, order="F") @ r.reshape(16, 16, order="F")).ravel(order="F")
    np.testing.assert_allclose(got, ex, rtol=1e-13, atol=1e-14)
    for t in (td, tl, tr):
        t.destruct()


def test_arena_back_pressure():
    """A body larger than the device arena is DEVICE_UNABLE (not an error) and
    leaves the tensor clean."""
    eng = engine()
    L = eng.lib
    cap = L.talshDeviceBufferSize(0, T.DEV_AMD_GPU)
    t = T.Tensor(L, T.R8, [int(cap // 8 // 4096) + 2, 4096], eng.gpu(0))
    assert t.err == T.DEVICE_UNABLE
    assert L.talshTensorIsEmpty(t.ptr) == T.YEP


@pytest.mark.parametrize("kind", [T.R8, T.C8])
def test_high_rank_contractions(oracle, kind):
    """Random contractions of tensors of rank 4 to 14 with extents 1-4 (ranks
    far above the configured ones: many unit extents, fusable and unfusable
    index runs, free and contracted indices in random orders), against
    long-double direct evaluation."""
    rng = np.random.default_rng(140 + kind)
    for case in range(12):
        nl, nr, nc = (int(v) for v in rng.integers(2, 8, 3))
        while nl + nc > 14 or nr + nc > 14:
            nc -= 1
        ext = {f"x{i}": int(rng.choice([1, 2, 2, 3, 3, 4])) for i in range(nl + nr + nc)}
        names = list(ext)
        fl, fr, cc = names[:nl], names[nl:nl + nr], names[nl + nr:]
        for group in (fl, fr, cc):
            while np.prod([ext[x] for x in group]) > 96:
                ext[max(group, key=ext.get)] //= 2
        dlab = list(rng.permutation(fl + fr))
        llab = list(rng.permutation(fl + cc))
        rlab = list(rng.permutation(cc + fr))
        pat = f"D({','.join(dlab)})+=L({','.join(llab)})*R({','.join(rlab)})"
        dd, ld, rd = ([ext[x] for x in lab] for lab in (dlab, llab, rlab))
        vol = lambda dims: int(np.prod(dims))
        def rand(n):
            x = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if kind == T.C8 else 0)
            return x.astype(T.NP_DTYPE[kind])
        d, l, r = rand(vol(dd)), rand(vol(ld)), rand(vol(rd))
        alpha = 0.75 - 0.5j if kind == T.C8 else -1.25
        e, got = contract(pat, d, dd, l, ld, r, rd, alpha)
        assert e == 0, (pat, e)
        ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, alpha)
        assert rel_err(got, ex) <= 1e-12, pat


@pytest.mark.parametrize("pat,a,b,k", [
    ("D(b,a)+=L(k,a)*R(k,b)", 256, 128, 200),   # planned with the operands exchanged
    ("D(a,b)+=L(k,a)*R(k,b)", 256, 128, 200),
    ("D(b,a)+=L(a,k)*R(b,k)", 130, 300, 96),
    ("D(b,a)+=L(k,a)*R(k,b)", 520, 70, 4100),   # more than one K chunk candidate set, ragged rows
])
def test_fast_math_every_plan_variant_unequal_m_n(oracle, pat, a, b, k):
    """R4 under the three-way BF16 split with m != n, through every
    matricization the planner offers (some exchange the operands, so the
    product runs as n x m): the split planes are padded differently along
    rows of A and B, and the workspace must be sized for the orientation that
    is launched (an undersized workspace is rejected by the launcher). Each
    variant is held to 1e-6 of the FP64 evaluation."""
    eng = engine()
    L = eng.lib
    import ctypes as C
    lab = {"a": a, "b": b, "k": k}
    sect = lambda s: s[s.index("(") + 1:s.index(")")].split(",")
    dpart, rest = pat.split("+=")
    lpart, rpart = rest.split("*")
    dd, ld, rd = ([lab[x] for x in sect(p)] for p in (dpart, lpart, rpart))
    l = oracle.uniform(T.R4, [int(np.prod(ld))], 21)
    r = oracle.uniform(T.R4, [int(np.prod(rd))], 22)
    d = oracle.uniform(T.R4, [int(np.prod(dd))], 23)
    moved = (C.c_double * 64)()
    nc = C.c_int()
    assert L.talshAmdPlanCandidates(pat.encode(), 2, T.ints(dd), 2, T.ints(ld), 2, T.ints(rd), T.R4, 1, 64, moved,
                                    C.byref(nc)) == 0
    assert nc.value >= 2
    for acc in (True, False):
        ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, 0.5, acc)
        assert L.talshEnableFastMath(T.DEV_AMD_GPU, 0) == 0
        try:
            for v in range(nc.value):
                L.talshAmdSetPlanVariant(v, -1)
                e, got = contract(pat, d, dd, l, ld, r, rd, 0.5, acc)
                assert e == 0, (v, e)
                assert rel_err(got, ex) <= 1e-6, (v, acc)
        finally:
            L.talshAmdSetPlanVariant(-1, -1)
            L.talshDisableFastMath(T.DEV_AMD_GPU, 0)


def test_config3_c8_full_size_sampled(oracle):
    """The configured C8 contraction at full size, a=b=i=j=96, k=l=128, real
    and imaginary parts N(0,1): 2048 sampled D elements against long-double
    sums (each over 16384 complex products), <= 1e-12."""
    pat = "D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)"
    dd, ld, rd = [96] * 4, [96, 128, 96, 128], [128, 96, 128, 96]
    l = oracle.normal(T.C8, int(np.prod(ld)), 31)
    r = oracle.normal(T.C8, int(np.prod(rd)), 32)
    d = np.zeros(96 ** 4, dtype=np.complex128)
    e, got = contract(pat, d, dd, l, ld, r, rd, 1.0)
    assert e == 0
    pos = np.concatenate([np.random.default_rng(33).choice(d.size, 2048, replace=False),
                          [0, d.size - 1]]).astype(np.uint64)
    ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, 1.0, True, pos)
    assert rel_err(got[pos.astype(np.int64)], ex) <= 1e-12


@pytest.mark.parametrize("fast", [False, True])
def test_config3_c4_modes_sampled(oracle, fast):
    """The configured contraction in C4 at full size (a=b=i=j=96, k=l=128,
    parts N(0,1)): native FP32 matrix cores (<= 1e-5) and the three-way BF16
    split (<= 1e-6 against the long-double evaluation of the same complex64
    inputs); 2048 sampled elements plus both ends."""
    eng = engine()
    L = eng.lib
    pat = "D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)"
    dd, ld, rd = [96] * 4, [96, 128, 96, 128], [128, 96, 128, 96]
    l = oracle.normal(T.C4, int(np.prod(ld)), 35)
    r = oracle.normal(T.C4, int(np.prod(rd)), 36)
    d = np.zeros(96 ** 4, dtype=np.complex64)
    if fast:
        assert L.talshEnableFastMath(T.DEV_AMD_GPU, 0) == 0
    try:
        e, got = contract(pat, d, dd, l, ld, r, rd, 1.0)
    finally:
        L.talshDisableFastMath(T.DEV_AMD_GPU, 0)
    assert e == 0
    pos = np.concatenate([np.random.default_rng(37).choice(d.size, 2048, replace=False),
                          [0, d.size - 1]]).astype(np.uint64)
    ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, 1.0, True, pos)
    err = rel_err(got[pos.astype(np.int64)], ex)
    assert err <= (1e-6 if fast else 1e-5), err


def test_c4_fast_math_conjugation_alpha_and_plans(oracle):
    """C4 under fast math with conjugated operands (L+, R+, D+), complex
    alpha, accumulating and not, through every plan variant (operands
    exchanged in some): <= 1e-6 against long-double evaluation."""
    import ctypes as C
    eng = engine()
    L = eng.lib
    rng = np.random.default_rng(44)
    for pat in ["D(a,b)+=L+(a,i)*R(i,b)", "D(a,b)+=L(a,i)*R+(i,b)", "D+(a,b)+=L(i,a)*R(b,i)",
                "D(a,b,c)+=L+(c,i,a)*R+(b,i)", "D(b,a)+=L(k,a)*R(k,b)"]:
        sect = lambda s: s[s.index("(") + 1:s.index(")")].split(",")
        dpart, rest = pat.split("+=")
        lpart, rpart = rest.split("*")
        labs = [sect(x) for x in (dpart, lpart, rpart)]
        hi = 300 if len(labs[0]) <= 2 else 40
        ext = {x: int(rng.integers(20, hi)) for x in sorted(set(sum(labs, [])))}
        dd, ld, rd = ([ext[x] for x in lab] for lab in labs)
        mk = lambda n: (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
        l, r, d = mk(int(np.prod(ld))), mk(int(np.prod(rd))), mk(int(np.prod(dd)))
        moved = (C.c_double * 64)()
        nc = C.c_int()
        assert L.talshAmdPlanCandidates(pat.encode(), len(dd), T.ints(dd), len(ld), T.ints(ld), len(rd), T.ints(rd),
                                        T.C4, 1, 64, moved, C.byref(nc)) == 0
        alpha = complex(0.7, -0.4)
        for acc in (True, False):
            ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, alpha, acc)
            assert L.talshEnableFastMath(T.DEV_AMD_GPU, 0) == 0
            try:
                for v in range(nc.value):
                    L.talshAmdSetPlanVariant(v, -1)
                    e, got = contract(pat, d, dd, l, ld, r, rd, alpha, acc)
                    assert e == 0, (pat, v, e)
                    assert rel_err(got, ex) <= 1e-6, (pat, v, acc)
            finally:
                L.talshAmdSetPlanVariant(-1, -1)
                L.talshDisableFastMath(T.DEV_AMD_GPU, 0)


@pytest.mark.parametrize("pat", ["D(a,b)+=L(i,a)*R+(b,i)", "D+(a,b)+=L+(a,i)*R(i,b)", "D(b,a)+=L(a,i)*R(b,i)"])
def test_large_complex_conjugated_contractions(oracle, pat):
    """C8 contractions large enough for the realified product path, with
    conjugated operands (applied while realifying) and complex alpha; 1024
    sampled elements against long-double sums."""
    sect = lambda s: s[s.index("(") + 1:s.index(")")].split(",")
    dpart, rest = pat.split("+=")
    lpart, rpart = rest.split("*")
    ext = {"a": 2304, "b": 2112, "i": 700}
    dd, ld, rd = ([ext[x] for x in sect(p)] for p in (dpart, lpart, rpart))
    l = oracle.normal(T.C8, int(np.prod(ld)), 71)
    r = oracle.normal(T.C8, int(np.prod(rd)), 72)
    d = oracle.normal(T.C8, int(np.prod(dd)), 73)
    alpha = complex(0.6, -0.8)
    e, got = contract(pat, d, dd, l, ld, r, rd, alpha)
    assert e == 0
    pos = np.random.default_rng(74).choice(d.size, 1024, replace=False).astype(np.uint64)
    ex = oracle.contract_exact(pat, d, dd, l, ld, r, rd, alpha, True, pos)
    assert rel_err(got[pos.astype(np.int64)], ex) <= 1e-12
