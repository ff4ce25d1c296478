"""Staged execution of a tensor operation on argument slices
(talshTensorOpActivate / LoadInput / Execute / Test / StoreOutput /
Deactivate / Progress, TALSH/talshc.cpp:2708-3000): the D slice of a bigger
tensor accumulates alpha * (L slice) * (R slice) of bigger tensors, against
numpy in FP64; the stage rules (an out-of-order call is TALSH_NOT_ALLOWED,
Execute needs a device, Deactivate of a resourced operation returns it to
DEFINED) and the progress state machine."""
import ctypes as C

import numpy as np
import pytest

import tensop_ctypes as TC
from exatensor_amd import talsh as T
from gpu_util import engine, gpu_tensor, rel_err

pytestmark = pytest.mark.gpu

P = C.POINTER
NOT_ALLOWED = 1000007  # TALSH_NOT_ALLOWED


def _lib():
    """A second handle on the engine library, with the tensor-operation
    prototypes over the ctypes layouts of tensop_ctypes (one engine state)."""
    lib = C.CDLL(T.LIB_PATH)
    lib.talshTensorOpClean.argtypes = [P(TC.TensOp)]
    lib.talshTensorOpSetArgument.argtypes = [P(TC.TensOp), P(TC.Tens), P(C.c_size_t), P(C.c_int)]
    lib.talshTensorOpSpecify.argtypes = [P(TC.TensOp), C.c_int, C.c_int, C.c_char_p, C.c_double, C.c_double]
    lib.talshTensorOpSetExecDevice.argtypes = [P(TC.TensOp), C.c_int, C.c_int]
    for f in ("talshTensorOpActivate", "talshTensorOpLoadInput", "talshTensorOpStoreOutput",
              "talshTensorOpDeactivate", "talshTensorOpDestruct"):
        getattr(lib, f).argtypes = [P(TC.TensOp)]
    lib.talshTensorOpExecute.argtypes = [P(TC.TensOp), C.c_int, C.c_int]
    lib.talshTensorOpTest.argtypes = [P(TC.TensOp), P(C.c_int), C.c_int]
    lib.talshTensorOpProgress.argtypes = [P(TC.TensOp), P(C.c_int)]
    return lib


def _as_tens(t):
    return C.cast(C.pointer(t.t), P(TC.Tens))


def _op(lib, kind, pattern, alpha, args, dev=None):
    """args: [(engine tensor, offsets, dims)] for D, L, R."""
    op = TC.TensOp()
    assert lib.talshTensorOpClean(C.byref(op)) == 0
    keep = []
    for t, offs, dims in args:
        o = (C.c_size_t * len(offs))(*offs)
        d = (C.c_int * len(dims))(*dims)
        keep += [o, d]
        assert lib.talshTensorOpSetArgument(C.byref(op), _as_tens(t), o, d) == 0
    pat = C.c_char_p(pattern.encode())
    keep.append(pat)
    assert lib.talshTensorOpSpecify(C.byref(op), TC.TENSOR_CONTRACT, kind, pat, alpha.real, alpha.imag) == 0
    if dev is not None:
        assert lib.talshTensorOpSetExecDevice(C.byref(op), dev, T.DEV_DEFAULT) == 0
    op._keep = keep
    return op


def _rand(rng, shape, kind):
    x = rng.uniform(-1, 1, shape)
    if kind in (T.C4, T.C8):
        x = x + 1j * rng.uniform(-1, 1, shape)
    return x.astype(T.NP_DTYPE[kind])


@pytest.mark.parametrize("kind", [T.R8, T.C8, T.R4])
def test_progress_runs_a_contraction_on_slices(kind):
    eng = engine()
    lib = _lib()
    rng = np.random.default_rng(kind)
    # D(a,b,c) += L(c,k,a) * R(k,b): slices of bigger tensors
    Db, Lb, Rb = _rand(rng, (9, 8, 7), kind), _rand(rng, (6, 11, 10), kind), _rand(rng, (11, 12), kind)
    td, tl, tr = (gpu_tensor(x.ravel(order="F"), list(x.shape)) for x in (Db, Lb, Rb))
    alpha = 0.5 - 0.25j if kind in (T.C4, T.C8) else 0.75
    a, b, c, k = 4, 5, 3, 6
    od, ol, orr = (2, 1, 3), (1, 2, 4), (2, 6)
    op = _op(lib, kind, "D(a,b,c)+=L(c,k,a)*R(k,b)", alpha,
             [(td, od, (a, b, c)), (tl, ol, (c, k, a)), (tr, orr, (k, b))], dev=eng.gpu(0))
    done = C.c_int(0)
    for _ in range(100000):
        assert lib.talshTensorOpProgress(C.byref(op), C.byref(done)) == 0
        if done.value == T.YEP:
            break
    assert done.value == T.YEP and op.stage == 8  # TALSH_OP_RETIRED
    assert op.time_finished >= op.time_completed >= op.time_scheduled >= op.time_started > 0
    got = td.download().reshape(Db.shape, order="F")
    want = Db.astype(np.complex128)
    Ls = Lb[ol[0]:ol[0] + c, ol[1]:ol[1] + k, ol[2]:ol[2] + a].astype(np.complex128)
    Rs = Rb[orr[0]:orr[0] + k, orr[1]:orr[1] + b].astype(np.complex128)
    want[od[0]:od[0] + a, od[1]:od[1] + b, od[2]:od[2] + c] += alpha * np.einsum("cka,kb->abc", Ls, Rs)
    assert rel_err(got, want) <= (1e-5 if kind == T.R4 else 1e-13)
    # everything outside the D slice is untouched
    mask = np.ones(Db.shape, bool)
    mask[od[0]:od[0] + a, od[1]:od[1] + b, od[2]:od[2] + c] = False
    assert np.array_equal(got[mask], Db[mask])
    assert lib.talshTensorOpDestruct(C.byref(op)) == 0
    for t in (td, tl, tr):
        t.destruct()


def test_stage_rules_and_manual_sequence():
    eng = engine()
    lib = _lib()
    rng = np.random.default_rng(5)
    Db, Lb, Rb = rng.uniform(-1, 1, (10, 10)), rng.uniform(-1, 1, (10, 10)), rng.uniform(-1, 1, (10, 10))
    td, tl, tr = (gpu_tensor(x.ravel(order="F"), [10, 10]) for x in (Db, Lb, Rb))
    args = [(td, (1, 2), (4, 5)), (tl, (3, 0), (4, 6)), (tr, (4, 2), (6, 5))]
    op = _op(lib, T.R8, "D(a,b)+=L(a,k)*R(k,b)", 2.0, args)
    ref = C.byref(op)
    done = C.c_int(0)
    # out of order
    assert lib.talshTensorOpLoadInput(ref) == NOT_ALLOWED
    assert lib.talshTensorOpExecute(ref, eng.gpu(0), T.DEV_DEFAULT) == NOT_ALLOWED
    assert lib.talshTensorOpTest(ref, C.byref(done), T.NOPE) == NOT_ALLOWED
    assert lib.talshTensorOpStoreOutput(ref) == NOT_ALLOWED
    assert lib.talshTensorOpDeactivate(ref) == NOT_ALLOWED
    # activate / deactivate returns the operation to DEFINED
    assert lib.talshTensorOpActivate(ref) == 0 and op.stage == 3
    assert lib.talshTensorOpDeactivate(ref) == 0 and op.stage == 2
    # manual sequence
    assert lib.talshTensorOpActivate(ref) == 0
    assert lib.talshTensorOpLoadInput(ref) == 0 and op.stage == 4
    # no device preset and none given: refused; with a device: scheduled
    assert lib.talshTensorOpExecute(ref, T.DEV_DEFAULT, T.DEV_DEFAULT) == T.TALSH_INVALID_ARGS
    assert lib.talshTensorOpExecute(ref, eng.gpu(0), T.DEV_DEFAULT) == 0 and op.stage == 5
    assert op.exec_dev_id == eng.gpu(0)
    completed = C.c_int(0)
    assert lib.talshTensorOpTest(ref, C.byref(completed), T.YEP) == 0
    assert completed.value == T.YEP and op.stage == 6
    assert lib.talshTensorOpStoreOutput(ref) == 0 and op.stage == 7
    assert lib.talshTensorOpDeactivate(ref) == 0 and op.stage == 8
    assert lib.talshTensorOpProgress(ref, C.byref(done)) == 0 and done.value == T.YEP
    want = Db.copy()
    want[1:5, 2:7] += 2.0 * Lb[3:7, 0:6] @ Rb[4:10, 2:7]
    got = td.download().reshape(10, 10, order="F")
    assert rel_err(got, want) <= 1e-13
    assert lib.talshTensorOpDestruct(ref) == 0
    for t in (td, tl, tr):
        t.destruct()
