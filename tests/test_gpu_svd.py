"""talshTensorDecomposeSVD (+ the L / R / LR absorbing forms) on the GPU,
against the mathematics of a (truncated) SVD: singular values equal numpy's
(descending), the left factor's columns and the right factor's rows are
orthonormal, L * S * R reproduces D at full rank and is the best rank-nv
approximation otherwise (Eckart-Young: the residual norm is the norm of the
dropped singular values), for R8, R4, C8 and C4, tall and wide
matricizations, and permuted index orders (TALSH/talshc.cpp:5650-6130)."""
import ctypes as C

import numpy as np
import pytest

from exatensor_amd import talsh as T
from gpu_util import engine, gpu_tensor

pytestmark = pytest.mark.gpu

# D(a,b,c,d) = L(c,i,a) * R(b,i,d): rows (L's open indices) c,a; columns b,d
PAT = "D(a,b,c,d)=L(c,i,a)*R(b,i,d)"


def _rand(rng, shape, kind):
    x = rng.uniform(-1, 1, shape)
    if kind in (T.C4, T.C8):
        x = x + 1j * rng.uniform(-1, 1, shape)
    return x.astype(T.NP_DTYPE[kind])


def _decompose(kind, a, b, c, d, nv, absorb="N", seed=0, form=None):
    eng = engine()
    lib = eng.lib
    rng = np.random.default_rng(seed)
    D = _rand(rng, (a, b, c, d), kind)
    td = gpu_tensor(D.ravel(order="F"), [a, b, c, d])
    tl = gpu_tensor(np.zeros(c * nv * a, T.NP_DTYPE[kind]), [c, nv, a])
    tr = gpu_tensor(np.zeros(b * nv * d, T.NP_DTYPE[kind]), [b, nv, d])
    ts = gpu_tensor(np.zeros(nv, T.NP_DTYPE[kind]), [nv])
    if form is None:
        e = lib.talshTensorDecomposeSVD(PAT.encode(), td.ptr, tl.ptr, tr.ptr, ts.ptr, absorb.encode(), eng.gpu(0),
                                        T.DEV_NULL)
    else:
        e = getattr(lib, form)(PAT.encode(), td.ptr, tl.ptr, tr.ptr, eng.gpu(0), T.DEV_NULL)
    out = None
    if e == 0:
        Lf = tl.download().reshape((c, nv, a), order="F")
        Rf = tr.download().reshape((b, nv, d), order="F")
        S = ts.download()
        Dafter = td.download().reshape((a, b, c, d), order="F")
        out = (D, Lf, Rf, S, Dafter)
    for t in (td, tl, tr, ts):
        t.destruct()
    return e, out


def _mat(D):
    a, b, c, d = D.shape
    # rows (c, a) with c fastest, columns (b, d) with b fastest
    return np.transpose(D, (2, 0, 1, 3)).reshape((c * a, b * d), order="F")


TOL = {T.R8: 1e-12, T.C8: 1e-12, T.R4: 2e-5, T.C4: 2e-5}


@pytest.mark.parametrize("kind", [T.R8, T.R4, T.C8, T.C4])
@pytest.mark.parametrize("a,b,c,d", [(3, 4, 5, 2), (2, 6, 3, 5)])  # tall (15 x 8) and wide (6 x 30)
def test_full_rank_decomposition(kind, a, b, c, d):
    nv = min(c * a, b * d)
    e, out = _decompose(kind, a, b, c, d, nv, seed=a * 10 + kind)
    assert e == 0
    D, Lf, Rf, S, Dafter = out
    tol = TOL[kind]
    assert np.array_equal(Dafter, D)  # D is left unchanged
    A = _mat(D.astype(np.complex128))
    sv = np.linalg.svd(A, compute_uv=False)
    assert np.allclose(S.real, sv[:nv], rtol=tol * 10, atol=tol * sv[0])
    assert np.all(np.diff(S.real) <= 0)
    Lm = np.transpose(Lf, (0, 2, 1)).reshape((c * a, nv), order="F").astype(np.complex128)
    Rm = np.transpose(Rf, (1, 0, 2)).reshape((nv, b * d), order="F").astype(np.complex128)
    assert np.allclose(Lm.conj().T @ Lm, np.eye(nv), atol=tol * 10)
    assert np.allclose(Rm @ Rm.conj().T, np.eye(nv), atol=tol * 10)
    rec = np.einsum("cia,i,bid->abcd", Lf.astype(np.complex128), S.astype(np.complex128), Rf.astype(np.complex128))
    assert np.linalg.norm(rec - D) <= tol * 10 * np.linalg.norm(D)


@pytest.mark.parametrize("kind", [T.R8, T.C8])
def test_truncated_is_best_approximation(kind):
    a, b, c, d, nv = 4, 3, 5, 4, 5  # 20 x 12 matrix, rank 5 kept
    e, out = _decompose(kind, a, b, c, d, nv, seed=7)
    assert e == 0
    D, Lf, Rf, S, _ = out
    sv = np.linalg.svd(_mat(D.astype(np.complex128)), compute_uv=False)
    assert np.allclose(S.real, sv[:nv], rtol=1e-11)
    rec = np.einsum("cia,i,bid->abcd", Lf, S, Rf)
    resid = np.linalg.norm(rec - D)
    assert abs(resid - np.sqrt(np.sum(sv[nv:] ** 2))) <= 1e-10 * np.linalg.norm(D)


@pytest.mark.parametrize("form,absorb", [("talshTensorDecomposeSVDL", "L"), ("talshTensorDecomposeSVDR", "R"),
                                         ("talshTensorDecomposeSVDLR", "S"), (None, "L"), (None, "S")])
def test_absorbed_forms_reproduce_d(form, absorb):
    a, b, c, d = 3, 4, 4, 2
    nv = min(c * a, b * d)
    e, out = _decompose(T.R8, a, b, c, d, nv, absorb=absorb, seed=3, form=form)
    assert e == 0
    D, Lf, Rf, S, _ = out
    rec = np.einsum("cia,bid->abcd", Lf, Rf)
    assert np.linalg.norm(rec - D) <= 1e-12 * np.linalg.norm(D)
    sv = np.linalg.svd(_mat(D), compute_uv=False)
    lnorm = np.linalg.norm(np.transpose(Lf, (0, 2, 1)).reshape((c * a, nv), order="F"), axis=0)
    want = {"L": sv, "R": np.ones(nv), "S": np.sqrt(sv)}[absorb]
    assert np.allclose(lnorm, want, rtol=1e-11)


def test_empty_s_is_constructed_and_bad_rank_rejected():
    eng = engine()
    lib = eng.lib
    rng = np.random.default_rng(5)
    a, b, c, d = 2, 3, 3, 2
    D = rng.uniform(-1, 1, (a, b, c, d))
    td = gpu_tensor(D.ravel(order="F"), [a, b, c, d])
    tl = gpu_tensor(np.zeros(c * 6 * a), [c, 6, a])
    tr = gpu_tensor(np.zeros(b * 6 * d), [b, 6, d])
    s = T.TalshTens()
    assert lib.talshTensorClean(C.byref(s)) == 0
    assert lib.talshTensorDecomposeSVD(PAT.encode(), td.ptr, tl.ptr, tr.ptr, C.byref(s), b"N", eng.gpu(0),
                                       T.DEV_NULL) == 0
    assert lib.talshTensorRank(C.byref(s)) == 1 and lib.talshTensorVolume(C.byref(s)) == 6
    assert lib.talshTensorDestruct(C.byref(s)) == 0
    # 7 singular values from a 6 x 6 matricization: rejected
    tl7 = gpu_tensor(np.zeros(c * 7 * a), [c, 7, a])
    tr7 = gpu_tensor(np.zeros(b * 7 * d), [b, 7, d])
    assert lib.talshTensorClean(C.byref(s)) == 0
    assert lib.talshTensorDecomposeSVD(PAT.encode(), td.ptr, tl7.ptr, tr7.ptr, C.byref(s), b"N", eng.gpu(0),
                                       T.DEV_NULL) == T.TALSH_INVALID_ARGS
    for t in (td, tl, tr, tl7, tr7):
        t.destruct()


@pytest.mark.parametrize("kind,dims", [(T.R8, (3, 4, 5, 2)), (T.C8, (2, 6, 3, 5)), (T.R4, (4, 2, 3, 3))])
def test_orthogonalize_svd(kind, dims):
    eng = engine()
    lib = eng.lib
    a, b, c, d = dims
    rng = np.random.default_rng(17)
    D = _rand(rng, dims, kind)
    td = gpu_tensor(D.ravel(order="F"), list(dims))
    assert lib.talshTensorOrthogonalizeSVD(PAT.encode(), td.ptr, eng.gpu(0), T.DEV_NULL) == 0
    Q = _mat(td.download().reshape(dims, order="F").astype(np.complex128))
    td.destruct()
    A = _mat(D.astype(np.complex128))
    U, sv, Vh = np.linalg.svd(A, full_matrices=False)
    tol = TOL[kind] * 10
    assert np.linalg.norm(Q - U @ Vh) <= tol * np.linalg.norm(U @ Vh)  # the polar factor of D
    k = min(Q.shape)
    gram = Q.conj().T @ Q if Q.shape[0] >= Q.shape[1] else Q @ Q.conj().T
    assert np.allclose(gram, np.eye(k), atol=tol)
    assert lib.talshTensorOrthogonalizeMGS(td.ptr, 1, (C.c_int * 1)(0), eng.gpu(0), T.DEV_NULL) in (
        T.TALSH_NOT_IMPLEMENTED, T.TALSH_OBJECT_IS_EMPTY)
