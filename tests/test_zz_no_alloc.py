"""The allocation-free matrix product (talshAmdGemmNoAlloc) against
talshAmdGemm with a caller-provided workspace: the same kernels run on the
same operands, so the outputs must be identical bit for bit. Two calls in a
row share the GPU's scratch array, the second on a smaller product, for the
three-way BF16 split (R4 and C4 realified), a split-K real product and a
realified native complex product; a product whose scratch need exceeds the
array is refused before anything runs."""
import numpy as np
import pytest

from exatensor_amd import talsh as T
from gpu_util import engine, gpu_tensor, rel_err

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _operands(kind, m, n, k, seed):
    rng = np.random.default_rng(seed)
    if kind in (T.C4, T.C8):
        dt = np.complex64 if kind == T.C4 else np.complex128
        mk = lambda size: (rng.uniform(-1, 1, size) + 1j * rng.uniform(-1, 1, size)).astype(dt)
    else:
        dt = np.float32 if kind == T.R4 else np.float64
        mk = lambda size: rng.uniform(-1, 1, size).astype(dt)
    return mk(m * k), mk(n * k), mk(m * n)


def _both(kind, fast, m, n, k, ak, bk, beta_zero, seed, alpha=(0.75, 0.0)):
    """(no-alloc output, workspace output, float64/complex128 expectation)."""
    eng = engine()
    L = eng.lib
    a, b, c0 = _operands(kind, m, n, k, seed)
    ta, tb = gpu_tensor(a, [a.size]), gpu_tensor(b, [b.size])
    tc1, tc2 = gpu_tensor(c0, [c0.size]), gpu_tensor(c0, [c0.size])
    ptr = lambda t: t.body(eng.gpu(0))[1]
    wsb = L.talshAmdGemmWorkspace(kind, fast, m, n, k)
    assert 0 < wsb <= SCRATCH, wsb
    tw = T.Tensor(L, T.R4, [(wsb + 3) // 4], eng.gpu(0), 0.0)
    bz = 1 if beta_zero else 0
    e1 = L.talshAmdGemmNoAlloc(kind, fast, m, n, k, ptr(ta), ak, ptr(tb), bk, ptr(tc1), alpha[0], alpha[1], bz, -1)
    e2 = L.talshAmdGemm(kind, fast, m, n, k, ptr(ta), ak, ptr(tb), bk, ptr(tc2), alpha[0], alpha[1], bz, -1,
                        ptr(tw), wsb)
    assert (e1, e2) == (0, 0), (e1, e2)
    got1, got2 = tc1.download(), tc2.download()
    for t in (ta, tb, tc1, tc2, tw):
        t.destruct()
    wide = np.complex128 if np.iscomplexobj(a) else np.float64
    A = a.astype(wide).reshape(m, k) if ak else a.astype(wide).reshape(k, m).T
    B = b.astype(wide).reshape(n, k) if bk else b.astype(wide).reshape(k, n).T
    al = complex(*alpha) if np.iscomplexobj(a) else alpha[0]
    want = al * (A @ B.T)
    if not beta_zero:
        want = want + c0.astype(wide).reshape(n, m).T
    return got1, got2, want.T.reshape(-1)


# (kind, fast math, first product, second smaller product, tolerance against float64)
CASES = [
    (T.R4, 1, (700, 500, 3000, 1, 0, False), (130, 70, 300, 0, 1, True), 1e-6),
    (T.C4, 1, (300, 260, 2000, 1, 1, True), (64, 40, 100, 1, 0, False), 1e-6),
    (T.R8, 0, (256, 192, 5000, 1, 1, False), (64, 64, 600, 0, 0, True), 1e-13),
    (T.R4, 0, (1024, 1024, 1024, 1, 1, True), (37, 5, 4099, 0, 0, False), 1e-5),
    (T.C8, 0, (2048, 2100, 600, 1, 0, True), (2048, 2048, 512, 0, 1, False), 1e-13),
]


@pytest.mark.parametrize("kind,fast,first,second,tol", CASES)
def test_no_alloc_matches_workspace_entry(kind, fast, first, second, tol):
    alpha = (0.75, -0.4) if kind in (T.C4, T.C8) else (0.75, 0.0)
    for i, (m, n, k, ak, bk, bz) in enumerate((first, second)):
        got1, got2, want = _both(kind, fast, m, n, k, ak, bk, bz, seed=17 + i, alpha=alpha)
        assert got1.tobytes() == got2.tobytes(), (kind, fast, m, n, k)
        assert rel_err(got1, want) <= tol, (kind, fast, m, n, k, rel_err(got1, want))


def test_no_alloc_refuses_products_beyond_its_scratch():
    eng = engine()
    L = eng.lib
    m = n = k = 4096  # three-way split planes alone: 6 * 8192 * 4096 bytes
    assert L.talshAmdGemmWorkspace(T.R4, 1, m, n, k) > SCRATCH
    c = gpu_tensor(np.full(16, 3.0, np.float32), [16])
    p = c.body(eng.gpu(0))[1]
    # refused before any launch: the (too small) operand pointers are never read
    e = L.talshAmdGemmNoAlloc(T.R4, 1, m, n, k, p, 1, p, 1, p, 1.0, 0.0, 1, -1)
    assert e == T.TALSH_LIMIT_EXCEEDED, e
    assert np.all(c.download() == 3.0)
    c.destruct()
