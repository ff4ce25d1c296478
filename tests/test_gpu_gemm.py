"""Matrix kernels on awkward shapes through talshAmdGemm: the three-way BF16
split (tile-blocked operands, split K) against FP64, ragged m/n/k, every
operand orientation, beta 0/1 and K chunk edges; the tile argument must not
change the three-way split's result. Real FP64/FP32 products with few output
tiles split K into FP64-summed partials; with and without that workspace they
agree with FP64."""
import numpy as np
import pytest

from exatensor_amd import talsh as T
from gpu_util import engine, gpu_tensor, rel_err

pytestmark = pytest.mark.gpu


def _gemm(m, n, k, ak, bk, beta_zero, tile, fast, kind=T.R4, seed=0, workspace=True):
    eng = engine()
    L = eng.lib
    rng = np.random.default_rng(seed)
    dt = np.float32 if kind == T.R4 else np.float64
    a = rng.uniform(-1, 1, m * k).astype(dt)
    b = rng.uniform(-1, 1, n * k).astype(dt)
    c0 = rng.uniform(-1, 1, m * n).astype(dt)
    A = a.reshape(m, k) if ak else a.reshape(k, m).T
    B = b.reshape(n, k) if bk else b.reshape(k, n).T
    alpha = 0.75
    want = alpha * (A.astype(np.float64) @ B.astype(np.float64).T)
    if not beta_zero:
        want += c0.reshape(n, m).T.astype(np.float64)
    ta, tb, tc = gpu_tensor(a, [a.size]), gpu_tensor(b, [b.size]), gpu_tensor(c0, [c0.size])
    wsb = L.talshAmdGemmWorkspace(kind, fast, m, n, k) if workspace else 0
    tw = T.Tensor(L, T.R4, [max(1, (wsb + 3) // 4)], eng.gpu(0), 0.0) if wsb else None
    ptr = lambda t: t.body(eng.gpu(0))[1] if t is not None else None
    e = L.talshAmdGemm(kind, fast, m, n, k, ptr(ta), ak, ptr(tb), bk, ptr(tc), alpha, 0.0, 1 if beta_zero else 0,
                       tile, ptr(tw), wsb)
    got = tc.download().reshape(n, m).T if e == 0 else None
    for t in (ta, tb, tc, tw):
        if t is not None:
            t.destruct()
    return e, got, want


@pytest.mark.parametrize("tile", [-1, 1])
@pytest.mark.parametrize("m,n,k,ak,bk,beta_zero", [
    (1, 1, 1, 1, 1, True),
    (300, 200, 70, 1, 1, True),
    (257, 129, 2100, 1, 1, False),
    (129, 300, 33, 0, 1, True),
    (200, 130, 4500, 1, 0, True),
    (513, 255, 2048, 0, 0, False),
    (1000, 700, 6200, 1, 1, True),
    (264, 136, 8500, 1, 1, False),
    (130, 70, 17000, 0, 1, True),
])
def test_bf16x3_shapes_against_fp64(m, n, k, ak, bk, beta_zero, tile):
    e, got, want = _gemm(m, n, k, ak, bk, beta_zero, tile, fast=1)
    assert e == 0
    assert rel_err(got, want) <= 1e-6, (m, n, k, tile)


@pytest.mark.parametrize("kind,tile", [(T.R4, 0), (T.R4, 2), (T.R8, 0), (T.R8, 1), (T.R8, 2)])
def test_native_kernels_ragged(kind, tile):
    e, got, want = _gemm(333, 177, 1001, 1, 0, False, tile, fast=0, kind=kind, seed=3)
    assert e == 0
    assert rel_err(got, want) <= (1e-5 if kind == T.R4 else 1e-13)


@pytest.mark.parametrize("kind", [T.R8, T.R4])
@pytest.mark.parametrize("m,n,k,ak,bk,beta_zero,tile", [
    (256, 192, 5000, 1, 1, True, 2),     # 12 tiles: K in 8 parts
    (1000, 700, 2100, 0, 1, False, 2),   # 176 tiles: 2 parts, accumulate
    (64, 64, 513, 1, 0, False, -1),      # one tile, default choice: 2 parts
    (1024, 1024, 1024, 1, 1, True, -1),  # the extent-32 rank-4 contraction's product: 2 parts
    (37, 5, 4099, 0, 0, True, 2),        # ragged everything
])
@pytest.mark.parametrize("workspace", [True, False])
def test_real_split_k(kind, m, n, k, ak, bk, beta_zero, tile, workspace):
    e, got, want = _gemm(m, n, k, ak, bk, beta_zero, tile, fast=0, kind=kind, seed=m + n, workspace=workspace)
    assert e == 0
    assert rel_err(got, want) <= (1e-5 if kind == T.R4 else 1e-13)


def test_split_k_workspace_is_reported():
    L = engine().lib
    # few output tiles, deep K: partials for up to 8 parts
    assert L.talshAmdGemmWorkspace(T.R8, 0, 256, 192, 5000) == 8 * 256 * 192 * 8 + 16
    assert L.talshAmdGemmWorkspace(T.R4, 0, 256, 192, 5000) == 8 * 256 * 192 * 4 + 16
    # many tiles or complex: none
    assert L.talshAmdGemmWorkspace(T.R8, 0, 9216, 9216, 16384) == 0
    assert L.talshAmdGemmWorkspace(T.C8, 0, 64, 64, 5000) == 0


def _gemm_c4(m, n, k, ak, bk, beta_zero, fast, seed=0, alpha=0.75 - 0.4j, kind=T.C4):
    eng = engine()
    L = eng.lib
    rng = np.random.default_rng(seed)
    cdt = np.complex64 if kind == T.C4 else np.complex128
    cx = lambda size: (rng.uniform(-1, 1, size) + 1j * rng.uniform(-1, 1, size)).astype(cdt)
    a, b, c0 = cx(m * k), cx(n * k), cx(m * n)
    A = a.reshape(m, k) if ak else a.reshape(k, m).T
    B = b.reshape(n, k) if bk else b.reshape(k, n).T
    want = alpha * (A.astype(np.complex128) @ B.astype(np.complex128).T)
    if not beta_zero:
        want += c0.reshape(n, m).T.astype(np.complex128)
    ta, tb, tc = gpu_tensor(a, [a.size]), gpu_tensor(b, [b.size]), gpu_tensor(c0, [c0.size])
    wsb = L.talshAmdGemmWorkspace(kind, fast, m, n, k)
    tw = T.Tensor(L, T.R4, [max(1, (wsb + 3) // 4)], eng.gpu(0), 0.0) if wsb else None
    ptr = lambda t: t.body(eng.gpu(0))[1] if t is not None else None
    e = L.talshAmdGemm(kind, fast, m, n, k, ptr(ta), ak, ptr(tb), bk, ptr(tc), alpha.real, alpha.imag,
                       1 if beta_zero else 0, -1, ptr(tw), wsb)
    got = tc.download().reshape(n, m).T if e == 0 else None
    for t in (ta, tb, tc, tw):
        if t is not None:
            t.destruct()
    return e, got, want, wsb


@pytest.mark.parametrize("m,n,k,ak,bk,beta_zero", [
    (1, 1, 1, 1, 1, True),
    (130, 70, 33, 1, 1, False),
    (129, 300, 4100, 0, 1, True),    # k' = 8200: two K chunks
    (257, 129, 2100, 1, 0, False),
    (200, 130, 5000, 0, 0, True),
])
def test_c4_three_way_split_against_fp64(m, n, k, ak, bk, beta_zero):
    """C4 under fast math: one realified real product (2m x 2k and n x 2k
    split planes) with the complex alpha applied in the FP64 reduction,
    every operand orientation, beta 0/1, against complex128 products of
    the same complex64 inputs: <= 1e-6 relative."""
    e, got, want, wsb = _gemm_c4(m, n, k, ak, bk, beta_zero, fast=1, seed=m + k)
    assert e == 0 and wsb > 0
    assert rel_err(got, want) <= 1e-6, (m, n, k)


def test_c4_native_unchanged_without_fast_math():
    e, got, want, wsb = _gemm_c4(257, 129, 2100, 1, 0, False, fast=0, seed=5)
    assert e == 0 and wsb == 0
    assert rel_err(got, want) <= 1e-5


def test_c4_workspace_is_reported():
    L = engine().lib
    m, n, k = 300, 200, 5000
    kp = -(-2 * k // 32) * 32
    mp = -(-2 * m // 256) * 256
    np_ = -(-n // 128) * 128
    nz = -(-kp // 8192)
    assert L.talshAmdGemmWorkspace(T.C4, 1, m, n, k) == 3 * (mp + np_) * kp * 2 + nz * 2 * m * n * 4 + 16


@pytest.mark.parametrize("kind", [T.C8, T.C4])
@pytest.mark.parametrize("m,n,k,ak,bk,beta_zero,alpha", [
    (2048, 2048, 600, 1, 1, True, 0.75),           # realified: 512 real 128x128 tiles
    (2100, 1950, 1000, 0, 1, False, 0.75 - 0.4j),  # ragged, complex alpha (applied to A), accumulate
    (2048, 2100, 777, 1, 0, False, -1.5),
    (2050, 2040, 513, 0, 0, True, 0.3 + 1.1j),
])
def test_complex_realified_native(kind, m, n, k, ak, bk, beta_zero, alpha):
    """Large complex products in native precision run realified (2m x 2k and
    n x 2k real operands, one real 128x128-tile product written straight into
    the interleaved complex output): every orientation, complex and real
    alpha, beta 0/1, against complex128 products; the workspace query reports
    the realified operands; small products keep the complex kernel (no
    workspace)."""
    L = engine().lib
    es = 16 if kind == T.C8 else 8
    assert L.talshAmdGemmWorkspace(kind, 0, m, n, k) >= (2 * m + n) * 2 * k * es // 2
    assert L.talshAmdGemmWorkspace(kind, 0, 200, 200, k) == 0
    e, got, want, wsb = _gemm_c4(m, n, k, ak, bk, beta_zero, fast=0, seed=m + k, alpha=complex(alpha), kind=kind)
    assert e == 0 and wsb > 0
    assert rel_err(got, want) <= (1e-13 if kind == T.C8 else 1e-5)
