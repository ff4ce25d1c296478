"""DLF permutation (talshTensorCopy) on the GPU: bit-exact against the oracle
and against numpy's own transpose, from tiny random cases to the configured
rank-6 extent-24 and rank-4 extent-128 tensors."""
import itertools
import random

import numpy as np
import pytest

from exatensor_amd import talsh as T
from gpu_util import engine, gpu_tensor

pytestmark = pytest.mark.gpu


def _pattern(perm):
    """Copy pattern for O2N perm (input dim i -> output position perm[i])."""
    n = len(perm)
    lab = [f"i{k}" for k in range(n)]
    dlab = [None] * n
    for k in range(n):
        dlab[perm[k]] = lab[k]
    return "D(" + ",".join(dlab) + ")=L(" + ",".join(lab) + ")"


def _gpu_copy(arr, dims, perm, conj=False):
    eng = engine()
    L = eng.lib
    ddims = [0] * len(dims)
    for i, p in enumerate(perm):
        ddims[p] = dims[i]
    tl = gpu_tensor(arr, dims)
    td = T.Tensor(L, T.KIND_OF[arr.dtype], ddims, eng.gpu(0))
    assert td.err == 0
    pat = _pattern(perm)
    if conj:
        pat = pat.replace("=L(", "=L+(")
    e = L.talshTensorCopy(pat.encode(), td.ptr, tl.ptr, eng.gpu(0), T.DEV_NULL, T.COPY_MT, None)
    assert e == 0, (pat, e)
    out = td.download()
    td.destruct(); tl.destruct()
    return out


def _numpy_perm(arr, dims, perm):
    a = arr.reshape(dims, order="F")
    n2o = [0] * len(perm)
    for i, p in enumerate(perm):
        n2o[p] = i
    return np.ascontiguousarray(a.transpose(n2o)).ravel(order="F") if False else a.transpose(n2o).ravel(order="F")


@pytest.mark.parametrize("kind", [T.R4, T.R8, T.C4, T.C8])
def test_random_small_permutations_bit_exact(oracle, kind):
    rng = random.Random(kind)
    nrs = np.random.default_rng(kind)
    for case in range(40):
        rank = rng.randint(1, 7)
        dims = [rng.choice([1, 2, 3, 5, 7, 8, 13, 16, 24, 31, 33, 64]) for _ in range(rank)]
        while int(np.prod(dims)) > 3_000_000:
            dims[dims.index(max(dims))] //= 2
        perm = list(range(rank))
        rng.shuffle(perm)
        n = int(np.prod(dims))
        x = nrs.uniform(-1, 1, n) + (1j * nrs.uniform(-1, 1, n) if kind in (T.C4, T.C8) else 0)
        x = x.astype(T.NP_DTYPE[kind])
        conj = kind in (T.C4, T.C8) and case % 3 == 0
        got = _gpu_copy(x, dims, perm, conj)
        exp = oracle.permute(dims, [p + 1 for p in perm], x, 1 if conj else 0)
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), (dims, perm, conj)


def test_all_rank4_permutations_mixed_extents(oracle):
    dims = [5, 12, 7, 9]
    x = np.arange(int(np.prod(dims)), dtype=np.float64)
    for perm in itertools.permutations(range(4)):
        got = _gpu_copy(x, dims, list(perm))
        assert np.array_equal(got, oracle.permute(dims, [p + 1 for p in perm], x)), perm


def test_ragged_tiles_large_odd_extents(oracle):
    # extents with no convenient divisors exercise masked edge tiles
    for dims, perm in [([1031, 777], [1, 0]), ([97, 89, 83], [2, 0, 1]), ([61, 3, 67, 5], [3, 1, 0, 2])]:
        x = np.random.default_rng(1).uniform(-1, 1, int(np.prod(dims))).astype(np.float32)
        got = _gpu_copy(x, dims, perm)
        assert np.array_equal(got, oracle.permute(dims, [p + 1 for p in perm], x)), (dims, perm)


@pytest.mark.parametrize("rank,ext", [(6, 24), (4, 128)])
@pytest.mark.parametrize("kind", [T.R4, T.R8])
def test_configured_shapes_bit_exact(oracle, rank, ext, kind):
    """The configured transposes at full size: a few seeded non-identity
    permutations checked bit for bit against numpy's transpose (an independent
    implementation); the integer-valued R8 variant is covered by R8 iota data."""
    dims = [ext] * rank
    n = ext ** rank
    rng = random.Random(100 * rank + kind)
    if kind == T.R8:
        x = np.arange(n, dtype=np.float64)  # integer valued, every element distinct
    else:
        x = oracle.uniform(T.R4, [n], 77)
    for _ in range(2):
        perm = list(range(rank))
        while perm == sorted(perm):
            rng.shuffle(perm)
        got = _gpu_copy(x, dims, perm)
        assert np.array_equal(got, _numpy_perm(x, dims, perm)), perm


@pytest.mark.parametrize("kind", [T.R4, T.R8, T.C4, T.C8])
@pytest.mark.parametrize("n", [1, 3, 64, 1000, 4099, 262147])
def test_identity_copies_bit_exact(oracle, kind, n):
    """Identity permutations take the flat copy (16-byte bodies: one-wave
    workgroups, a partial last workgroup when the vector count is not a
    multiple of 64) or the element copy (other sizes); conjugated complex
    copies flip the imaginary signs."""
    nrs = np.random.default_rng(n + kind)
    x = nrs.uniform(-1, 1, n) + (1j * nrs.uniform(-1, 1, n) if kind in (T.C4, T.C8) else 0)
    x = x.astype(T.NP_DTYPE[kind])
    for conj in ((False, True) if kind in (T.C4, T.C8) else (False,)):
        got = _gpu_copy(x, [n], [0], conj)
        exp = oracle.permute([n], [1], x, 1 if conj else 0)
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), (n, conj)


@pytest.mark.parametrize("kind", [T.R8, T.C8])
def test_high_rank_permutations_bit_exact(oracle, kind):
    """Ranks 8 to 56 (the maximum) with extents 1-5: shapes whose fused rank
    stays within the tiled kernel's 16 grid dimensions take the tiled path,
    the others the per-element path (both are hit, checked on the host
    plan); reversed orders defeat index fusion. Bit-exact against the
    oracle, conjugated for every other complex case."""
    import math
    eng = engine()
    rng = random.Random(56 + kind)
    nrs = np.random.default_rng(kind)
    paths = set()
    for case in range(24):
        rank = 56 if case < 2 else rng.randint(8, 40)
        if rank == 56:
            dims = [2] * 18 + [1] * 38
            rng.shuffle(dims)
        else:
            dims = [rng.choice((1, 2, 2, 2, 3, 3, 4, 5)) for _ in range(rank)]
        while math.prod(dims) > 1 << 20:
            i = max(range(rank), key=lambda j: dims[j])
            dims[i] = max(1, dims[i] // 2)
        perm = list(range(rank))
        rng.shuffle(perm)
        if case % 4 == 1:
            perm = perm[::-1]
        info = T.ints([0] * 7)
        assert eng.lib.talshAmdPermutePlan(kind, rank, T.ints(dims), T.ints(perm), info) == 0
        paths.add(info[1])
        n = math.prod(dims)
        if kind == T.R8:
            x = np.arange(n, dtype=np.float64)
        else:
            x = (nrs.uniform(-1, 1, n) + 1j * nrs.uniform(-1, 1, n)).astype(np.complex128)
        conj = kind == T.C8 and case % 2 == 0
        got = _gpu_copy(x, dims, perm, conj)
        exp = oracle.permute(dims, [p + 1 for p in perm], x, 1 if conj else 0)
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), (dims, perm, conj)
    assert {1, 2} <= paths, paths


def test_wide_offsets_and_unfusable_runs():
    """A transpose of 4.3 billion R4 elements (17 GB each way): element
    offsets beyond 32 bits, and two leading extents whose fused extent would
    leave the int range, so they stay separate dimensions. Every element
    carries its input index (mod 2^32) as its bit pattern; 4M random output
    positions and both ends of the output are checked against the
    permutation's index formula."""
    dims, perm = [65536, 32769, 2], [1, 2, 0]  # output dims (2, 65536, 32769)
    n = 65536 * 32769 * 2
    assert n >= 1 << 32
    info = T.ints([0] * 7)
    assert engine().lib.talshAmdPermutePlan(T.R4, 3, T.ints(dims), T.ints(perm), info) == 0
    assert info[0] == 3 and info[1] == 1, list(info)  # unfused, tiled
    x = np.empty(n, dtype=np.uint32)
    step = 1 << 28
    for s in range(0, n, step):
        e = min(n, s + step)
        x[s:e] = np.arange(s, e, dtype=np.uint64).astype(np.uint32)
    got = _gpu_copy(x.view(np.float32), dims, perm).view(np.uint32)
    del x
    rng = np.random.default_rng(32)
    o = np.concatenate([rng.integers(0, n, 1 << 22, dtype=np.uint64),
                        np.arange(1 << 16, dtype=np.uint64), np.arange(n - (1 << 16), n, dtype=np.uint64)])
    j2, t = o % 2, o // 2
    j0, j1 = t % 65536, t // 65536
    want = ((j0 + 65536 * (j1 + 32769 * j2)) & 0xFFFFFFFF).astype(np.uint32)
    assert np.array_equal(got[o.astype(np.int64)], want)


@pytest.mark.parametrize("rank,ext", [(6, 24), (4, 128)])
@pytest.mark.parametrize("kind", [T.R4, T.R8])
def test_configured_workload_every_permutation(rank, ext, kind):
    """Every transpose of the benchmark workload at full size: the 64 seeded
    non-identity rank-6 permutations (extent 24) or all 23 non-identity
    rank-4 permutations (extent 128), in R4 and R8. Every element carries
    its own input index: integer-valued doubles in R8, the index as the bit
    pattern in R4. For each permutation 1M random output positions plus the
    first and last 64K are checked bit for bit against the permutation's
    index formula."""
    import bench
    from gpu_util import pinned
    eng = engine()
    L = eng.lib
    perms = bench.perms_rank6() if rank == 6 else bench.perms_rank4()
    assert len(perms) == (64 if rank == 6 else 23)
    dims = [ext] * rank
    n = ext ** rank
    if kind == T.R8:
        x = np.arange(n, dtype=np.float64)
    else:
        x = np.arange(n, dtype=np.uint32).view(np.float32)
    tl = gpu_tensor(x, dims)
    del x
    td = T.Tensor(L, kind, dims, eng.gpu(0))
    assert td.err == 0
    out = pinned(n * (8 if kind == T.R8 else 4), np.float64 if kind == T.R8 else np.float32)
    rng = np.random.default_rng(1000 * rank + kind)
    edge = 1 << 16
    for perm in perms:
        e = L.talshTensorCopy(bench.copy_pattern(perm), td.ptr, tl.ptr, eng.gpu(0), T.DEV_NULL, T.COPY_MT, None)
        assert e == 0, (perm, e)
        td.download(out)
        o = np.concatenate([rng.integers(0, n, 1 << 20, dtype=np.int64), np.arange(edge, dtype=np.int64),
                            np.arange(n - edge, n, dtype=np.int64)])
        # output multi-index j (dimension 0 fastest); input dimension k sits at output position perm[k]
        j = [(o // ext ** p) % ext for p in range(rank)]
        want = sum(j[perm[k]] * ext ** k for k in range(rank))
        got = out[o]
        if kind == T.R8:
            assert np.array_equal(got, want.astype(np.float64)), perm
        else:
            assert np.array_equal(got.view(np.uint32), want.astype(np.uint32)), perm
    td.destruct(); tl.destruct()
