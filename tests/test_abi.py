"""The drop-in boundary: the engine library exports every entry point its
headers declare, with the reference's struct layouts and constant values, and
its host-side index logic (pattern parser, matricization planner) matches the
reference's golden fixtures, and the transpose planner's tiles fit the kernel.
CPU only."""
import ctypes as C
import gzip
import json
import os
import re
import subprocess

import pytest

from exatensor_amd import talsh as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
REF_INC = "/root/reference/TALSH"


@pytest.fixture(scope="module")
def lib():
    return T.load_library()


def _declared_functions():
    names = set()
    for h in ("talsh.h", "tensor_algebra.h", "talsh_amd.h"):
        text = open(os.path.join(INC, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*\*?\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                             text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof", "typedef", "int"):
                names.add(name)
    return sorted(names)


def test_every_declared_symbol_is_exported(lib):
    names = _declared_functions()
    assert len(names) > 150
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


_PROBE = r'''
#include <cstdio>
#include <cstddef>
#include "talsh.h"
#define S(t) printf(#t " %zu\n", sizeof(t));
#define O(t, f) printf(#t "." #f " %zu\n", offsetof(t, f));
#define V(c) printf(#c " %lld\n", (long long)(c));
int main() {
  S(talsh_tens_t) S(talsh_task_t) S(talsh_tens_op_t) S(talsh_tens_slice_t) S(talshTensArg_t)
  S(talsh_tens_shape_t) S(talsh_tens_signature_t) S(talsh_tens_data_t) S(talsh_tens_dense_t) S(talsh_dev_rsc_t)
  O(talsh_task_t, tens_args) O(talsh_task_t, data_vol) O(talsh_task_t, exec_time)
  O(talsh_tens_op_t, symb_pattern) O(talsh_tens_op_t, tens_arg) O(talsh_tens_op_t, task_handle)
  O(talsh_tens_op_t, exec_dev_id) O(talsh_tens_op_t, time_finished) O(talsh_dev_rsc_t, mem_attached)
  V(TALSH_SUCCESS) V(TALSH_FAILURE) V(TALSH_NOT_AVAILABLE) V(TALSH_NOT_IMPLEMENTED) V(TALSH_NOT_INITIALIZED)
  V(TALSH_INVALID_ARGS) V(TALSH_OBJECT_IS_EMPTY) V(TALSH_IN_PROGRESS) V(TALSH_INVALID_REQUEST)
  V(TALSH_TASK_ERROR) V(TALSH_TASK_EMPTY) V(TALSH_TASK_COMPLETED) V(TALSH_TENSOR_CONTRACT) V(TALSH_OP_RETIRED)
  V(TALSH_MAX_DEV_PRESENT) V(MAX_TENSOR_RANK) V(MAX_TENSOR_OPERANDS) V(R4) V(R8) V(C4) V(C8) V(NO_TYPE)
  V(DEV_HOST) V(DEV_NVIDIA_GPU) V(DEV_INTEL_MIC) V(DEV_AMD_GPU) V(DEV_MAX) V(DEVICE_UNABLE) V(TRY_LATER)
  V(NOT_CLEAN) V(COPY_MTT) V(COPY_KKK) V(COPY_TT) V(COPY_K) V(YEP) V(NOPE)
  return 0;
}
'''


def _probe(inc, tmp_path, tag):
    src = tmp_path / f"probe_{tag}.cpp"
    src.write_text(_PROBE)
    exe = tmp_path / f"probe_{tag}"
    subprocess.run(["g++", "-std=c++11", "-w", "-I", inc, str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_struct_layouts_and_constants_match_reference(tmp_path):
    ours = _probe(INC, tmp_path, "ours")
    assert "talsh_tens_t 40" in ours
    if not os.path.isdir(REF_INC):
        pytest.skip("reference headers not present")
    assert ours == _probe(REF_INC, tmp_path, "ref")


def test_parser_matches_reference_fixtures(lib):
    rows = json.load(open(os.path.join(ROOT, "tests", "golden", "patterns.json")))["rows"]
    for row in rows:
        e, dig, dr, lr, rr, cj = T.str2dig(lib, row["pattern"])
        assert e == row["ierr"], row
        if e == 0:
            assert (dig, dr, lr, rr, cj) == (row["dig"], row["drank"], row["lrank"], row["rrank"], row["conj"]), row


def test_permutation_planner_matches_reference_fixtures(lib):
    with gzip.open(os.path.join(ROOT, "tests", "golden", "permutations.json.gz"), "rt") as f:
        rows = json.load(f)["rows"]
    for row in rows:
        d, l, r = (C.c_int * 64)(), (C.c_int * 64)(), (C.c_int * 64)()
        ncd, nlu, nru, nhu, ierr = (C.c_int() for _ in range(5))
        dig = T.ints(row["dig"])
        if row["hyper"]:
            lib.get_contraction_permutations(row["tl"], row["tr"], row["lrank"], row["rrank"], dig, row["conj"], d, l,
                                             r, C.byref(ncd), C.byref(nlu), C.byref(nru), C.byref(nhu), C.byref(ierr))
        else:
            lib.get_contr_permutations(row["tl"], row["tr"], row["lrank"], row["rrank"], dig, row["conj"], d, l, r,
                                       C.byref(ncd), C.byref(nlu), C.byref(nru), C.byref(ierr))
        assert ierr.value == row["ierr"], row
        if ierr.value:
            continue
        assert (ncd.value, nlu.value, nru.value) == (row["ncd"], row["nlu"], row["nru"]), row
        if row["hyper"]:
            assert nhu.value == row["nhu"], row
        assert list(l[1:row["lrank"] + 1]) == row["lprm"][1:], row
        assert list(r[1:row["rrank"] + 1]) == row["rprm"][1:], row
        assert list(d[1:len(row["dprm"])]) == row["dprm"][1:], row


def test_transpose_plans_fit_the_kernel(lib):
    """Every tiled transpose plan fits the kernel it is launched with: the LDS
    tile is at most 64 KiB before padding, a 512-thread workgroup covers it
    with at most 8 vectors per thread, and the vector width divides 16 bytes.
    On the rank-6 extent-24 and rank-4 extent-128 workloads every plan is
    tiled, uses 16-byte vectors, and both contiguous runs are whole 128-byte
    lines. Random shapes of ranks 2-40 and all four data kinds, CPU only."""
    import math
    import random
    es_of = {T.R4: 4, T.R8: 8, T.C4: 8, T.C8: 16}
    rng = random.Random(11)

    def plan(kind, dims, perm):
        info = T.ints([0] * 7)
        assert lib.talshAmdPermutePlan(kind, len(dims), T.ints(dims), T.ints(perm), info) == 0
        return list(info)

    for _ in range(600):
        kind = rng.choice(list(es_of))
        es = es_of[kind]
        rank = rng.randint(2, 8)
        dims = [rng.choice((1, 2, 3, 5, 7, 8, 16, 24, 31, 64, 100, 128, 257)) for _ in range(rank)]
        while math.prod(dims) > 1 << 28:
            dims[dims.index(max(dims))] //= 2
        perm = list(range(rank))
        rng.shuffle(perm)
        fused, path, tile, in_run, out_run, grid, vec = plan(kind, dims, perm)
        if path != 1:
            continue
        assert tile * es <= 65536, (dims, perm, tile)
        assert -(-tile // (512 * vec)) <= 8, (dims, perm, tile, vec)
        assert vec * es <= 16 and grid >= 1
    # ranks 8-40 of small extents: fused ranks above 16 take the per-element
    # path; tiled plans keep the same limits
    paths = set()
    for _ in range(400):
        kind = rng.choice(list(es_of))
        es = es_of[kind]
        rank = rng.randint(8, 40)
        dims = [rng.choice((1, 2, 2, 3, 3, 4, 5)) for _ in range(rank)]
        while math.prod(dims) > 1 << 22:
            dims[dims.index(max(dims))] //= 2
        perm = list(range(rank))
        rng.shuffle(perm)
        fused, path, tile, in_run, out_run, grid, vec = plan(kind, dims, perm)
        paths.add(path)
        assert path in (1, 2) and (fused <= 16 or path == 2), (dims, perm, fused, path)
        if path == 1:
            assert tile * es <= 65536 and -(-tile // (512 * vec)) <= 8 and vec * es <= 16, (dims, perm, tile)
    assert paths == {1, 2}
    import bench
    for kind in (T.R4, T.R8):
        es = es_of[kind]
        for dims, perms in (([24] * 6, bench.perms_rank6()), ([128] * 4, bench.perms_rank4())):
            for p in perms:
                fused, path, tile, in_run, out_run, grid, vec = plan(kind, dims, p)
                assert path == 1 and vec == 16 // es, (kind, p)
                assert (in_run * es) % 128 == 0 and (out_run * es) % 128 == 0, (kind, p, in_run, out_run)


@pytest.fixture(scope="module")
def host_engine(lib):
    eng = T.Engine(gpus=[])
    yield eng
    eng.shutdown()


def test_host_only_engine_semantics(lib, host_engine):
    t = T.Tensor(lib, T.R8, [3, 4, 5], 0, 2.5)
    assert t.err == 0
    assert lib.talshTensorRank(t.ptr) == 3 and lib.talshTensorVolume(t.ptr) == 60
    assert (t.download() == 2.5).all()
    # a second construct on a non-empty tensor is refused
    assert lib.talshTensorConstruct(t.ptr, T.R8, 1, T.ints([2]), 0, None, -1, None, 0.0, 0.0) == \
        T.TALSH_OBJECT_NOT_EMPTY
    # tensor operations execute on GPUs only
    assert lib.talshTensorInit(t.ptr, 1.0, 0.0, 0, T.DEV_HOST, T.COPY_M, None) == T.TALSH_NOT_AVAILABLE
    assert lib.talshTensorContract(b"D(a,b,c)+=L(a,b,c)*R()", t.ptr, t.ptr, t.ptr, 1.0, 0.0, T.DEV_DEFAULT,
                                   T.DEV_DEFAULT, T.COPY_MTT, T.YEP, None) == T.TALSH_NOT_AVAILABLE
    task = T.TalshTask()
    lib.talshTaskClean(C.byref(task))
    assert lib.talshTaskStatus(C.byref(task)) == T.TALSH_TASK_EMPTY
    assert t.destruct() == 0
    assert lib.talshTensorIsEmpty(t.ptr) == T.YEP


def test_device_id_codec(lib):
    assert lib.talshFlatDevId(T.DEV_HOST, 0) == 0
    assert lib.talshFlatDevId(T.DEV_NVIDIA_GPU, 0) == 1
    assert lib.talshFlatDevId(T.DEV_AMD_GPU, 0) == 17
    k = C.c_int()
    assert lib.talshKindDevId(18, C.byref(k)) == 1 and k.value == T.DEV_AMD_GPU


def test_split_k_workspace_rule(lib):
    """Real FP64/FP32 products split K only when the 64x64 output tiles are
    fewer than two per CU (256 CUs), into at most 8 parts of >= 256 k-values
    (16-aligned); the workspace holds one m x n partial per part. BF16x3 and
    complex workspaces follow their own rules. CPU only (the library's
    device query falls back to 256 CUs without a GPU)."""
    import random
    rng = random.Random(4)
    for _ in range(300):
        kind = rng.choice([T.R8, T.R4])
        es = 8 if kind == T.R8 else 4
        m, n = rng.randint(1, 3000), rng.randint(1, 3000)
        k = rng.randint(1, 40000)
        ws = lib.talshAmdGemmWorkspace(kind, 0, m, n, k)
        tiles = -(-m // 64) * -(-n // 64)
        if ws == 0:
            continue
        nz, rem = divmod(ws - 16, m * n * es)
        assert rem == 0 and 2 <= nz <= 8, (m, n, k, ws)
        assert tiles < 512 and k >= 512, (m, n, k, nz)
        chunk = -(-(-(-k // min(8, 512 // tiles, k // 256))) // 16) * 16
        assert -(-k // chunk) == nz and chunk >= 256, (m, n, k, nz, chunk)
    assert lib.talshAmdGemmWorkspace(T.R8, 0, 9216, 9216, 16384) == 0
    assert lib.talshAmdGemmWorkspace(T.C8, 0, 64, 64, 5000) == 0
