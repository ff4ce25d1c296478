"""The task scheduler and device arena through the talsh.h C-ABI, as a
block-level worker drives them (INTERVIRT/tavp_worker.F90:9440-9480): many
asynchronous contraction tasks in flight over the engine's streams, chained
on shared output blocks, retried on TALSH_IN_PROGRESS / TRY_LATER; the
in-flight task contract (TALSH/talshc.cpp:2210); and the arena's back
pressure (TALSH/mem_manager.cpp:635) in a child process with a small arena."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from exatensor_amd import blocksparse as BS
from exatensor_amd import talsh as T
from gpu_util import engine, gpu_tensor, rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_small_arena_back_pressure_contract():
    """80 MiB arena: bodies until TRY_LATER (the refused tensor stays
    empty), DEVICE_UNABLE beyond the arena, fragmentation refused and
    coalescing after a release, and a contraction whose temporaries do not
    fit returning TRY_LATER with D, the task handle and the arena unchanged,
    then succeeding after memory is released."""
    print("child: tests/gpu_small_arena.py", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "gpu_small_arena.py")], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "SMALL_ARENA_OK" in r.stdout


def _finish(lib, tk):
    st = C.c_int()
    assert lib.talshTaskWait(C.byref(tk), C.byref(st)) == 0 and st.value == T.TALSH_TASK_COMPLETED
    assert lib.talshTaskDestruct(C.byref(tk)) == 0


def test_chained_async_block_tasks(oracle):
    """More than 300 asynchronous talshTensorContract tasks (one per block
    contraction, each with its own transposes and product), up to 64 in
    flight over the engine's streams. Blocks that accumulate into the same D
    block are chained: a task whose D is still held by an earlier task gets
    TALSH_IN_PROGRESS (or TRY_LATER), the oldest task in flight is retired
    and the call issued again, as a block-level worker does. Every D block is
    checked against FP64 sums over its chain (R8, <= 1e-12)."""
    eng = engine()
    lib = eng.lib
    ext, ops = BS.generate(seed=17, segments=5, fill=0.35, max_ops=420, extents=(3, 5, 8, 13, 16, 24))
    assert len(ops) >= 300
    rng = np.random.default_rng(17)
    host, tens = {}, {}

    def get(tag, key, modes):
        if (tag, key) not in tens:
            dims = BS.dims(ext, key, modes)
            arr = rng.uniform(-1, 1, int(np.prod(dims)))
            host[(tag, key)] = (arr, dims)
            tens[(tag, key)] = gpu_tensor(arr, dims)
        return tens[(tag, key)]

    triples = [(get("D", o[0], "abcd"), get("L", o[1], "akcl"), get("R", o[2], "lbkd")) for o in ops]
    chains = {}
    for dk, _, _ in ops:
        chains[dk] = chains.get(dk, 0) + 1
    assert max(chains.values()) >= 3
    pat = BS.PATTERN.encode()
    window, inflight, retries = 64, [], 0
    for i, (D, L, R) in enumerate(triples):
        tk = T.TalshTask()
        lib.talshTaskClean(C.byref(tk))
        while True:
            e = lib.talshTensorContract(pat, D.ptr, L.ptr, R.ptr, 0.5, 0.0, eng.gpu(0), T.DEV_NULL, T.COPY_MTT,
                                        T.YEP, C.byref(tk))
            if e in (T.TALSH_IN_PROGRESS, T.TRY_LATER):
                assert inflight, (i, e)
                assert lib.talshTaskIsEmpty(C.byref(tk)) == T.YEP
                _finish(lib, inflight.pop(0))
                retries += 1
                continue
            break
        assert e == 0, (i, e)
        inflight.append(tk)
        if len(inflight) >= window:
            _finish(lib, inflight.pop(0))
    for tk in inflight:
        _finish(lib, tk)
    assert retries > 0  # chained blocks were met while their D was in flight
    exp = {k: v[0].copy() for k, v in host.items() if k[0] == "D"}
    for dk, lk, rk in ops:
        dd = BS.dims(ext, dk, "abcd")
        l, ld = host[("L", lk)]
        r, rd = host[("R", rk)]
        assert oracle.contract(BS.PATTERN, exp[("D", dk)], dd, l, ld, r, rd, 0.5, True) == 0
    for key, val in exp.items():
        assert rel_err(tens[key].download(), val) <= 1e-12, key
    for t in tens.values():
        t.destruct()


def test_in_flight_task_contract():
    """A task still running cannot be destructed (TALSH_IN_PROGRESS, the task
    stays valid), its output tensor cannot be used by another operation
    (TALSH_IN_PROGRESS, nothing enqueued), and after the wait both succeed."""
    eng = engine()
    lib = eng.lib
    n = 8192  # 1.1e12 flop: tens of milliseconds on the device
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, n * n)
    td, tl, tr = gpu_tensor(np.zeros(n * n), [n, n]), gpu_tensor(a, [n, n]), gpu_tensor(a, [n, n])
    t2 = gpu_tensor(np.zeros(n * n), [n, n])
    tk = T.TalshTask()
    lib.talshTaskClean(C.byref(tk))
    assert lib.talshTensorContract(b"D(a,b)+=L(a,i)*R(i,b)", td.ptr, tl.ptr, tr.ptr, 1.0, 0.0, eng.gpu(0), T.DEV_NULL,
                                   T.COPY_MTT, T.YEP, C.byref(tk)) == 0
    assert lib.talshTaskDestruct(C.byref(tk)) == T.TALSH_IN_PROGRESS
    assert lib.talshTaskStatus(C.byref(tk)) in (T.TALSH_TASK_SCHEDULED, 2000002, 2000003, 2000004)
    assert lib.talshTensorCopy(b"D(a,b)=L(a,b)", t2.ptr, td.ptr, eng.gpu(0), T.DEV_NULL, T.COPY_MT,
                               None) == T.TALSH_IN_PROGRESS
    _finish(lib, tk)
    assert lib.talshTensorCopy(b"D(a,b)=L(a,b)", t2.ptr, td.ptr, eng.gpu(0), T.DEV_NULL, T.COPY_MT, None) == 0
    A = a.reshape(n, n, order="F")
    probe = rng.integers(0, n, (64, 2))
    got = t2.download().reshape(n, n, order="F")
    want = np.array([A[i, :] @ A[:, j] for i, j in probe])
    assert rel_err(got[probe[:, 0], probe[:, 1]], want) <= 1e-12
    for t in (td, tl, tr, t2):
        t.destruct()
