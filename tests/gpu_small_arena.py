"""Run as a child process by tests/test_gpu_scheduler.py (needs its own engine
with an 80 MiB device arena): the arena's back-pressure contract through the
talsh.h C-ABI. Prints one line per check and "SMALL_ARENA_OK" at the end."""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from exatensor_amd import talsh as T  # noqa: E402

MiB = 1 << 20


def main():
    eng = T.Engine(gpus=[0], arena_bytes=80 * MiB)
    L = eng.lib
    g = eng.gpu(0)
    cap = L.talshDeviceBufferSize(0, T.DEV_AMD_GPU)
    assert cap == 80 * MiB, cap
    assert L.talshDeviceBufferFreeSize(0, T.DEV_AMD_GPU) == cap
    free = lambda: L.talshDeviceBufferFreeSize(0, T.DEV_AMD_GPU)
    largest = lambda: L.talshDeviceTensorSize(0, T.DEV_AMD_GPU)

    # 1. fill with 8 MiB bodies until TRY_LATER; the refused tensor stays empty
    blocks = []
    while True:
        t = T.Tensor(L, T.R8, [1024, 1024], g)
        if t.err:
            break
        blocks.append(t)
    assert t.err == T.TRY_LATER and L.talshTensorIsEmpty(t.ptr) == T.YEP, t.err
    assert len(blocks) == 10 and free() == 0
    big = T.Tensor(L, T.R8, [1024, 1024 * 11], g)
    assert big.err == T.DEVICE_UNABLE and L.talshTensorIsEmpty(big.ptr) == T.YEP
    print("fill: 10 bodies of 8 MiB, then TRY_LATER; larger than the arena: DEVICE_UNABLE", flush=True)

    # 2. fragmentation and coalescing
    for i in range(0, 10, 2):
        assert blocks[i].destruct() == 0
    assert free() == 40 * MiB and largest() == 8 * MiB
    t16 = T.Tensor(L, T.R8, [2048, 1024], g)
    assert t16.err == T.TRY_LATER and free() == 40 * MiB
    assert blocks[1].destruct() == 0  # joins the holes of blocks 0, 1, 2
    assert largest() == 24 * MiB and free() == 48 * MiB
    t16 = T.Tensor(L, T.R8, [2048, 1024], g)
    assert t16.err == 0
    assert t16.destruct() == 0
    for i in range(3, 10, 2):
        assert blocks[i].destruct() == 0
    assert free() == cap and largest() == cap
    print("fragmented free space refused (TRY_LATER), coalesced after a release", flush=True)

    # 3. a contraction whose temporaries do not fit: TRY_LATER with every
    # object unchanged, then the same call succeeds once memory is released
    rng = np.random.default_rng(8)
    dims = [32] * 4
    d0 = rng.uniform(-1, 1, 32 ** 4)
    l = rng.uniform(-1, 1, 32 ** 4)
    r = rng.uniform(-1, 1, 32 ** 4)
    td, tl, tr = (T.Tensor(L, T.R8, dims, g) for _ in range(3))
    for t, a in ((td, d0), (tl, l), (tr, r)):
        assert t.err == 0
        t.upload(a)
    fillers = []
    while True:
        f = T.Tensor(L, T.R8, [1024, 1024], g)
        if f.err:
            break
        fillers.append(f)
    assert len(fillers) == 7 and free() == 0
    pat = b"D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)"
    task = T.TalshTask()
    L.talshTaskClean(C.byref(task))
    e = L.talshTensorContract(pat, td.ptr, tl.ptr, tr.ptr, 0.5, 0.0, g, T.DEV_NULL, T.COPY_MTT, T.YEP, C.byref(task))
    assert e == T.TRY_LATER, e
    assert L.talshTaskIsEmpty(C.byref(task)) == T.YEP  # the handle can be used again
    assert free() == 0
    assert np.array_equal(td.download(), d0) and td.images() == [g]
    e = L.talshTensorContract(pat, td.ptr, tl.ptr, tr.ptr, 0.5, 0.0, g, T.DEV_NULL, T.COPY_MTT, T.YEP, None)
    assert e == T.TRY_LATER, e
    for f in fillers:
        assert f.destruct() == 0
    e = L.talshTensorContract(pat, td.ptr, tl.ptr, tr.ptr, 0.5, 0.0, g, T.DEV_NULL, T.COPY_MTT, T.YEP, C.byref(task))
    assert e == 0, e
    st = C.c_int()
    assert L.talshTaskWait(C.byref(task), C.byref(st)) == 0 and st.value == T.TALSH_TASK_COMPLETED
    assert L.talshTaskDestruct(C.byref(task)) == 0
    Lf, Rf, Df = (x.reshape(dims, order="F") for x in (l, r, d0))
    want = (Df + 0.5 * np.einsum("dibj,jcia->abcd", Lf, Rf)).ravel(order="F")
    got = td.download()
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel <= 1e-12, rel
    for t in (td, tl, tr):
        assert t.destruct() == 0
    assert free() == cap  # every temporary returned
    print(f"contraction: TRY_LATER with objects unchanged, then done after release (rel. error {rel:.2e})",
          flush=True)
    eng.shutdown()
    print("SMALL_ARENA_OK", flush=True)


if __name__ == "__main__":
    main()
