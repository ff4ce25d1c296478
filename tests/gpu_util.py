"""Helpers shared by the GPU tests: one engine per test process, tensors made
from numpy arrays on GPU 0, the oracle for expected values."""
import ctypes as C

import numpy as np

from exatensor_amd import talsh as T

_ENG = None
_PINNED = {}


def engine():
    global _ENG
    if _ENG is None:
        _ENG = T.Engine(gpus=[0], arena_bytes=48 << 30)
    return _ENG


def gpu_tensor(arr, dims, kind=None, dev=None):
    eng = engine()
    kind = kind or T.KIND_OF[np.asarray(arr).dtype]
    t = T.Tensor(eng.lib, kind, dims, eng.gpu(0) if dev is None else dev)
    assert t.err == 0, t.err
    t.upload(arr)
    return t


def rel_err(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def contract(pattern, d, dd, l, ld, r, rd, alpha=1.0, acc=True, copy=T.COPY_MTT):
    eng = engine()
    L = eng.lib
    td, tl, tr = gpu_tensor(d, dd), gpu_tensor(l, ld), gpu_tensor(r, rd)
    a = complex(alpha)
    e = L.talshTensorContract(pattern.encode(), td.ptr, tl.ptr, tr.ptr, a.real, a.imag, eng.gpu(0), T.DEV_NULL,
                              copy, T.YEP if acc else T.NOPE, None)
    out = td.download() if e == 0 else None
    for t in (td, tl, tr):
        t.destruct()
    return e, out


def pinned(nbytes, dtype):
    """A numpy array over pinned host memory (hipHostMalloc), reused per size
    for the life of the process: device-to-host copies of multi-GB bodies run
    at the link's rate instead of through pageable staging."""
    if nbytes not in _PINNED:
        hip = C.CDLL("libamdhip64.so")
        hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        hip.hipHostMalloc.restype = C.c_int
        p = C.c_void_p()
        assert hip.hipHostMalloc(C.byref(p), nbytes, 0) == 0, "hipHostMalloc failed"
        _PINNED[nbytes] = (C.c_char * nbytes).from_address(p.value)
    return np.frombuffer(_PINNED[nbytes], dtype=dtype)
