"""ctypes binding of the engine's talsh.h C-ABI (exatensor_amd/libtalsh.so).

Mirrors the reference API names (TALSH/talsh.h) one to one; Python objects only
wrap the caller-owned talsh_tens_t / talsh_task_t structs. There is no other
execution path: if the shared library is missing, loading fails loudly.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtalsh.so")

# ---- constants (talsh.h / tensor_algebra.h) ---------------------------------
TALSH_SUCCESS = 0
TALSH_FAILURE = -666
TALSH_NOT_AVAILABLE = -888
TALSH_NOT_IMPLEMENTED = -999
TALSH_NOT_INITIALIZED = 1000000
TALSH_ALREADY_INITIALIZED = 1000001
TALSH_INVALID_ARGS = 1000002
TALSH_OBJECT_NOT_EMPTY = 1000004
TALSH_OBJECT_IS_EMPTY = 1000005
TALSH_IN_PROGRESS = 1000006
TALSH_NOT_FOUND = 1000009
TALSH_LIMIT_EXCEEDED = 1000008
TALSH_NOT_ALLOWED = 1000007
TALSH_TASK_ERROR = 1999999
TALSH_TASK_EMPTY = 2000000
TALSH_TASK_SCHEDULED = 2000001
TALSH_TASK_COMPLETED = 2000005
NO_TYPE, R4, R8, C4, C8 = 0, 4, 8, 14, 18
DEV_NULL, DEV_DEFAULT, DEV_HOST, DEV_NVIDIA_GPU, DEV_INTEL_MIC, DEV_AMD_GPU = -1, -1, 0, 1, 2, 3
NOPE, YEP = 0, 1
DEVICE_UNABLE = -546372819
TRY_LATER = -918273645
NOT_CLEAN = -192837465
COPY_D, COPY_M, COPY_T, COPY_K = 0, 1, 2, 3
COPY_MT, COPY_TT, COPY_KK, COPY_MM = 6, 10, 15, 5
COPY_MTT, COPY_MMM, COPY_KKK, COPY_TTT, COPY_MKK = 26, 21, 63, 42, 31
FUSED_VARIANT = 1000  # plan variants of the fused-gather candidates start here
MAX_TENSOR_RANK = 56
MAX_TENSOR_OPERANDS = 4
NP_DTYPE = {R4: np.float32, R8: np.float64, C4: np.complex64, C8: np.complex128}
KIND_OF = {np.dtype(np.float32): R4, np.dtype(np.float64): R8, np.dtype(np.complex64): C4,
           np.dtype(np.complex128): C8}
KERNEL_TRANSPOSE, KERNEL_GEMM, KERNEL_GROUPED = 0, 1, 2


class TalshTens(C.Structure):
    """talsh_tens_t (layout of TALSH/talsh.h:87-94)."""
    _fields_ = [("shape_p", C.c_void_p), ("dev_rsc", C.c_void_p), ("data_kind", C.POINTER(C.c_int)),
                ("avail", C.POINTER(C.c_int)), ("dev_rsc_len", C.c_int), ("ndev", C.c_int)]


class TalshTensArg(C.Structure):
    _fields_ = [("tens_p", C.POINTER(TalshTens)), ("source_image", C.c_int)]


class TalshTask(C.Structure):
    """talsh_task_t (layout of TALSH/talsh.h:110-121)."""
    _fields_ = [("task_p", C.c_void_p), ("task_error", C.c_int), ("dev_kind", C.c_int), ("data_kind", C.c_int),
                ("coherence", C.c_int), ("num_args", C.c_int), ("tens_args", TalshTensArg * MAX_TENSOR_OPERANDS),
                ("data_vol", C.c_double), ("flops", C.c_double), ("exec_time", C.c_double)]


_IP = C.POINTER(C.c_int)
_FP = C.POINTER(C.c_float)
_TP = C.POINTER(TalshTens)
_KP = C.POINTER(TalshTask)
_lib = None


def load_library(path=LIB_PATH):
    """Loads libtalsh.so and declares the signatures used from Python."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"engine library not built: {path} (run __graft_entry__.build())")
    L = C.CDLL(path)
    sig = {
        "talshInit": (C.c_int, [C.POINTER(C.c_size_t), _IP, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP]),
        "talshShutdown": (C.c_int, []),
        "talshDeviceCount": (C.c_int, [C.c_int, _IP]),
        "talshFlatDevId": (C.c_int, [C.c_int, C.c_int]),
        "talshKindDevId": (C.c_int, [C.c_int, _IP]),
        "talshEnableFastMath": (C.c_int, [C.c_int, C.c_int]),
        "talshDisableFastMath": (C.c_int, [C.c_int, C.c_int]),
        "talshQueryFastMath": (C.c_int, [C.c_int, C.c_int]),
        "talshDeviceBufferSize": (C.c_size_t, [C.c_int, C.c_int]),
        "talshDeviceBufferFreeSize": (C.c_size_t, [C.c_int, C.c_int]),
        "talshDeviceTensorSize": (C.c_size_t, [C.c_int, C.c_int]),
        "talshTensorClean": (C.c_int, [_TP]),
        "talshTensorIsEmpty": (C.c_int, [_TP]),
        "talshTensorConstruct": (C.c_int, [_TP, C.c_int, C.c_int, _IP, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_double, C.c_double]),
        "talshTensorImportData": (C.c_int, [_TP, C.c_int, C.c_void_p]),
        "talshTensorDestruct": (C.c_int, [_TP]),
        "talshTensorRank": (C.c_int, [_TP]),
        "talshTensorVolume": (C.c_size_t, [_TP]),
        "talshTensorGetBodyAccess": (C.c_int, [_TP, C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int]),
        "talshTensorGetScalar": (C.c_int, [_TP, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "talshTensorPresence": (C.c_int, [_TP, _IP, _IP, _IP, C.c_int, C.c_int]),
        "talshTensorPlace": (C.c_int, [_TP, C.c_int, C.c_int, C.c_void_p, C.c_int, _KP]),
        "talshTensorDiscard": (C.c_int, [_TP, C.c_int, C.c_int]),
        "talshTensorInit": (C.c_int, [_TP, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _KP]),
        "talshTensorScale": (C.c_int, [_TP, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _KP]),
        "talshTensorCopy": (C.c_int, [C.c_char_p, _TP, _TP, C.c_int, C.c_int, C.c_int, _KP]),
        "talshTensorAdd": (C.c_int, [C.c_char_p, _TP, _TP, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _KP]),
        "talshTensorContract": (C.c_int, [C.c_char_p, _TP, _TP, _TP, C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_int, C.c_int, _KP]),
        "talshTaskClean": (C.c_int, [_KP]),
        "talshTaskStatus": (C.c_int, [_KP]),
        "talshTaskIsEmpty": (C.c_int, [_KP]),
        "talshTaskComplete": (C.c_int, [_KP, _IP, _IP]),
        "talshTaskWait": (C.c_int, [_KP, _IP]),
        "talshTaskDestruct": (C.c_int, [_KP]),
        "talshTaskTime": (C.c_int, [_KP, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "talsh_get_contr_ptrn_str2dig": (C.c_int, [C.c_char_p, _IP, _IP, _IP, _IP, _IP]),
        "get_contr_permutations": (None, [C.c_int, C.c_int, C.c_int, C.c_int, _IP, C.c_int, _IP, _IP, _IP, _IP,
                                          _IP, _IP, _IP]),
        "get_contraction_permutations": (None, [C.c_int, C.c_int, C.c_int, C.c_int, _IP, C.c_int, _IP, _IP, _IP,
                                                _IP, _IP, _IP, _IP, _IP]),
        "talshAmdProfile": (C.c_int, [C.c_int]),
        "talshAmdProfileQuery": (C.c_int, [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
        "talshAmdSetPlanVariant": (C.c_int, [C.c_int, C.c_int]),
        "talshAmdSetPlanMode": (C.c_int, [C.c_int]),
        "talshAmdGetPlanMode": (C.c_int, []),
        "talshAmdAutotuneChoice": (C.c_int, [C.c_char_p, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, C.c_int, C.c_int,
                                             C.c_int, _IP, _IP]),
        "talshAmdAutotuneClear": (C.c_int, []),
        "talshAmdPlanCandidates": (C.c_int, [C.c_char_p, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, C.c_int, C.c_int,
                                             C.c_int, C.POINTER(C.c_double), _IP]),
        "talshAmdPlanFusedCandidates": (C.c_int, [C.c_char_p, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, _IP]),
        "talshAmdPlanFusedMaps": (C.c_int, [C.c_char_p, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, _IP,
                                            C.POINTER(C.c_longlong)]),
        "talshAmdDeviceSynchronize": (C.c_int, [C.c_int]),
        "talshAmdPermutePlan": (C.c_int, [C.c_int, C.c_int, _IP, _IP, _IP]),
        "talshAmdGemm": (C.c_int, [C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p, C.c_int,
                                   C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int,
                                   C.c_void_p, C.c_size_t]),
        "talshAmdGemmWorkspace": (C.c_size_t, [C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t]),
        "talshAmdGemmNoAlloc": (C.c_int, [C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int]),
        "talshAmdBatchCreate": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_TP), C.POINTER(_TP), C.POINTER(_TP),
                                          C.c_double, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
        "talshAmdBatchExecute": (C.c_int, [C.c_void_p, C.c_int]),
        "talshAmdBatchInfo": (C.c_int, [C.c_void_p, _IP, _IP, C.POINTER(C.c_double)]),
        "talshAmdBatchDestroy": (C.c_int, [C.c_void_p]),
        "talshTensorContractXL": (C.c_int, [C.c_char_p, _TP, _TP, _TP, C.c_double, C.c_double, C.c_int, C.c_int,
                                            C.c_int]),
        "talshTensorSlice": (C.c_int, [_TP, _TP, _IP, C.c_int, C.c_int, C.c_int, C.c_int, _KP]),
        "talshTensorDecomposeSVD": (C.c_int, [C.c_char_p, _TP, _TP, _TP, _TP, C.c_char, C.c_int, C.c_int]),
        "talshTensorDecomposeSVDL": (C.c_int, [C.c_char_p, _TP, _TP, _TP, C.c_int, C.c_int]),
        "talshTensorDecomposeSVDR": (C.c_int, [C.c_char_p, _TP, _TP, _TP, C.c_int, C.c_int]),
        "talshTensorDecomposeSVDLR": (C.c_int, [C.c_char_p, _TP, _TP, _TP, C.c_int, C.c_int]),
        "talshTensorOrthogonalizeSVD": (C.c_int, [C.c_char_p, _TP, C.c_int, C.c_int]),
        "talshTensorInsert": (C.c_int, [_TP, _TP, _IP, C.c_int, C.c_int, C.c_int, C.c_int, _KP]),
        "talshAmdSetXLBudget": (C.c_int, [C.c_size_t]),
        "talshAmdMlpCreate": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)]),
        "talshAmdMlpDestroy": (C.c_int, [C.c_void_p]),
        "talshAmdMlpNumParams": (C.c_int, [C.c_void_p, _IP]),
        "talshAmdMlpGetParams": (C.c_int, [C.c_void_p, _FP]),
        "talshAmdMlpSetParams": (C.c_int, [C.c_void_p, _FP]),
        "talshAmdMlpSetNormalization": (C.c_int, [C.c_void_p, _FP, _FP, C.c_float, C.c_float]),
        "talshAmdMlpGradient": (C.c_int, [C.c_void_p, _FP, _FP, C.c_int, _FP, _FP]),
        "talshAmdMlpPredict": (C.c_int, [C.c_void_p, _FP, C.c_longlong, _FP]),
        "talshAmdMlpTrain": (C.c_int, [C.c_void_p, _FP, _FP, C.c_longlong, C.c_int, C.c_int, C.c_float,
                                       C.c_ulonglong, _FP]),
        "talshAmdMlpSave": (C.c_int, [C.c_void_p, C.c_char_p]),
        "talshAmdMlpLoad": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
        "talshAmdPlanFeatureCount": (C.c_int, []),
        "talshAmdPlanFeatures": (C.c_int, [C.c_char_p, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, C.c_int, C.c_int,
                                           C.c_int, C.c_int, _IP, _IP, _FP, _IP]),
        "talshAmdSelectorInstall": (C.c_int, [C.c_void_p]),
        "talshAmdSelectorChoice": (C.c_int, [C.c_char_p, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP, C.c_int, C.c_int,
                                             C.c_int, _IP, _IP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ints(vals, n=None):
    vals = list(vals)
    arr = (C.c_int * (n if n is not None else max(len(vals), 1)))()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


class Engine:
    """Initializes the engine on GPU 0 (or host-only with gpus=[])."""

    def __init__(self, gpus=(0,), arena_bytes=None):
        self.lib = load_library()
        if arena_bytes is not None:
            os.environ["TALSH_GPU_ARENA_BYTES"] = str(int(arena_bytes))
        sz = C.c_size_t(1 << 24)
        mx = C.c_int()
        g = ints(gpus) if gpus else None
        e = self.lib.talshInit(C.byref(sz), C.byref(mx), 0, None, 0, None, len(gpus), g)
        if e not in (TALSH_SUCCESS, TALSH_ALREADY_INITIALIZED):
            raise RuntimeError(f"talshInit failed: {e}")
        self.ngpus = len(gpus)

    def gpu(self, n=0):
        return self.lib.talshFlatDevId(DEV_AMD_GPU, n)

    def shutdown(self):
        return self.lib.talshShutdown()


class Tensor:
    """Owns one talsh_tens_t."""

    def __init__(self, lib, kind, dims, dev_id=0, init=0.0):
        self.lib = lib
        self.t = TalshTens()
        lib.talshTensorClean(C.byref(self.t))
        self.kind = kind
        self.dims = [int(x) for x in dims]
        v = complex(init)
        self.err = lib.talshTensorConstruct(C.byref(self.t), kind, len(self.dims), ints(self.dims), dev_id, None, -1,
                                            None, v.real, v.imag)

    @property
    def ptr(self):
        return C.byref(self.t)

    @property
    def volume(self):
        return int(np.prod(self.dims)) if self.dims else 1

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=NP_DTYPE[self.kind]).ravel()
        assert a.size == self.volume
        e = self.lib.talshTensorImportData(self.ptr, self.kind, a.ctypes.data)
        if e:
            raise RuntimeError(f"talshTensorImportData: {e}")

    def body(self, dev_id):
        p = C.c_void_p()
        e = self.lib.talshTensorGetBodyAccess(self.ptr, C.byref(p), self.kind, dev_id, DEV_NULL)
        return e, p.value

    def download(self, out=None):
        """Copies the body (from whichever image exists) into a numpy array
        (into `out` when given: a contiguous array of the body's size and
        dtype, e.g. a view of pinned host memory)."""
        if out is None:
            out = np.empty(self.volume, dtype=NP_DTYPE[self.kind])
        elif out.size != self.volume or out.dtype != NP_DTYPE[self.kind] or not out.flags.c_contiguous:
            raise ValueError("download: output array does not match the body")
        n = self.t.ndev
        if n <= 0:
            raise RuntimeError("tensor has no body image")
        copies = ints([0] * 16)
        kinds = ints([0] * 16)
        nc = C.c_int()
        self.lib.talshTensorPresence(self.ptr, C.byref(nc), copies, kinds, DEV_NULL, -1)
        dev = copies[0]
        e, p = self.body(dev)
        if e:
            raise RuntimeError(f"talshTensorGetBodyAccess: {e}")
        if dev == 0:
            C.memmove(out.ctypes.data, p, out.nbytes)
        else:
            hip = _hip()
            if hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(p), C.c_size_t(out.nbytes), 2) != 0:
                raise RuntimeError("hipMemcpy D2H failed")
        return out

    def images(self):
        copies = ints([0] * 16)
        kinds = ints([0] * 16)
        nc = C.c_int()
        self.lib.talshTensorPresence(self.ptr, C.byref(nc), copies, kinds, DEV_NULL, -1)
        return [copies[i] for i in range(nc.value)]

    def destruct(self):
        return self.lib.talshTensorDestruct(self.ptr)


_hiplib = None


def _hip():
    global _hiplib
    if _hiplib is None:
        _hiplib = C.CDLL("libamdhip64.so")
        _hiplib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hiplib.hipMemcpy.restype = C.c_int
    return _hiplib


def str2dig(lib, pattern):
    dig = (C.c_int * 128)()
    dr, lr, rr, cj = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    e = lib.talsh_get_contr_ptrn_str2dig(pattern.encode(), dig, C.byref(dr), C.byref(lr), C.byref(rr),
                                         C.byref(cj))
    n = max(lr.value, 0) + max(rr.value, 0)
    return e, list(dig[:n]) if e == 0 else [], dr.value, lr.value, rr.value, cj.value
