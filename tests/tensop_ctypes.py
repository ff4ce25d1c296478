"""ctypes layouts of talsh_tens_op_t and its parts (TALSH/talsh.h:98-140,
TALSH/tensor_algebra.h:214-249) and a recursive contraction decomposition
driven through talshTensorOpDecompose2 of a given TAL-SH library (the
reference's or the engine's): test infrastructure."""
import ctypes as C

MAX_OPERANDS = 4
TENSOR_CONTRACT = 82


class Tens(C.Structure):
    _fields_ = [("shape_p", C.c_void_p), ("dev_rsc", C.c_void_p), ("data_kind", C.POINTER(C.c_int)),
                ("avail", C.POINTER(C.c_int)), ("dev_rsc_len", C.c_int), ("ndev", C.c_int)]


class TensArg(C.Structure):
    _fields_ = [("tens_p", C.c_void_p), ("source_image", C.c_int)]


class Task(C.Structure):
    _fields_ = [("task_p", C.c_void_p), ("task_error", C.c_int), ("dev_kind", C.c_int), ("data_kind", C.c_int),
                ("coherence", C.c_int), ("num_args", C.c_int), ("tens_args", TensArg * MAX_OPERANDS),
                ("data_vol", C.c_double), ("flops", C.c_double), ("exec_time", C.c_double)]


class Signature(C.Structure):
    _fields_ = [("num_dim", C.c_int), ("offsets", C.POINTER(C.c_size_t))]


class Shape(C.Structure):
    _fields_ = [("num_dim", C.c_int), ("dims", C.POINTER(C.c_int)), ("divs", C.POINTER(C.c_int)),
                ("grps", C.POINTER(C.c_int))]


class TensSlice(C.Structure):
    _fields_ = [("tensor", C.c_void_p), ("bases", Signature), ("shape", Shape)]


class TensOp(C.Structure):
    _fields_ = [("opkind", C.c_int), ("data_kind", C.c_int), ("num_args", C.c_uint),
                ("tens_slice", TensSlice * MAX_OPERANDS), ("symb_pattern", C.c_char_p), ("alpha_real", C.c_double),
                ("alpha_imag", C.c_double), ("tens_arg", Tens * MAX_OPERANDS), ("task_handle", Task),
                ("exec_dev_id", C.c_int), ("stage", C.c_int), ("time_started", C.c_double),
                ("time_scheduled", C.c_double), ("time_completed", C.c_double), ("time_finished", C.c_double)]


def _ia(v):
    return (C.c_int * max(len(v), 1))(*v)


def decompose(lib, pattern, kind, ddims, ldims, rdims, limit_elems):
    """Leaves of the recursive halving of D+=L*R until every leaf's D+L+R
    volume is at most limit_elems: [[(offsets, dims) of D, L, R], ...] in
    depth-first order (first child first). Host tensors, no data."""
    P = C.POINTER
    lib.talshTensorClean.argtypes = [P(Tens)]
    lib.talshTensorConstruct.argtypes = [P(Tens), C.c_int, C.c_int, P(C.c_int), C.c_int, C.c_void_p, C.c_int,
                                         C.c_void_p, C.c_double, C.c_double]
    lib.talshTensorDestruct.argtypes = [P(Tens)]
    lib.talshTensorOpClean.argtypes = [P(TensOp)]
    lib.talshTensorOpSetArgument.argtypes = [P(TensOp), P(Tens), P(C.c_size_t), P(C.c_int)]
    lib.talshTensorOpSpecify.argtypes = [P(TensOp), C.c_int, C.c_int, C.c_char_p, C.c_double, C.c_double]
    lib.talshTensorOpDecompose2.argtypes = [P(TensOp), P(TensOp), P(TensOp)]
    lib.talshTensorOpDestruct.argtypes = [P(TensOp)]
    pat = pattern.encode()
    tens = []
    for dims in (ddims, ldims, rdims):
        t = Tens()
        lib.talshTensorClean(C.byref(t))
        e = lib.talshTensorConstruct(C.byref(t), kind, len(dims), _ia(dims), 0, None, -1, None, 0.0, 0.0)
        assert e == 0, e
        tens.append(t)
    root = TensOp()
    lib.talshTensorOpClean(C.byref(root))
    for t, dims in zip(tens, (ddims, ldims, rdims)):
        z = (C.c_size_t * max(len(dims), 1))()
        assert lib.talshTensorOpSetArgument(C.byref(root), C.byref(t), z, _ia(dims)) == 0
    assert lib.talshTensorOpSpecify(C.byref(root), TENSOR_CONTRACT, kind, pat, 1.0, 0.0) == 0
    leaves = []

    def rec(op, depth):
        parts, vol = [], 0
        for a in range(3):
            sl = op.tens_slice[a]
            r = sl.shape.num_dim
            offs = [int(sl.bases.offsets[i]) for i in range(r)]
            dims = [int(sl.shape.dims[i]) for i in range(r)]
            parts.append((offs, dims))
            v = 1
            for x in dims:
                v *= x
            vol += v
        if vol <= limit_elems or depth > 40:
            leaves.append(parts)
            return
        c1, c2 = TensOp(), TensOp()
        lib.talshTensorOpClean(C.byref(c1))
        lib.talshTensorOpClean(C.byref(c2))
        e = lib.talshTensorOpDecompose2(C.byref(op), C.byref(c1), C.byref(c2))
        assert e == 0, e
        rec(c1, depth + 1)
        rec(c2, depth + 1)
        lib.talshTensorOpDestruct(C.byref(c1))
        lib.talshTensorOpDestruct(C.byref(c2))

    rec(root, 0)
    lib.talshTensorOpDestruct(C.byref(root))
    for t in tens:
        lib.talshTensorDestruct(C.byref(t))
    return leaves
