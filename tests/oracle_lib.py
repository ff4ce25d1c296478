"""ctypes bindings of the CP-TAL oracle (oracle/libcptal.so) and, when it has
been built, of the reference TAL-SH CPU library (oracle/_ref/libtalsh_ref.so).

Test infrastructure only: the product library never imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libcptal.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtalsh_ref.so")

R4, R8, C4, C8 = 4, 8, 14, 18
NP_DTYPE = {R4: np.float32, R8: np.float64, C4: np.complex64, C8: np.complex128}
KIND_OF = {np.dtype(np.float32): R4, np.dtype(np.float64): R8,
           np.dtype(np.complex64): C4, np.dtype(np.complex128): C8}

_IP = C.POINTER(C.c_int)


def _ia(vals, n=None):
    vals = list(vals)
    if n is None:
        n = max(len(vals), 1)
    arr = (C.c_int * n)()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


class Oracle:
    """The C restatement of CP-TAL."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        self.lib = L = C.CDLL(path)
        L.cptal_str2dig.argtypes = [C.c_char_p, _IP, _IP, _IP, _IP, _IP]
        L.cptal_dig2str.argtypes = [C.c_int, C.c_int, C.c_int, _IP, C.c_char_p, C.c_int]
        L.cptal_copy_dlf.argtypes = [C.c_int, C.c_int, _IP, _IP, C.c_void_p, C.c_void_p, C.c_int]
        L.cptal_tensor_copy.argtypes = [C.c_int, C.c_int, _IP, _IP, C.c_void_p, C.c_void_p, C.c_int]
        L.cptal_slice.argtypes = [C.c_int, C.c_int, _IP, C.c_void_p, _IP, C.c_void_p, _IP, C.c_int]
        L.cptal_insert.argtypes = [C.c_int, C.c_int, _IP, C.c_void_p, _IP, C.c_void_p, _IP, C.c_int]
        L.cptal_tensor_add.argtypes = [C.c_int, C.c_int, _IP, _IP, C.c_void_p, C.c_void_p,
                                       C.c_double, C.c_double, C.c_int]
        L.cptal_init.argtypes = [C.c_int, C.c_size_t, C.c_void_p, C.c_double, C.c_double]
        L.cptal_scale.argtypes = [C.c_int, C.c_size_t, C.c_void_p, C.c_double, C.c_double]
        L.cptal_contract.argtypes = [_IP, C.c_int, C.c_int, _IP, C.c_void_p, C.c_int, _IP, C.c_void_p,
                                     C.c_int, _IP, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int]
        L.cptal_contract_ld_sample.argtypes = [_IP, C.c_int, C.c_int, _IP, C.c_void_p, C.c_int, _IP,
                                               C.c_void_p, C.c_int, _IP, C.c_void_p, C.c_double,
                                               C.c_double, C.c_int, C.c_int, C.c_size_t,
                                               C.POINTER(C.c_size_t), C.c_void_p]
        L.cptal_fill_uniform.argtypes = [C.c_int, C.c_size_t, C.c_void_p, C.c_ulonglong, C.c_double,
                                         C.c_double]
        L.cptal_fill_int.argtypes = [C.c_size_t, C.c_void_p, C.c_ulonglong]
        L.cptal_fill_normal.argtypes = [C.c_int, C.c_size_t, C.c_void_p, C.c_ulonglong]
        for fn in ("cptal_get_contr_permutations", "cptal_get_contraction_permutations"):
            getattr(L, fn).restype = None

    # --- patterns -----------------------------------------------------------
    def str2dig(self, s):
        dig = (C.c_int * 128)()
        dr, lr, rr, cj = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        e = self.lib.cptal_str2dig(s.encode(), dig, C.byref(dr), C.byref(lr), C.byref(rr), C.byref(cj))
        n = max(lr.value, 0) + max(rr.value, 0)
        return e, list(dig[:n]) if e == 0 else [], dr.value, lr.value, rr.value, cj.value

    def dig2str(self, lrank, rrank, conj, dig):
        buf = C.create_string_buffer(512)
        e = self.lib.cptal_dig2str(lrank, rrank, conj, _ia(dig), buf, 512)
        return e, buf.value.decode()

    def permutations(self, hyper, tl, tr, lrank, rrank, dig, conj):
        d, l, r = (C.c_int * 64)(), (C.c_int * 64)(), (C.c_int * 64)()
        ncd, nlu, nru, nhu, ierr = (C.c_int() for _ in range(5))
        if hyper:
            self.lib.cptal_get_contraction_permutations(tl, tr, lrank, rrank, _ia(dig), conj, d, l, r,
                                                        C.byref(ncd), C.byref(nlu), C.byref(nru),
                                                        C.byref(nhu), C.byref(ierr))
        else:
            self.lib.cptal_get_contr_permutations(tl, tr, lrank, rrank, _ia(dig), conj, d, l, r,
                                                  C.byref(ncd), C.byref(nlu), C.byref(nru), C.byref(ierr))
        drank = sum(1 for v in dig if v > 0) if not hyper else max([v for v in dig if v > 0] + [0])
        return dict(ierr=ierr.value, dprm=list(d[:drank + 1]), lprm=list(l[:lrank + 1]),
                    rprm=list(r[:rrank + 1]), ncd=ncd.value, nlu=nlu.value, nru=nru.value,
                    nhu=nhu.value if hyper else 0)

    # --- data ---------------------------------------------------------------
    def uniform(self, kind, shape, seed, lo=-1.0, hi=1.0):
        a = np.empty(int(np.prod(shape)) if len(shape) else 1, dtype=NP_DTYPE[kind])
        self.lib.cptal_fill_uniform(kind, a.size, a.ctypes.data, seed, lo, hi)
        return a

    def integers(self, n, seed):
        a = np.empty(n, dtype=np.float64)
        self.lib.cptal_fill_int(n, a.ctypes.data, seed)
        return a

    def normal(self, kind, n, seed):
        a = np.empty(n, dtype=NP_DTYPE[kind])
        self.lib.cptal_fill_normal(kind, n, a.ctypes.data, seed)
        return a

    # --- operations ---------------------------------------------------------
    def slice(self, tens, tdims, slc, sdims, offsets, accumulative):
        out = np.array(slc, copy=True)
        e = self.lib.cptal_slice(KIND_OF[tens.dtype], len(tdims), _ia(tdims), tens.ctypes.data, _ia(sdims),
                                 out.ctypes.data, _ia(offsets), 1 if accumulative else 0)
        return e, out

    def insert(self, tens, tdims, slc, sdims, offsets, accumulative):
        out = np.array(tens, copy=True)
        e = self.lib.cptal_insert(KIND_OF[tens.dtype], len(tdims), _ia(tdims), out.ctypes.data, _ia(sdims),
                                  slc.ctypes.data, _ia(offsets), 1 if accumulative else 0)
        return e, out

    def copy(self, pattern, ldims, l, conj=None):
        e, dig, dr, lr, rr, cj = self.str2dig(pattern)
        assert e == 0, (pattern, e)
        kind = KIND_OF[l.dtype]
        d = np.empty_like(l)
        e = self.lib.cptal_tensor_copy(kind, len(ldims), _ia(ldims), _ia(dig), l.ctypes.data, d.ctypes.data,
                                       cj if conj is None else conj)
        assert e == 0
        return d

    def permute(self, ldims, o2n_1based, l, conj=0):
        kind = KIND_OF[l.dtype]
        d = np.empty_like(l)
        e = self.lib.cptal_copy_dlf(kind, len(ldims), _ia(ldims), _ia([1] + list(o2n_1based)),
                                    l.ctypes.data, d.ctypes.data, conj)
        assert e == 0, e
        return d

    def add(self, pattern, ldims, d, l, scale=1.0):
        e, dig, dr, lr, rr, cj = self.str2dig(pattern)
        assert e == 0
        kind = KIND_OF[l.dtype]
        s = complex(scale)
        e = self.lib.cptal_tensor_add(kind, len(ldims), _ia(ldims), _ia(dig), l.ctypes.data, d.ctypes.data,
                                      s.real, s.imag, cj)
        assert e == 0, e
        return d

    def contract(self, pattern, d, ddims, l, ldims, r, rdims, alpha=1.0, accumulative=True):
        e, dig, dr, lr, rr, cj = self.str2dig(pattern)
        if e:
            return e
        kind = KIND_OF[d.dtype]
        a = complex(alpha)
        return self.lib.cptal_contract(_ia(dig), kind, len(ddims), _ia(ddims), d.ctypes.data, len(ldims),
                                       _ia(ldims), l.ctypes.data, len(rdims), _ia(rdims), r.ctypes.data,
                                       a.real, a.imag, cj, 1 if accumulative else 0)

    def contract_exact(self, pattern, d, ddims, l, ldims, r, rdims, alpha=1.0, accumulative=True,
                       positions=None):
        """Long-double direct evaluation at DLF offsets `positions` (all if None).
        Returns complex128 values (the long double result rounded to double)."""
        e, dig, dr, lr, rr, cj = self.str2dig(pattern)
        assert e == 0
        kind = KIND_OF[d.dtype]
        if positions is None:
            positions = np.arange(d.size, dtype=np.uint64)
        pos = np.ascontiguousarray(positions, dtype=np.uint64)
        out = np.zeros(2 * pos.size, dtype=np.longdouble)
        a = complex(alpha)
        e = self.lib.cptal_contract_ld_sample(_ia(dig), kind, len(ddims), _ia(ddims), d.ctypes.data, len(ldims),
                                              _ia(ldims), l.ctypes.data, len(rdims), _ia(rdims), r.ctypes.data,
                                              a.real, a.imag, cj, 1 if accumulative else 0, pos.size,
                                              pos.ctypes.data_as(C.POINTER(C.c_size_t)), out.ctypes.data)
        assert e == 0, e
        return out[0::2].astype(np.float64) + 1j * out[1::2].astype(np.float64)


class TalshTens(C.Structure):
    # talsh_tens_t (TALSH/talsh.h:87-94)
    _fields_ = [("shape_p", C.c_void_p), ("dev_rsc", C.c_void_p), ("data_kind", C.c_void_p),
                ("avail", C.c_void_p), ("dev_rsc_len", C.c_int), ("ndev", C.c_int)]


class Reference:
    """The reference TAL-SH CPU library (built from its own sources into oracle/_ref)."""

    _inited = False

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise RuntimeError("reference CPU library not built")
        self.lib = L = C.CDLL(path)
        L.talsh_get_contr_ptrn_str2dig.argtypes = [C.c_char_p, _IP, _IP, _IP, _IP, _IP]
        L.get_contr_permutations.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _IP, C.c_int, _IP, _IP, _IP,
                                             _IP, _IP, _IP, _IP]
        L.get_contr_permutations.restype = None
        L.get_contraction_permutations.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _IP, C.c_int, _IP, _IP,
                                                   _IP, _IP, _IP, _IP, _IP, _IP]
        L.get_contraction_permutations.restype = None
        L.talshInit.argtypes = [C.POINTER(C.c_size_t), _IP, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP]
        L.talshTensorConstruct.argtypes = [C.POINTER(TalshTens), C.c_int, C.c_int, _IP, C.c_int, C.c_void_p,
                                           C.c_int, C.c_void_p, C.c_double, C.c_double]
        L.talshTensorClean.argtypes = [C.POINTER(TalshTens)]
        L.talshTensorDestruct.argtypes = [C.POINTER(TalshTens)]
        L.talshTensorGetBodyAccess.argtypes = [C.POINTER(TalshTens), C.POINTER(C.c_void_p), C.c_int, C.c_int,
                                               C.c_int]
        L.talshTensorContract.argtypes = [C.c_char_p, C.POINTER(TalshTens), C.POINTER(TalshTens),
                                          C.POINTER(TalshTens), C.c_double, C.c_double, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_void_p]
        L.talshTensorCopy.argtypes = [C.c_char_p, C.POINTER(TalshTens), C.POINTER(TalshTens), C.c_int, C.c_int,
                                      C.c_int, C.c_void_p]
        L.talshTensorAdd.argtypes = [C.c_char_p, C.POINTER(TalshTens), C.POINTER(TalshTens), C.c_double,
                                     C.c_double, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.talshTensorInit.argtypes = [C.POINTER(TalshTens), C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.c_void_p]
        L.talshTensorScale.argtypes = [C.POINTER(TalshTens), C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                       C.c_void_p]
        for fn in (L.talshTensorSlice, L.talshTensorInsert):
            fn.argtypes = [C.POINTER(TalshTens), C.POINTER(TalshTens), _IP, C.c_int, C.c_int, C.c_int, C.c_int,
                           C.c_void_p]
        if not Reference._inited:
            sz = C.c_size_t(1 << 28)
            mx = C.c_int()
            e = L.talshInit(C.byref(sz), C.byref(mx), 0, None, 0, None, 0, None)
            assert e == 0, e
            Reference._inited = True

    def str2dig(self, s):
        dig = (C.c_int * 128)()
        dr, lr, rr, cj = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        e = self.lib.talsh_get_contr_ptrn_str2dig(s.encode(), dig, C.byref(dr), C.byref(lr), C.byref(rr),
                                                  C.byref(cj))
        n = max(lr.value, 0) + max(rr.value, 0)
        return e, list(dig[:n]) if e == 0 else [], dr.value, lr.value, rr.value, cj.value

    def permutations(self, hyper, tl, tr, lrank, rrank, dig, conj):
        d, l, r = (C.c_int * 64)(), (C.c_int * 64)(), (C.c_int * 64)()
        ncd, nlu, nru, nhu, ierr = (C.c_int() for _ in range(5))
        if hyper:
            self.lib.get_contraction_permutations(tl, tr, lrank, rrank, _ia(dig), conj, d, l, r, C.byref(ncd),
                                                  C.byref(nlu), C.byref(nru), C.byref(nhu), C.byref(ierr))
        else:
            self.lib.get_contr_permutations(tl, tr, lrank, rrank, _ia(dig), conj, d, l, r, C.byref(ncd),
                                            C.byref(nlu), C.byref(nru), C.byref(ierr))
        drank = sum(1 for v in dig if v > 0) if not hyper else max([v for v in dig if v > 0] + [0])
        return dict(ierr=ierr.value, dprm=list(d[:drank + 1]), lprm=list(l[:lrank + 1]),
                    rprm=list(r[:rrank + 1]), ncd=ncd.value, nlu=nlu.value, nru=nru.value,
                    nhu=nhu.value if hyper else 0)

    def _tensor(self, arr, dims):
        t = TalshTens()
        self.lib.talshTensorClean(C.byref(t))
        kind = KIND_OF[arr.dtype]
        e = self.lib.talshTensorConstruct(C.byref(t), kind, len(dims), _ia(dims), 0, None, -1, None, 0.0, 0.0)
        assert e == 0, e
        p = C.c_void_p()
        e = self.lib.talshTensorGetBodyAccess(C.byref(t), C.byref(p), kind, 0, 0)
        assert e == 0, e
        C.memmove(p, arr.ctypes.data, arr.nbytes)
        return t, p

    def _read(self, p, like):
        out = np.empty_like(like)
        C.memmove(out.ctypes.data, p, out.nbytes)
        return out

    def contract(self, pattern, d, ddims, l, ldims, r, rdims, alpha=1.0, accumulative=True):
        td, pd = self._tensor(d, ddims)
        tl, _ = self._tensor(l, ldims)
        tr, _ = self._tensor(r, rdims)
        a = complex(alpha)
        e = self.lib.talshTensorContract(pattern.encode(), C.byref(td), C.byref(tl), C.byref(tr), a.real, a.imag,
                                         0, 0, 26, 1 if accumulative else 0, None)
        out = self._read(pd, d) if e == 0 else None
        for t in (td, tl, tr):
            self.lib.talshTensorDestruct(C.byref(t))
        return e, out

    def copy(self, pattern, ddims, ldims, l):
        td, pd = self._tensor(np.zeros(int(np.prod(ddims)) if ddims else 1, dtype=l.dtype), ddims)
        tl, _ = self._tensor(l, ldims)
        e = self.lib.talshTensorCopy(pattern.encode(), C.byref(td), C.byref(tl), 0, 0, 6, None)
        out = self._read(pd, l) if e == 0 else None
        for t in (td, tl):
            self.lib.talshTensorDestruct(C.byref(t))
        return e, out

    def add(self, pattern, d, ddims, l, ldims, scale=1.0):
        td, pd = self._tensor(d, ddims)
        tl, _ = self._tensor(l, ldims)
        s = complex(scale)
        e = self.lib.talshTensorAdd(pattern.encode(), C.byref(td), C.byref(tl), s.real, s.imag, 0, 0, 6, None)
        out = self._read(pd, d) if e == 0 else None
        for t in (td, tl):
            self.lib.talshTensorDestruct(C.byref(t))
        return e, out

    def slice(self, tens, tdims, slc, sdims, offsets, accumulative):
        """talshTensorSlice on the host: slice (sdims, initial values slc) from tens."""
        td, pd = self._tensor(slc, sdims)
        tl, _ = self._tensor(tens, tdims)
        e = self.lib.talshTensorSlice(C.byref(td), C.byref(tl), _ia(offsets), 0, 0, 6, 1 if accumulative else 0,
                                      None)
        out = self._read(pd, slc) if e == 0 else None
        for t in (td, tl):
            self.lib.talshTensorDestruct(C.byref(t))
        return e, out

    def insert(self, tens, tdims, slc, sdims, offsets, accumulative):
        """talshTensorInsert on the host: slc inserted into tens at offsets."""
        td, pd = self._tensor(tens, tdims)
        tl, _ = self._tensor(slc, sdims)
        e = self.lib.talshTensorInsert(C.byref(td), C.byref(tl), _ia(offsets), 0, 0, 6, 1 if accumulative else 0,
                                       None)
        out = self._read(pd, tens) if e == 0 else None
        for t in (td, tl):
            self.lib.talshTensorDestruct(C.byref(t))
        return e, out

    def init(self, d, ddims, val):
        td, pd = self._tensor(d, ddims)
        v = complex(val)
        e = self.lib.talshTensorInit(C.byref(td), v.real, v.imag, 0, 0, 1, None)
        out = self._read(pd, d) if e == 0 else None
        self.lib.talshTensorDestruct(C.byref(td))
        return e, out

    def scale(self, d, ddims, val):
        td, pd = self._tensor(d, ddims)
        v = complex(val)
        e = self.lib.talshTensorScale(C.byref(td), v.real, v.imag, 0, 0, 1, None)
        out = self._read(pd, d) if e == 0 else None
        self.lib.talshTensorDestruct(C.byref(td))
        return e, out
