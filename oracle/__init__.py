"""TEST INFRASTRUCTURE ONLY -- the oracle.

ctypes loader for

* ``oracle/liboracle.so``: the CPU restatement of the reference's solve phase
  (``bicgstab_oracle.c``; EPBiCGSTAB / PBiCGSTAB control flow, SpMV, the
  ACIFIM apply modes, triangular solves), and
* ``oracle/_ref/libref_afiim.so`` / ``libref_acifim.so``: the reference's own
  AFIIM / ACIFIM construction and Matrix Market readers, compiled unmodified
  from the reference tree by ``oracle/Makefile`` (its solve loops and PDFIIM's
  construction need MKL headers the image does not ship and are not built).

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
package, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)

_lib = None
_ref = {}


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        L.or_set_threads.argtypes = [C.c_int]
        L.or_csr_mv.argtypes = [C.c_int, PI, PI, PD, PD, PD]
        L.or_csr_mv_sym_upper.argtypes = [C.c_int, PI, PI, PD, PD, PD]
        L.or_csr_mv_trans.argtypes = [C.c_int, PI, PI, PD, PD, PD]
        L.or_trsv_lower.argtypes = [C.c_int, PI, PI, PD, PD, PD]
        L.or_trsv_upper.argtypes = [C.c_int, PI, PI, PD, PD, PD]
        L.or_dot.argtypes = [C.c_int, PD, PD]
        L.or_dot.restype = C.c_double
        L.or_nrm2.argtypes = [C.c_int, PD]
        L.or_nrm2.restype = C.c_double
        L.or_apply.argtypes = [C.c_int, C.c_int, C.c_char, PI, PI, PD, PD, PI, PI, PD, PI, PI, PD, PI, PI,
                               PD, PD, PD, PD]
        L.or_bicgstab.argtypes = [C.c_int, C.c_int, C.c_int, PI, PI, PD, PD, C.c_double, C.c_int, PI, PI,
                                  PD, PD, PI, PI, PD, PI, PI, PD, PI, PI, PD, C.c_int, C.c_char, PD, PD,
                                  PI, PI, PD, C.c_int]
        L.or_bicgstab.restype = C.c_int
        _lib = L
    return _lib


def set_threads(t):
    lib().or_set_threads(int(t))


def _c(m):
    if m is None:
        return None, None, None
    return (np.ascontiguousarray(m.ptr, np.int32).ctypes.data_as(PI),
            np.ascontiguousarray(m.idx, np.int32).ctypes.data_as(PI),
            np.ascontiguousarray(m.val, np.float64).ctypes.data_as(PD))


def _keep(*ms):
    out = []
    for m in ms:
        if m is None:
            out.append(None)
        else:
            out.append((np.ascontiguousarray(m.ptr, np.int32), np.ascontiguousarray(m.idx, np.int32),
                        np.ascontiguousarray(m.val, np.float64)))
    return out


def _ptrs(t):
    if t is None:
        return None, None, None
    return t[0].ctypes.data_as(PI), t[1].ctypes.data_as(PI), t[2].ctypes.data_as(PD)


def csr_mv(A, x, kind="general"):
    """y = A x; kind 'general', 'sym_upper' (A holds the upper triangle of a
    symmetric matrix) or 'trans' (y = A^T x)."""
    (a,) = _keep(A)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(A.n, np.float64)
    f = {"general": lib().or_csr_mv, "sym_upper": lib().or_csr_mv_sym_upper,
         "trans": lib().or_csr_mv_trans}[kind]
    f(A.n, *_ptrs(a), x.ctypes.data_as(PD), y.ctypes.data_as(PD))
    return y


def trsv(T, b, lower):
    (t,) = _keep(T)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty(T.n, np.float64)
    f = lib().or_trsv_lower if lower else lib().or_trsv_upper
    f(T.n, *_ptrs(t), b.ctypes.data_as(PD), x.ctypes.data_as(PD))
    return x


def apply(F, v, nr=1, mode="E"):
    """ACIFIM apply(): out = M v (modes E, C, I, J)."""
    g, h, l_, u = _keep(F.G, F.H, F.L, F.U)
    n = F.G.n
    idv = np.ascontiguousarray(F.idv, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    out = np.empty(n, np.float64)
    work = np.empty(3 * n, np.float64)
    lib().or_apply(n, nr, mode.encode(), *_ptrs(g), idv.ctypes.data_as(PD), *_ptrs(h), *_ptrs(l_),
                   *_ptrs(u), v.ctypes.data_as(PD), out.ctypes.data_as(PD), work.ctypes.data_as(PD))
    return out


def bicgstab(variant, A, b, F, x0=None, tol=1e-8, nmax=500, nr=1, mode="E", issym=False, trace_cap=0,
             info_in=0):
    """Reference solve loop.  variant: 'afiim', 'pdfiim' or 'acifim'.
    Returns (x, resval, iter, info, trace)."""
    var = {"afiim": 0, "pdfiim": 1, "acifim": 2}[variant]
    a, g, h, l_, u = _keep(A, F.G, F.H, F.L, F.U)
    n = A.n
    idv = np.ascontiguousarray(F.idv, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    resval, it, info = C.c_double(), C.c_int(), C.c_int(info_in)
    tr = np.zeros(max(trace_cap, 1), np.float64)
    nt = lib().or_bicgstab(var, int(bool(issym)), n, *_ptrs(a), b.ctypes.data_as(PD), tol, nmax, *_ptrs(g),
                           idv.ctypes.data_as(PD), *_ptrs(h), *_ptrs(l_), *_ptrs(u), nr, mode.encode(),
                           x.ctypes.data_as(PD), C.byref(resval), C.byref(it), C.byref(info),
                           tr.ctypes.data_as(PD) if trace_cap else None, trace_cap)
    trace = tr[:min(nt, trace_cap)].copy() if trace_cap else None
    return x, resval.value, it.value, info.value, trace


# ------------------------------------------------------------ reference builds

def ref_available(name="afiim"):
    return os.path.exists(os.path.join(HERE, "_ref", f"libref_{name}.so"))


def _reflib(name):
    if name not in _ref:
        path = os.path.join(HERE, "_ref", f"libref_{name}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle with the reference tree present)")
        L = C.CDLL(path)
        vp = C.c_void_p
        if name == "afiim":
            L.ref_afiim.argtypes = [C.c_int, PD, PI, PI] + [vp] * 7 + [C.c_double, C.c_int, C.c_int,
                                                                      C.c_double, C.c_double]
            L.ref_read_coeff_matrix.argtypes = [C.c_char_p, PI, vp, vp, vp]
            L.ref_read_rhs_vector.argtypes = [C.c_char_p, PI, vp]
        else:
            L.ref_acifim.argtypes = [C.c_int, PD, PI, PI] + [vp] * 13 + [
                C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                C.c_int]
            L.ref_acifim.restype = C.c_int
        L.ref_free_d.argtypes = [vp]
        L.ref_free_i.argtypes = [vp]
        _ref[name] = L
    return _ref[name]


def _take(L, p, count, dtype):
    ctype = C.c_double if dtype == np.float64 else C.c_int
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), (max(count, 1),))[:count].copy()
    (L.ref_free_d if dtype == np.float64 else L.ref_free_i)(p)
    return arr


def _take_csr(L, n, pv, pj, pi):
    from cmamd import Csr
    ptr = _take(L, pi.value, n + 1, np.int32)
    nnz = int(ptr[-1])
    return Csr(n, ptr, _take(L, pj.value, nnz, np.int32), _take(L, pv.value, nnz, np.float64))


def ref_read_coeff_matrix(path):
    from cmamd import Csr
    L = _reflib("afiim")
    n = C.c_int()
    pv, pj, pi = C.c_void_p(), C.c_void_p(), C.c_void_p()
    L.ref_read_coeff_matrix(os.fsencode(path), C.byref(n), C.byref(pv), C.byref(pj), C.byref(pi))
    return _take_csr(L, n.value, pv, pj, pi)


def ref_read_rhs_vector(path):
    L = _reflib("afiim")
    n, pb = C.c_int(), C.c_void_p()
    L.ref_read_rhs_vector(os.fsencode(path), C.byref(n), C.byref(pb))
    return _take(L, pb.value, n.value, np.float64)


def ref_afiim(A, dtol, elem_per_row_col=10, growth=5, eta=0.0, shift=0.0):
    from cmamd import Factors
    L = _reflib("afiim")
    (a,) = _keep(A)
    outs = [C.c_void_p() for _ in range(7)]
    L.ref_afiim(A.n, a[2].ctypes.data_as(PD), a[1].ctypes.data_as(PI), a[0].ctypes.data_as(PI),
                *[C.byref(o) for o in outs], dtol, elem_per_row_col, growth, eta, shift)
    Gv, Gj, Gi, IDv, Hv, Hj, Hi = outs
    return Factors(_take_csr(L, A.n, Gv, Gj, Gi), _take(L, IDv.value, A.n, np.float64),
                   _take_csr(L, A.n, Hv, Hj, Hi))


def ref_acifim(A, dtol1, dtol2, elem_per_row_col=30, growth=5, eta=0.0, shift=0.0, coeff=2.0,
               element_limit=200, filtit=30):
    """Returns (rc, Factors or None)."""
    from cmamd import Factors
    L = _reflib("acifim")
    (a,) = _keep(A)
    outs = [C.c_void_p() for _ in range(13)]
    rc = L.ref_acifim(A.n, a[2].ctypes.data_as(PD), a[1].ctypes.data_as(PI), a[0].ctypes.data_as(PI),
                      *[C.byref(o) for o in outs], dtol1, dtol2, elem_per_row_col, growth, eta, shift,
                      coeff, element_limit, filtit)
    if rc != 0:
        return rc, None
    Gv, Gj, Gi, IDv, Hv, Hj, Hi, Lv, Lj, Li, Uv, Uj, Ui = outs
    n = A.n
    return 0, Factors(_take_csr(L, n, Gv, Gj, Gi), _take(L, IDv.value, n, np.float64),
                      _take_csr(L, n, Hv, Hj, Hi), _take_csr(L, n, Lv, Lj, Li), _take_csr(L, n, Uv, Uj, Ui))
