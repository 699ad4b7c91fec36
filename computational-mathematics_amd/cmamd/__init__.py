"""ctypes binding of the MI355X BiCGSTAB solve engine.

Two shared libraries, both built in-tree (``make -C computational-mathematics_amd``):

* ``lib/libcmsolve.so`` -- the HIP solve engine behind the C ABI of
  ``include/cm_solve.h`` (drop-in replacements for the reference's
  ``EPBiCGSTAB`` / ``PBiCGSTAB`` plus resident systems).
* ``lib/libcmhost.so`` -- host factor producers (AFIIM, ACIFIM, PDFIIM),
  Matrix Market ingest and the seeded workload generators
  (``include/cm_host.h``).

Nothing here computes: every call goes to one of the two libraries, and a
missing library raises at import time of the function that needs it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(os.path.dirname(_HERE), "lib")

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)

CM_OK = 0
CM_ERR_HIP = -1
CM_ERR_ARG = -2
CM_ERR_UNSUPPORTED = -3
CM_ERR_NODEV = -4

VARIANT = {"afiim": 0, "pdfiim": 1, "acifim": 2}


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"engine error {code}: {msg}")
        self.code = code


_libs = {}


def _load(name):
    if name not in _libs:
        path = os.path.join(LIBDIR, name)
        if name == "libcmhost.so" and os.environ.get("CMAMD_HOST_LIB"):
            path = os.environ["CMAMD_HOST_LIB"]  # e.g. the sanitized build, lib/asan/libcmhost.so
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build() "
                               f"or make -C computational-mathematics_amd")
        _libs[name] = C.CDLL(path)
    return _libs[name]


def solve_lib():
    lib = _load("libcmsolve.so")
    if not getattr(lib, "_typed", False):
        _type_solve(lib)
        lib._typed = True
    return lib


def host_lib():
    lib = _load("libcmhost.so")
    if not getattr(lib, "_typed", False):
        _type_host(lib)
        lib._typed = True
    return lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(PD)


def _ip(a):
    return None if a is None else a.ctypes.data_as(PI)


@dataclass
class Csr:
    """CSR matrix: 0-based int32 indices, float64 values, columns ascending."""
    n: int
    ptr: np.ndarray
    idx: np.ndarray
    val: np.ndarray

    @property
    def nnz(self):
        return int(self.ptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.idx, self.ptr), shape=(self.n, self.n))

    @staticmethod
    def from_scipy(m):
        m = m.tocsr()
        m.sort_indices()
        return Csr(m.shape[0], np.ascontiguousarray(m.indptr, dtype=np.int32),
                   np.ascontiguousarray(m.indices, dtype=np.int32),
                   np.ascontiguousarray(m.data, dtype=np.float64))


def _csr_c(m):
    return (np.ascontiguousarray(m.ptr, dtype=np.int32), np.ascontiguousarray(m.idx, dtype=np.int32),
            np.ascontiguousarray(m.val, dtype=np.float64))


# --------------------------------------------------------------------------- host

def _type_host(lib):
    vp = C.c_void_p
    lib.cm_afiim.argtypes = [C.c_int, PD, PI, PI] + [vp] * 7 + [C.c_double, C.c_int, C.c_int,
                                                               C.c_double, C.c_double]
    lib.cm_afiim.restype = None
    lib.cm_acifim.argtypes = [C.c_int, PD, PI, PI] + [vp] * 13 + [
        C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
    lib.cm_acifim.restype = C.c_int
    lib.cm_pdfiim.argtypes = [C.c_int, PD, PI, PI] + [vp] * 7 + [
        C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_char, C.c_char, C.c_int]
    lib.cm_pdfiim.restype = C.c_int
    lib.cm_free_d.argtypes = [vp]
    lib.cm_free_i.argtypes = [vp]
    lib.cm_read_mtx.argtypes = [C.c_char_p, C.c_int, PI, PI, vp, vp, vp, PI]
    lib.cm_read_mtx.restype = C.c_int
    lib.cm_read_mtx_vector.argtypes = [C.c_char_p, PI, vp]
    lib.cm_read_mtx_vector.restype = C.c_int
    lib.cm_gen_convdiff.argtypes = [C.c_int, C.c_int, C.c_double, C.c_ulonglong, PI,
                                    C.POINTER(C.c_longlong), PI, PI, PD]
    lib.cm_gen_convdiff.restype = C.c_int
    lib.cm_gen_stencil27.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_ulonglong, PI,
                                     C.POINTER(C.c_longlong), PI, PI, PD]
    lib.cm_gen_stencil27.restype = C.c_int
    lib.cm_gen_uniform.argtypes = [C.c_ulonglong, C.c_longlong, PD]
    lib.cm_gen_uniform.restype = None
    lib.cm_csr_mv_host.argtypes = [C.c_int, PI, PI, PD, PD, PD]
    lib.cm_csr_mv_host.restype = None


def _take(lib, p, count, dtype):
    """Copy a new[]-allocated array into numpy and release it."""
    if not p:
        return None
    ctype = C.c_double if dtype == np.float64 else C.c_int
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(ctype)), (max(count, 1),))[:count].copy()
    (lib.cm_free_d if dtype == np.float64 else lib.cm_free_i)(p)
    return arr


def _take_csr(lib, n, pv, pj, pi):
    ptr = _take(lib, pi.value, n + 1, np.int32)
    nnz = int(ptr[-1])
    return Csr(n, ptr, _take(lib, pj.value, nnz, np.int32), _take(lib, pv.value, nnz, np.float64))


@dataclass
class Factors:
    """Factor interface of the solve: M = G diag(idv) H (+ L, U for ACIFIM)."""
    G: Csr
    idv: np.ndarray
    H: Csr | None
    L: Csr | None = None
    U: Csr | None = None


def afiim(A: Csr, dtol, elem_per_row_col=10, growth=5, eta=0.0, shift=0.0) -> Factors:
    lib = host_lib()
    ptr, idx, val = _csr_c(A)
    outs = [C.c_void_p() for _ in range(7)]
    lib.cm_afiim(A.n, _dp(val), _ip(idx), _ip(ptr), *[C.byref(o) for o in outs], dtol,
                 elem_per_row_col, growth, eta, shift)
    Gv, Gj, Gi, IDv, Hv, Hj, Hi = outs
    return Factors(_take_csr(lib, A.n, Gv, Gj, Gi), _take(lib, IDv.value, A.n, np.float64),
                   _take_csr(lib, A.n, Hv, Hj, Hi))


def acifim(A: Csr, dtol1, dtol2, elem_per_row_col=30, growth=5, eta=0.0, shift=0.0, coeff=2.0,
           element_limit=200, filtit=30) -> Factors:
    lib = host_lib()
    ptr, idx, val = _csr_c(A)
    outs = [C.c_void_p() for _ in range(13)]
    rc = lib.cm_acifim(A.n, _dp(val), _ip(idx), _ip(ptr), *[C.byref(o) for o in outs], dtol1, dtol2,
                       elem_per_row_col, growth, eta, shift, coeff, element_limit, filtit)
    if rc != 0:
        raise EngineError(rc, "acifim: NaN Schur complement")
    Gv, Gj, Gi, IDv, Hv, Hj, Hi, Lv, Lj, Li, Uv, Uj, Ui = outs
    n = A.n
    return Factors(_take_csr(lib, n, Gv, Gj, Gi), _take(lib, IDv.value, n, np.float64),
                   _take_csr(lib, n, Hv, Hj, Hi), _take_csr(lib, n, Lv, Lj, Li),
                   _take_csr(lib, n, Uv, Uj, Ui))


def pdfiim(A: Csr, prefill, lfill, postfill, ftol, warpsize=-1, filttype="I", patt="B",
           nthreads=0) -> Factors:
    lib = host_lib()
    ptr, idx, val = _csr_c(A)
    outs = [C.c_void_p() for _ in range(7)]
    rc = lib.cm_pdfiim(A.n, _dp(val), _ip(idx), _ip(ptr), *[C.byref(o) for o in outs], prefill, lfill,
                       postfill, ftol, warpsize, filttype.encode(), patt.encode(), nthreads)
    if rc != 0:
        raise EngineError(rc, "pdfiim: unsupported arguments")
    Gv, Gj, Gi, IDv, Hv, Hj, Hi = outs
    return Factors(_take_csr(lib, A.n, Gv, Gj, Gi), _take(lib, IDv.value, A.n, np.float64),
                   _take_csr(lib, A.n, Hv, Hj, Hi))


def pdfiim_device(A: Csr, prefill, lfill, postfill, ftol, warpsize=-1, filttype="I", patt="B"):
    """PDFIIM factors built on the GPU (libcmsolve.so); the host producer's
    factors bit for bit.  Returns (Factors, construction kernels' device ms)."""
    lib = solve_lib()
    hl = host_lib()  # releases the returned new[] arrays
    ptr, idx, val = _csr_c(A)
    outs = [C.c_void_p() for _ in range(7)]
    ms = C.c_double()
    _check(lib.cm_pdfiim_device(A.n, _dp(val), _ip(idx), _ip(ptr), *[C.byref(o) for o in outs], prefill, lfill,
                                postfill, ftol, warpsize, filttype.encode(), patt.encode(), C.byref(ms)))
    Gv, Gj, Gi, IDv, Hv, Hj, Hi = outs
    return (Factors(_take_csr(hl, A.n, Gv, Gj, Gi), _take(hl, IDv.value, A.n, np.float64),
                    _take_csr(hl, A.n, Hv, Hj, Hi)), ms.value)


def read_mtx(path, symmetric_mode=0):
    lib = host_lib()
    n, nnz, sym = C.c_int(), C.c_int(), C.c_int()
    pi, pj, pv = C.c_void_p(), C.c_void_p(), C.c_void_p()
    rc = lib.cm_read_mtx(os.fsencode(path), symmetric_mode, C.byref(n), C.byref(nnz), C.byref(pi),
                         C.byref(pj), C.byref(pv), C.byref(sym))
    if rc != 0:
        raise EngineError(rc, f"cannot read {path}")
    return _take_csr(lib, n.value, pv, pj, pi), bool(sym.value)


def read_mtx_vector(path):
    lib = host_lib()
    n, pb = C.c_int(), C.c_void_p()
    rc = lib.cm_read_mtx_vector(os.fsencode(path), C.byref(n), C.byref(pb))
    if rc != 0:
        raise EngineError(rc, f"cannot read {path}")
    return _take(lib, pb.value, n.value, np.float64)


def gen_convdiff(dim, N, peclet, seed) -> Csr:
    lib = host_lib()
    n, nnz = C.c_int(), C.c_longlong()
    rc = lib.cm_gen_convdiff(dim, N, peclet, seed, C.byref(n), C.byref(nnz), None, None, None)
    if rc != 0:
        raise EngineError(rc, "gen_convdiff")
    ptr = np.empty(n.value + 1, np.int32)
    idx = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    rc = lib.cm_gen_convdiff(dim, N, peclet, seed, C.byref(n), C.byref(nnz), _ip(ptr), _ip(idx), _dp(val))
    if rc != 0:
        raise EngineError(rc, "gen_convdiff")
    return Csr(n.value, ptr, idx, val)


def gen_stencil27(N, extra_frac, extra_lo, extra_hi, seed) -> Csr:
    lib = host_lib()
    n, nnz = C.c_int(), C.c_longlong()
    args = (N, extra_frac, extra_lo, extra_hi, seed, C.byref(n), C.byref(nnz))
    rc = lib.cm_gen_stencil27(*args, None, None, None)
    if rc != 0:
        raise EngineError(rc, "gen_stencil27")
    ptr = np.empty(n.value + 1, np.int32)
    idx = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    rc = lib.cm_gen_stencil27(*args, _ip(ptr), _ip(idx), _dp(val))
    if rc != 0:
        raise EngineError(rc, "gen_stencil27")
    return Csr(n.value, ptr, idx, val)


def gen_uniform(seed, count):
    x = np.empty(count, np.float64)
    host_lib().cm_gen_uniform(seed, count, _dp(x))
    return x


def csr_mv_host(A: Csr, x):
    ptr, idx, val = _csr_c(A)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(A.n, np.float64)
    host_lib().cm_csr_mv_host(A.n, _ip(ptr), _ip(idx), _dp(val), _dp(x), _dp(y))
    return y


# -------------------------------------------------------------------------- solve

SOLVE_SYMBOLS = [
    "cm_afiim_epbicgstab", "cm_pdfiim_epbicgstab", "cm_acifim_pbicgstab", "cm_system_create",
    "cm_system_destroy", "cm_system_set_rhs", "cm_system_set_x", "cm_system_get_x", "cm_system_solve",
    "cm_system_spmv", "cm_system_apply", "cm_system_nnz", "cm_system_profile", "cm_system_kernel_stats",
    "cm_system_kernel_format_bytes",
    "cm_system_enable_mixed", "cm_system_solve_mixed",
    "cm_system_block_set", "cm_system_block_get", "cm_system_block_solve", "cm_system_block_solve_mixed",
    "cm_system_graphs", "cm_system_block_graphs", "cm_system_block_release",
    "cm_system_block_time_spmm", "cm_system_dia_info", "cm_batch_solve", "cm_system_autotune", "cm_system_block_autotune",
    "cm_system_spmv_format", "cm_system_spmv_formats", "cm_system_profile_sampled", "cm_system_block_dia", "cm_last_error", "cm_device_count",
    "cm_server_create", "cm_server_configure", "cm_server_destroy", "cm_server_add_system", "cm_server_submit", "cm_server_wait",
    "cm_server_stats", "cm_batch_solve_noalloc", "cm_pdfiim_device",
]


def _type_solve(lib):
    vp = C.c_void_p
    lib.cm_last_error.restype = C.c_char_p
    lib.cm_device_count.restype = C.c_int
    lib.cm_afiim_epbicgstab.argtypes = [C.c_int, PD, PI, PI, PD, C.c_double, C.c_int, PD, PI, PI, PD,
                                        PD, PI, PI, PD, PD, PI, C.c_int, PI]
    lib.cm_pdfiim_epbicgstab.argtypes = [C.c_int, C.c_int, PI, PI, PD, PD, C.c_double, C.c_int, PI, PI,
                                         PD, PD, PI, PI, PD, PD, PD, PI, C.c_int, PI]
    lib.cm_acifim_pbicgstab.argtypes = [C.c_int, PD, PI, PI, PD, C.c_double, C.c_int, PD, PI, PI, PD,
                                        PD, PI, PI, PD, PI, PI, PD, PI, PI, C.c_int, C.c_char, PD, PD,
                                        PI, C.c_int, PI]
    for f in ("cm_afiim_epbicgstab", "cm_pdfiim_epbicgstab", "cm_acifim_pbicgstab"):
        getattr(lib, f).restype = C.c_int
    lib.cm_system_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, PI, PI, PD, PI, PI, PD,
                                     PD, PI, PI, PD, PI, PI, PD, PI, PI, PD, C.c_int, C.c_char]
    lib.cm_system_create.restype = C.c_int
    lib.cm_system_destroy.argtypes = [vp]
    lib.cm_system_set_rhs.argtypes = [vp, PD]
    lib.cm_system_set_x.argtypes = [vp, PD]
    lib.cm_system_get_x.argtypes = [vp, PD]
    lib.cm_system_solve.argtypes = [vp, C.c_double, C.c_int, PD, PI, PI, PD, C.c_int, PI]
    lib.cm_system_spmv.argtypes = [vp, C.c_char, PD, PD]
    lib.cm_system_apply.argtypes = [vp, PD, PD]
    lib.cm_system_nnz.argtypes = [vp, C.c_char]
    lib.cm_system_nnz.restype = C.c_longlong
    lib.cm_system_enable_mixed.argtypes = [vp, C.c_int]
    lib.cm_system_enable_mixed.restype = C.c_int
    lib.cm_system_solve_mixed.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, PD, PI, PI, PI]
    lib.cm_system_solve_mixed.restype = C.c_int
    lib.cm_system_block_set.argtypes = [vp, C.c_int, PD, PD]
    lib.cm_system_block_get.argtypes = [vp, PD]
    lib.cm_system_block_solve.argtypes = [vp, C.c_double, C.c_int, PD, PI, PI]
    lib.cm_system_block_solve_mixed.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, PD, PI,
                                                PI, PI]
    lib.cm_system_block_graphs.argtypes = [vp, C.c_int]
    lib.cm_system_graphs.argtypes = [vp, C.c_int]
    lib.cm_system_graphs.restype = C.c_int
    lib.cm_system_block_release.argtypes = [vp]
    lib.cm_system_dia_info.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_longlong)]
    lib.cm_system_dia_info.restype = C.c_int
    lib.cm_system_block_time_spmm.argtypes = [vp, C.c_char, C.c_int, C.c_int, PD, PD]
    lib.cm_system_block_time_spmm.restype = C.c_int
    lib.cm_batch_solve.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(PD), C.POINTER(PD), PD, PI, C.c_int, PD,
                                   PI, PI]
    lib.cm_batch_solve.restype = C.c_int
    lib.cm_pdfiim_device.argtypes = [C.c_int, PD, PI, PI] + [C.POINTER(vp)] * 7 + [
        C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_char, C.c_char, PD]
    lib.cm_pdfiim_device.restype = C.c_int
    lib.cm_system_autotune.argtypes = [vp, C.c_int, PI]
    lib.cm_system_autotune.restype = C.c_int
    lib.cm_system_spmv_format.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    lib.cm_system_spmv_format.restype = C.c_int
    lib.cm_system_spmv_formats.argtypes = [vp, PI]
    lib.cm_system_spmv_formats.restype = C.c_int
    lib.cm_system_block_autotune.argtypes = [vp, C.c_int, C.c_int, PI]
    lib.cm_system_block_autotune.restype = C.c_int
    for f in ("cm_system_block_set", "cm_system_block_get", "cm_system_block_solve",
              "cm_system_block_solve_mixed", "cm_system_block_graphs", "cm_system_block_release"):
        getattr(lib, f).restype = C.c_int
    PLL = C.POINTER(C.c_longlong)
    lib.cm_server_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_longlong]
    lib.cm_server_destroy.argtypes = [vp]
    lib.cm_server_configure.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_int]
    lib.cm_server_configure.restype = C.c_int
    lib.cm_server_add_system.argtypes = [vp, vp, PI]
    lib.cm_server_submit.argtypes = [vp, C.c_int, C.c_int, PD, C.c_double, C.c_int, PLL]
    lib.cm_server_wait.argtypes = [vp, C.c_longlong, PD, PD, PI, PI, PD]
    lib.cm_server_stats.argtypes = [vp, PLL]
    for f in ("cm_server_create", "cm_server_destroy", "cm_server_add_system", "cm_server_submit",
              "cm_server_wait", "cm_server_stats"):
        getattr(lib, f).restype = C.c_int
    lib.cm_system_profile.argtypes = [vp, C.c_int]
    lib.cm_system_profile.restype = C.c_int
    lib.cm_system_profile_sampled.argtypes = [vp, C.c_int, C.c_int]
    lib.cm_system_profile_sampled.restype = C.c_int
    lib.cm_system_block_dia.argtypes = [vp, C.c_int]
    lib.cm_system_block_dia.restype = C.c_int
    lib.cm_system_kernel_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong), PD, PD]
    lib.cm_system_kernel_stats.restype = C.c_int
    lib.cm_system_kernel_format_bytes.argtypes = [vp, C.c_int, PD]
    lib.cm_system_kernel_format_bytes.restype = C.c_int
    for f in ("cm_system_destroy", "cm_system_set_rhs", "cm_system_set_x", "cm_system_get_x",
              "cm_system_solve", "cm_system_spmv", "cm_system_apply"):
        getattr(lib, f).restype = C.c_int


def _check(rc):
    if rc != CM_OK:
        raise EngineError(rc, solve_lib().cm_last_error().decode(errors="replace"))


def device_count():
    return solve_lib().cm_device_count()


@dataclass
class SolveResult:
    x: np.ndarray
    resval: float
    iter: int
    info: int
    trace: np.ndarray | None = None


class System:
    """A system resident in HBM: A and its factors uploaded once."""

    def __init__(self, variant, A: Csr, F: Factors, a_symmetric=False, nr=1, mode="E"):
        lib = solve_lib()
        self._lib = lib
        self.n = A.n
        self._keep = []

        def csr(m):
            if m is None:
                return None, None, None
            t = _csr_c(m)
            self._keep.append(t)
            return _ip(t[0]), _ip(t[1]), _dp(t[2])

        a = csr(A)
        g = csr(F.G)
        h = csr(F.H)
        l_ = csr(F.L)
        u = csr(F.U)
        idv = np.ascontiguousarray(F.idv, np.float64)
        self._keep.append(idv)
        h_ = C.c_void_p()
        var = VARIANT[variant] if isinstance(variant, str) else int(variant)
        rc = lib.cm_system_create(C.byref(h_), var, A.n, int(bool(a_symmetric)), *a, *g, _dp(idv), *h,
                                  *l_, *u, nr, mode.encode())
        _check(rc)
        self._h = h_
        self._keep = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cm_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rhs(self, b):
        b = np.ascontiguousarray(b, np.float64)
        _check(self._lib.cm_system_set_rhs(self._h, _dp(b)))

    def set_x(self, x=None):
        if x is None:
            _check(self._lib.cm_system_set_x(self._h, None))
        else:
            x = np.ascontiguousarray(x, np.float64)
            _check(self._lib.cm_system_set_x(self._h, _dp(x)))

    def get_x(self):
        x = np.empty(self.n, np.float64)
        _check(self._lib.cm_system_get_x(self._h, _dp(x)))
        return x

    def solve(self, tol=1e-8, nmax=500, trace_cap=0, info_in=0):
        resval, it, info, nt = C.c_double(), C.c_int(), C.c_int(info_in), C.c_int()
        tr = np.zeros(max(trace_cap, 1), np.float64)
        _check(self._lib.cm_system_solve(self._h, tol, nmax, C.byref(resval), C.byref(it), C.byref(info),
                                         _dp(tr) if trace_cap else None, trace_cap, C.byref(nt)))
        trace = tr[:min(nt.value, trace_cap)].copy() if trace_cap else None
        return resval.value, it.value, info.value, trace

    def enable_mixed(self, factor_bits=32):
        _check(self._lib.cm_system_enable_mixed(self._h, factor_bits))

    def solve_mixed(self, factor_bits=32, tol=1e-8, inner_tol=1e-4, nmax=500, max_outer=30):
        """FP64 refinement around the FP32 / FP16-factor inner solve on the
        resident b and x.  Returns (true relative residual, outer steps,
        inner passes, info)."""
        rv, o, it, info = C.c_double(), C.c_int(), C.c_int(), C.c_int()
        _check(self._lib.cm_system_solve_mixed(self._h, factor_bits, tol, inner_tol, nmax, max_outer,
                                               C.byref(rv), C.byref(o), C.byref(it), C.byref(info)))
        return rv.value, o.value, it.value, info.value

    def block_set(self, B, X0=None):
        """Resident block of k right-hand sides (B: n x k), initial guesses
        X0 (n x k) or zeros.  B = None keeps the resident right-hand sides
        and only resets the initial guesses."""
        if B is None:
            if getattr(self, "_bk", 0) < 1:
                raise ValueError("no resident block: pass B")
            px = None if X0 is None else _dp(np.asfortranarray(np.asarray(X0, np.float64).reshape(self.n, self._bk)))
            _check(self._lib.cm_system_block_set(self._h, self._bk, None, px))
            return
        B = np.asfortranarray(np.asarray(B, np.float64).reshape(self.n, -1))
        self._bk = B.shape[1]
        if X0 is None:
            px = None
        else:
            X0 = np.asfortranarray(np.asarray(X0, np.float64).reshape(self.n, self._bk))
            px = _dp(X0)
        _check(self._lib.cm_system_block_set(self._h, self._bk, _dp(B), px))

    def graphs(self, on=True):
        """Single-RHS passes from captured hipGraphs (on) or kernel by kernel."""
        _check(self._lib.cm_system_graphs(self._h, 1 if on else 0))

    def block_dia(self, on=True):
        """Block products in the diagonal format where a matrix suits it (default on)."""
        _check(self._lib.cm_system_block_dia(self._h, 1 if on else 0))

    def block_graphs(self, on=True):
        _check(self._lib.cm_system_block_graphs(self._h, 1 if on else 0))

    def block_time_spmm(self, which="A", bits=64, reps=20):
        """(average microseconds, compulsory bytes) of the current block's
        SpMM with matrix `which` at `bits` precision."""
        us, by = C.c_double(), C.c_double()
        _check(self._lib.cm_system_block_time_spmm(self._h, which.encode(), bits, reps, C.byref(us),
                                                   C.byref(by)))
        return us.value, by.value

    def autotune(self, reps=10):
        """Tile per resident matrix (A, G, H, GT, L, U; 0 = absent)."""
        t = np.zeros(6, np.int32)
        _check(self._lib.cm_system_autotune(self._h, reps, _ip(t)))
        return dict(zip(("A", "G", "H", "GT", "L", "U"), (int(v) for v in t)))

    SPMV_FORMATS = ("staged", "staged16", "slot", "slot16", "dia", "dia2")
    MATRICES = ("A", "G", "H", "GT", "L", "U")

    def spmv_format(self, bits, which, fmt):
        """SpMV format of one resident copy: bits 64/32/16, which in MATRICES,
        fmt in SPMV_FORMATS (EngineError when not available)."""
        _check(self._lib.cm_system_spmv_format(self._h, int(bits), self.MATRICES.index(which),
                                               self.SPMV_FORMATS.index(fmt)))

    def dia_info(self, which="A"):
        """(offsets of the stored diagonals, remainder entries) of matrix
        'A', 'G', 'H', 'GT', 'L' or 'U' (the diagonal index is built if
        absent; no offsets when no diagonal is held by 60% of the rows)."""
        q = ("A", "G", "H", "GT", "L", "U").index(which)
        nd, nrem = C.c_int(), C.c_longlong()
        off = np.zeros(32, np.int32)
        _check(self._lib.cm_system_dia_info(self._h, q, C.byref(nd), _ip(off), C.byref(nrem)))
        return off[:nd.value].copy(), int(nrem.value)

    def spmv_formats(self):
        """{bits: {matrix: format}} of the resident copies."""
        f = np.zeros(18, np.int32)
        _check(self._lib.cm_system_spmv_formats(self._h, _ip(f)))
        out = {}
        for p, bits in enumerate((64, 32, 16)):
            d = {m: self.SPMV_FORMATS[int(f[6 * p + q])] for q, m in enumerate(self.MATRICES) if f[6 * p + q] >= 0}
            if d:
                out[bits] = d
        return out

    def block_autotune(self, bits=64, reps=10):
        """Short-row limit of the current block's SpMM bins for A, G, H."""
        t = np.zeros(3, np.int32)
        _check(self._lib.cm_system_block_autotune(self._h, bits, reps, _ip(t)))
        return dict(zip(("A", "G", "H"), (int(v) for v in t)))

    def block_release(self):
        _check(self._lib.cm_system_block_release(self._h))

    def block_get(self):
        X = np.empty((self.n, self._bk), np.float64, order="F")
        _check(self._lib.cm_system_block_get(self._h, _dp(X)))
        return X

    def block_solve(self, tol=1e-8, nmax=500, info_in=0):
        """FP64 block solve; per-column (resval, iter, info) arrays.  info
        starts at info_in (the engine leaves it so when nmax < 0)."""
        k = self._bk
        rv, it, info = np.empty(k), np.empty(k, np.int32), np.full(k, info_in, np.int32)
        _check(self._lib.cm_system_block_solve(self._h, tol, nmax, _dp(rv), _ip(it), _ip(info)))
        return rv, it, info

    def block_solve_mixed(self, factor_bits=32, tol=1e-8, inner_tol=1e-4, nmax=500, max_outer=30):
        """Per-column FP64 refinement around the low-precision inner block
        solve.  Returns (resval[k], outer steps, inner passes, info[k])."""
        k = self._bk
        rv, info = np.empty(k), np.empty(k, np.int32)
        o, inner = C.c_int(), C.c_int()
        _check(self._lib.cm_system_block_solve_mixed(self._h, factor_bits, tol, inner_tol, nmax, max_outer,
                                                     _dp(rv), C.byref(o), C.byref(inner), _ip(info)))
        return rv, o.value, inner.value, info

    def spmv(self, which, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n, np.float64)
        _check(self._lib.cm_system_spmv(self._h, which.encode(), _dp(x), _dp(y)))
        return y

    def apply(self, v):
        v = np.ascontiguousarray(v, np.float64)
        out = np.empty(self.n, np.float64)
        _check(self._lib.cm_system_apply(self._h, _dp(v), _dp(out)))
        return out

    def nnz(self, which):
        return self._lib.cm_system_nnz(self._h, which.encode())

    KERNEL_CATS = ("A", "F1", "F2", "LU", "VEC", "SETUP", "FIN")

    def profile(self, on=True, cats=None, sample=0):
        """Bracket launches with HIP events: all categories, or only `cats`;
        sample > 0 brackets only the first `sample` passes of each solve (the
        rest replay from graphs)."""
        if not on:
            mask = 0
        elif cats is None:
            mask = (1 << len(self.KERNEL_CATS)) - 1
        else:
            mask = sum(1 << self.KERNEL_CATS.index(c) for c in cats)
        _check(self._lib.cm_system_profile_sampled(self._h, mask, int(sample)))

    def kernel_stats(self):
        """{category: (launches, total_ms, total_bytes)} for the profiled solves;
        total_bytes is the format-independent algorithmic traffic (CSR figure
        for the SpMV categories)."""
        out = {}
        for c, name in enumerate(self.KERNEL_CATS):
            n, ms, by = C.c_longlong(), C.c_double(), C.c_double()
            _check(self._lib.cm_system_kernel_stats(self._h, c, C.byref(n), C.byref(ms), C.byref(by)))
            out[name] = (n.value, ms.value, by.value)
        return out

    def kernel_format_bytes(self):
        """{category: bytes the counted launches touch once in the storage
        format they ran in} for the profiled solves."""
        out = {}
        for c, name in enumerate(self.KERNEL_CATS):
            by = C.c_double()
            _check(self._lib.cm_system_kernel_format_bytes(self._h, c, C.byref(by)))
            out[name] = by.value
        return out


def afiim_epbicgstab(A: Csr, b, F: Factors, x0=None, tol=1e-8, nmax=500, verbose=0) -> SolveResult:
    """Drop-in AFIIM EPBiCGSTAB through the C ABI."""
    lib = solve_lib()
    ptr, idx, val = _csr_c(A)
    gp, gi, gv = _csr_c(F.G)
    hp, hi, hv = _csr_c(F.H)
    idv = np.ascontiguousarray(F.idv, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    resval, it, info = C.c_double(), C.c_int(), C.c_int(0)
    _check(lib.cm_afiim_epbicgstab(A.n, _dp(val), _ip(idx), _ip(ptr), _dp(b), tol, nmax, _dp(gv), _ip(gi),
                                   _ip(gp), _dp(idv), _dp(hv), _ip(hi), _ip(hp), _dp(x), C.byref(resval),
                                   C.byref(it), verbose, C.byref(info)))
    return SolveResult(x, resval.value, it.value, info.value)


def pdfiim_epbicgstab(A: Csr, b, F: Factors, x0=None, tol=1e-8, nmax=1200, verbose=0,
                      a_symmetric=False) -> SolveResult:
    lib = solve_lib()
    ptr, idx, val = _csr_c(A)
    gp, gi, gv = _csr_c(F.G)
    if F.H is not None:
        hp, hi, hv = _csr_c(F.H)
    else:
        hp = hi = hv = None
    idv = np.ascontiguousarray(F.idv, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    resval, it, info = C.c_double(), C.c_int(), C.c_int(0)
    _check(lib.cm_pdfiim_epbicgstab(A.n, int(bool(a_symmetric)), _ip(ptr), _ip(idx), _dp(val), _dp(b), tol,
                                    nmax, _ip(gp), _ip(gi), _dp(gv), _dp(idv), _ip(hp), _ip(hi), _dp(hv),
                                    _dp(x), C.byref(resval), C.byref(it), verbose, C.byref(info)))
    return SolveResult(x, resval.value, it.value, info.value)


def acifim_pbicgstab(A: Csr, b, F: Factors, nr=1, mode="E", x0=None, tol=1e-8, nmax=2000,
                     verbose=0) -> SolveResult:
    lib = solve_lib()
    ptr, idx, val = _csr_c(A)
    gp, gi, gv = _csr_c(F.G)
    hp, hi, hv = _csr_c(F.H)
    lp, li, lv = _csr_c(F.L)
    up, ui, uv = _csr_c(F.U)
    idv = np.ascontiguousarray(F.idv, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    resval, it, info = C.c_double(), C.c_int(), C.c_int(0)
    _check(lib.cm_acifim_pbicgstab(A.n, _dp(val), _ip(idx), _ip(ptr), _dp(b), tol, nmax, _dp(gv), _ip(gi),
                                   _ip(gp), _dp(idv), _dp(hv), _ip(hi), _ip(hp), _dp(lv), _ip(li), _ip(lp),
                                   _dp(uv), _ip(ui), _ip(up), nr, mode.encode(), _dp(x), C.byref(resval),
                                   C.byref(it), verbose, C.byref(info)))
    return SolveResult(x, resval.value, it.value, info.value)


@dataclass
class Served:
    X: np.ndarray
    resval: np.ndarray
    info: np.ndarray
    outer: int
    inner: int
    batch_cols: int
    queue_ms: float
    service_ms: float
    latency_ms: float


class Server:
    """Request loop over resident systems with continuous batching (slots
    worker streams; requests of one system merged into block solves of up to
    16 columns; admission by HBM footprint).  precision 64: FP64 BiCGSTAB per
    column; 32 / 16: FP64 refinement around the FP32 inner solve with FP32 or
    FP16 factors.  Registered systems must outlive the server and must not be
    used directly while it runs."""

    def __init__(self, slots=4, precision=32, hbm_budget=0, max_cols=16, inner_tol=1e-4, max_outer=30,
                 small_n=0):
        self._lib = solve_lib()
        h = C.c_void_p()
        _check(self._lib.cm_server_create(C.byref(h), slots, precision, int(hbm_budget)))
        self._h = h
        _check(self._lib.cm_server_configure(h, max_cols, inner_tol, max_outer, small_n))
        self._n = {}
        self._k = {}
        self._systems = []

    def add_system(self, system: "System") -> int:
        sid = C.c_int()
        _check(self._lib.cm_server_add_system(self._h, system._h, C.byref(sid)))
        self._n[sid.value] = system.n
        self._systems.append(system)
        return sid.value

    def submit(self, sys_id, B, tol=1e-8, nmax=500) -> int:
        if sys_id not in self._n:  # let the engine report the unknown id
            rid = C.c_longlong()
            _check(self._lib.cm_server_submit(self._h, sys_id, 1, _dp(np.zeros(1)), tol, nmax, C.byref(rid)))
        n = self._n[sys_id]
        B = np.asfortranarray(np.asarray(B, np.float64).reshape(n, -1))
        rid = C.c_longlong()
        _check(self._lib.cm_server_submit(self._h, sys_id, B.shape[1], _dp(B), tol, nmax, C.byref(rid)))
        self._k[rid.value] = (n, B.shape[1])
        return rid.value

    def wait(self, rid) -> Served:
        n, k = self._k.pop(rid)
        X = np.empty((n, k), np.float64, order="F")
        rv, info = np.empty(k), np.empty(k, np.int32)
        cnt, tms = np.empty(3, np.int32), np.empty(3)
        _check(self._lib.cm_server_wait(self._h, rid, _dp(X), _dp(rv), _ip(info), _ip(cnt), _dp(tms)))
        return Served(X, rv, info, int(cnt[0]), int(cnt[1]), int(cnt[2]), float(tms[0]), float(tms[1]),
                      float(tms[2]))

    def stats(self):
        s = (C.c_longlong * 8)()
        _check(self._lib.cm_server_stats(self._h, s))
        return dict(zip(("batches", "columns", "max_running", "evictions", "bytes", "budget",
                         "persistent_batches", "persistent_columns"), list(s)))

    def close(self):
        if self._h is not None:
            self._lib.cm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_solve(items, restart_true=False, info_in=0):
    """Persistent batched FP64 solves of small resident systems in one launch.
    items: list of (System, b, x0 or None, tol, nmax).  Returns a list of
    SolveResult (x, resval, iter, info); info starts at info_in (the engine
    leaves it so for an item with nmax < 0)."""
    lib = solve_lib()
    m = len(items)
    hs = (C.c_void_p * m)(*[it[0]._h for it in items])
    bs = [np.ascontiguousarray(it[1], np.float64) for it in items]
    xs = [np.zeros(len(b)) if it[2] is None else np.array(it[2], np.float64) for it, b in zip(items, bs)]
    pb = (PD * m)(*[_dp(b) for b in bs])
    px = (PD * m)(*[_dp(x) for x in xs])
    tol = np.array([it[3] for it in items], np.float64)
    nmax = np.array([it[4] for it in items], np.int32)
    rv, it_, info = np.empty(m), np.empty(m, np.int32), np.full(m, info_in, np.int32)
    _check(lib.cm_batch_solve(m, hs, pb, px, _dp(tol), _ip(nmax), 1 if restart_true else 0, _dp(rv), _ip(it_),
                              _ip(info)))
    return [SolveResult(xs[i], float(rv[i]), int(it_[i]), int(info[i])) for i in range(m)]
