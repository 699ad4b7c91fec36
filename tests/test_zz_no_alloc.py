"""cm_batch_solve_noalloc (the batched solve without any allocation) against
cm_batch_solve on the same items, through the C ABI: two calls in a row, the
second with fewer and smaller items, so the second runs on the array the
first left behind.  Both run the same kernel on the same resident systems and
each item is solved alone in its workgroup, so the results must agree bit for
bit.  Device vectors for the no-allocation call come from hipMalloc in the
same HIP runtime the engine links (test plumbing only)."""
import ctypes as C

import numpy as np
import pytest

import cmamd

pytestmark = pytest.mark.gpu

PD = C.POINTER(C.c_double)


def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


class _Dev:
    """One device vector of doubles (hipMalloc'ed by the test)."""

    def __init__(self, hip, host):
        self.hip = hip
        self.n = len(host)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 8 * self.n) == 0
        h = np.ascontiguousarray(host, np.float64)
        assert hip.hipMemcpy(self.p, h.ctypes.data, 8 * self.n, H2D) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 8 * self.n, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _systems():
    out = []
    for d, N, pe, seed in ((2, 60, 0.5, 81), (3, 14, 3.0, 82), (2, 90, 1.0, 83), (3, 9, 0.2, 84)):
        A = cmamd.gen_convdiff(d, N, pe, seed=seed)
        F = cmamd.afiim(A, 0.1)
        out.append((A, F, cmamd.System("afiim", A, F)))
    return out


def _items(systems, which, seed):
    rng = np.random.default_rng(seed)
    items = []
    for q in which:
        A, F, h = systems[q]
        b = cmamd.csr_mv_host(A, rng.uniform(-1, 1, A.n))
        x0 = None if rng.random() < 0.5 else rng.uniform(-1, 1, A.n)
        items.append((h, b, x0, 1e-8, 1000))
    items.append((systems[which[0]][2], items[0][1], None, 1e-8, 3))  # NMAX reached
    return items


def _noalloc(hip, items):
    lib = cmamd.solve_lib()
    lib.cm_batch_solve_noalloc.restype = C.c_int
    m = len(items)
    dB = [_Dev(hip, it[1]) for it in items]
    dX = [_Dev(hip, np.zeros(len(it[1])) if it[2] is None else it[2]) for it in items]
    hs = (C.c_void_p * m)(*[it[0]._h for it in items])
    pb = (C.c_void_p * m)(*[d.p.value for d in dB])
    px = (C.c_void_p * m)(*[d.p.value for d in dX])
    tol = np.array([it[3] for it in items], np.float64)
    nmax = np.array([it[4] for it in items], np.int32)
    rv, itn, info = np.empty(m), np.empty(m, np.int32), np.full(m, -7, np.int32)
    rc = lib.cm_batch_solve_noalloc(m, hs, pb, px, tol.ctypes.data_as(PD), nmax.ctypes.data_as(C.POINTER(C.c_int)),
                                    0, rv.ctypes.data_as(PD), itn.ctypes.data_as(C.POINTER(C.c_int)),
                                    info.ctypes.data_as(C.POINTER(C.c_int)), None)
    assert rc == 0, cmamd.solve_lib().cm_last_error()
    xs = [d.get() for d in dX]
    for d in dB + dX:
        d.free()
    return xs, rv, itn, info


def test_noalloc_batch_matches_batch_twice():
    hip = _hip()
    systems = _systems()
    for call, (which, seed) in enumerate((([0, 1, 2, 3, 0, 2, 1, 3, 2], 5), ([3, 1, 3], 6))):
        items = _items(systems, which, seed)
        print(f"no-alloc batch call {call}: {len(items)} items, n = {[it[0].n for it in items]}", flush=True)
        ref = cmamd.batch_solve(items)
        xs, rv, itn, info = _noalloc(hip, items)
        for i, r in enumerate(ref):
            assert (int(itn[i]), int(info[i])) == (r.iter, r.info), (call, i)
            assert rv[i] == r.resval, (call, i)
            assert np.array_equal(xs[i], r.x), (call, i)
        assert int(info[-1]) == 3 and int(itn[-1]) == 3
        assert np.all(info[:-1] == 0)
    for A, F, h in systems:
        h.close()


def test_noalloc_batch_rejects_oversize():
    lib = cmamd.solve_lib()
    A = cmamd.gen_convdiff(2, 8, 0.5, seed=1)
    h = cmamd.System("afiim", A, cmamd.afiim(A, 0.1))
    m = 4097
    hs = (C.c_void_p * m)(*([h._h] * m))
    dummy = (C.c_void_p * m)(*([1] * m))  # never dereferenced: rejected before any copy
    tol = np.full(m, 1e-8)
    nmax = np.full(m, 10, np.int32)
    rc = lib.cm_batch_solve_noalloc(m, hs, dummy, dummy, tol.ctypes.data_as(PD),
                                    nmax.ctypes.data_as(C.POINTER(C.c_int)), 0, None, None, None, None)
    assert rc == cmamd.CM_ERR_ARG
    h.close()
