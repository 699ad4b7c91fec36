"""PDFIIM construction on the GPU (cm_pdfiim_device) against the host
producer (cm_pdfiim, host/pdfiim.cpp) and against what the mathematics fixes.

The GPU builds every row with the host's operation sequence (the same
pattern order, the same pivots, column-oriented substitutions, no fused
multiply-adds), so G, H and D must equal the host's bit for bit: on 2D/3D
convection-diffusion, on the reference's fixture sherman2, for every filter
type, with warpsize truncation, for the 'L'/'U' patterns, and on rows whose
patterns outgrow the LDS tiers (33-64 indices, then up to 256 in global
memory).  Independently of the host
code: with ftol = 0 on a dense matrix nothing is dropped and G D H = A^{-1};
on a tridiagonal matrix g, h and D have closed forms.  Patterns beyond 256
indices are refused with CM_ERR_UNSUPPORTED."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import cmamd

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _dense(m):
    return m.to_scipy().toarray()


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def _csr_same(a, b):
    return (a.n == b.n and np.array_equal(a.ptr, b.ptr) and np.array_equal(a.idx, b.idx)
            and _same(a.val, b.val))


def _check_bitwise(A, *args, **kw):
    Fh = cmamd.pdfiim(A, *args, **kw)
    Fd, ms = cmamd.pdfiim_device(A, *args, **kw)
    assert _csr_same(Fd.G, Fh.G), "G differs from the host producer"
    assert _csr_same(Fd.H, Fh.H), "H differs from the host producer"
    assert _same(Fd.idv, Fh.idv), "D differs from the host producer"
    assert ms > 0.0
    return Fd, ms


def _sherman2():
    d = np.load(os.path.join(ROOT, "tests", "golden", "sherman2.npz"))
    return cmamd.Csr(len(d["ptr"]) - 1, d["ptr"], d["idx"], d["val"])


@pytest.mark.parametrize("case", ["cd2-p1", "cd3-p1", "cd3-pe5", "cd2-wide"])
def test_device_factors_match_host(case):
    A = {"cd2-p1": lambda: cmamd.gen_convdiff(2, 40, 1.0, seed=2),
         "cd3-p1": lambda: cmamd.gen_convdiff(3, 20, 1.0, seed=2),
         "cd3-pe5": lambda: cmamd.gen_convdiff(3, 16, 5.0, seed=4),
         "cd2-wide": lambda: cmamd.gen_convdiff(2, 60, 2.0, seed=9)}[case]()
    print(f"pdfiim device {case} n={A.n}", flush=True)
    _check_bitwise(A, 2, 2, 1, 0.05, -1, "I", "B")


@pytest.mark.parametrize("eps", [0.001, 0.01, 0.05])
def test_device_factors_sherman2(eps):
    _check_bitwise(_sherman2(), 2, 2, 1, eps, -1, "I", "B")


@pytest.mark.parametrize("filt", ["1", "2", "I"])
def test_device_filter_types(filt):
    A = cmamd.gen_convdiff(3, 12, 2.0, seed=6)
    _check_bitwise(A, 2, 2, 1, 0.01, -1, filt, "B")


def test_device_warpsize_and_patterns():
    A = cmamd.gen_convdiff(3, 12, 1.0, seed=7)
    _check_bitwise(A, 2, 2, 1, 0.01, 8, "I", "B")
    _check_bitwise(A, 2, 2, 1, 0.01, 1, "I", "B")
    for patt in ("L", "U"):
        _check_bitwise(A, 2, 1, 1, 0.01, -1, "I", patt)


def test_device_second_tier_rows():
    # patterns of up to 56 indices: rows go through the 64-index tier
    A = cmamd.gen_convdiff(2, 40, 1.0, seed=2)
    _check_bitwise(A, 3, 3, 2, 0.001, -1, "I", "B")
    _check_bitwise(A, 4, 2, 2, 0.001, -1, "2", "B")


def test_device_third_tier_rows():
    # patterns of up to 115 indices: the rows past 64 build in global memory
    A = cmamd.gen_convdiff(3, 14, 1.0, seed=2)
    _check_bitwise(A, 3, 2, 2, 0.01, -1, "I", "B")


def test_device_full_pattern_is_exact_inverse():
    # dense A, ftol = 0: row i's pattern is {0..i-1}, nothing is dropped, each
    # small solve is exact, so G D^{-1} H = A^{-1} (rows past 32 use tier 2)
    rng = np.random.default_rng(3)
    n = 48
    M = rng.uniform(-1, 1, (n, n)) + np.diag(np.full(n, 2.0 * n))
    A = cmamd.Csr.from_scipy(sp.csr_matrix(M))
    F, _ = _check_bitwise(A, 1, 1, 1, 0.0, -1, "I", "B")
    Minv = _dense(F.G) @ np.diag(F.idv) @ _dense(F.H)
    assert np.linalg.norm(Minv @ M - np.eye(n)) < 1e-10


def test_device_tridiagonal_closed_form():
    n = 300
    rng = np.random.default_rng(5)
    lo, di, up = rng.uniform(-1, 1, n - 1), rng.uniform(3, 4, n), rng.uniform(-1, 1, n - 1)
    T = np.diag(di) + np.diag(lo, -1) + np.diag(up, 1)
    A = cmamd.Csr.from_scipy(sp.csr_matrix(T))
    F, _ = cmamd.pdfiim_device(A, 1, 1, 1, 0.0, -1, "I", "B")
    G, H = _dense(F.G), _dense(F.H)
    assert F.idv[0] == 1.0 / T[0, 0]
    for i in range(1, n):
        g = -T[i - 1, i] / T[i - 1, i - 1]
        h = -T[i, i - 1] / T[i - 1, i - 1]
        s = h * (T[i - 1, i - 1] * g + T[i - 1, i]) + (T[i, i - 1] * g + T[i, i])
        assert G[i - 1, i] == pytest.approx(g, rel=1e-15)
        assert H[i, i - 1] == pytest.approx(h, rel=1e-15)
        assert F.idv[i] == pytest.approx(1.0 / s, rel=1e-14)


def test_device_refuses_long_patterns_and_bad_arguments():
    A = cmamd.gen_convdiff(3, 14, 1.0, seed=2)
    with pytest.raises(cmamd.EngineError) as e:
        cmamd.pdfiim_device(A, 3, 3, 2, 0.001, -1, "I", "B")  # patterns of up to 287 indices
    assert e.value.code == cmamd.CM_ERR_UNSUPPORTED
    with pytest.raises(cmamd.EngineError) as e:
        cmamd.pdfiim_device(A, 1, 2, 1, 0.05, -1, "I", "L")
    assert e.value.code == cmamd.CM_ERR_UNSUPPORTED
    with pytest.raises(cmamd.EngineError) as e:
        cmamd.pdfiim_device(A, 2, 2, 1, 0.05, 0, "I", "B")
    assert e.value.code == cmamd.CM_ERR_ARG


def test_device_factors_solve_like_host():
    # the solve on device-built factors is the solve on host-built ones
    A = cmamd.gen_convdiff(3, 20, 1.0, seed=5)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(6, A.n))
    F, _ = cmamd.pdfiim_device(A, 2, 2, 1, 0.05, -1, "I", "B")
    r1 = cmamd.pdfiim_epbicgstab(A, b, F, nmax=1200)
    r2 = cmamd.pdfiim_epbicgstab(A, b, cmamd.pdfiim(A, 2, 2, 1, 0.05, -1, "I", "B"), nmax=1200)
    assert r1.info == r2.info == 0 and r1.iter == r2.iter
    assert np.array_equal(r1.x, r2.x)
