"""Host-side pieces: workload generators, Matrix Market ingest, C ABI exports.

No GPU needed: the ABI test loads libcmsolve.so and checks that every function
include/*.h declares is exported, and that the entry points report a missing
device as an error instead of computing anything on the CPU.
"""
import os
import re

import numpy as np
import pytest

import cmamd
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_SHERMAN = "/root/reference/AFIIM/example/sherman2.mtx"


# ----------------------------------------------------------------- generators

@pytest.mark.parametrize("dim,N", [(2, 100), (3, 128), (2, 7), (3, 5)])
def test_convdiff_sizes(dim, N):
    n, nnz = N ** dim, (2 * dim + 1) * N ** dim - 2 * dim * N ** (dim - 1)
    if n > 100000:  # 128^3: only the size query
        import ctypes as C
        nn, zz = C.c_int(), C.c_longlong()
        assert cmamd.host_lib().cm_gen_convdiff(dim, N, 1.0, 2, C.byref(nn), C.byref(zz), None, None, None) == 0
        assert (nn.value, zz.value) == (n, nnz) == (2097152, 14581760)
        return
    A = cmamd.gen_convdiff(dim, N, 1.0, seed=3)
    assert (A.n, A.nnz) == (n, nnz)


@pytest.mark.parametrize("dim,pe", [(2, 0.5), (3, 1.0), (3, 5.0)])
def test_convdiff_is_nonsymmetric_m_matrix(dim, pe):
    A = cmamd.gen_convdiff(dim, 9, pe, seed=5)
    S = A.to_scipy()
    d = S.diagonal()
    off = S - __import__("scipy.sparse", fromlist=["diags"]).diags(d)
    assert np.all(d > 0) and off.max() <= 0
    rows = np.asarray(S.sum(axis=1)).ravel()
    assert np.all(rows > -1e-12 * d)  # weakly diagonally dominant
    assert np.any(rows > 1e-6)  # strictly on the boundary
    if pe > 0:
        assert abs(S - S.T).max() > 0.1
    for i in range(A.n):
        assert np.all(np.diff(A.idx[A.ptr[i]:A.ptr[i + 1]]) > 0)


def test_convdiff_symmetric_without_convection():
    S = cmamd.gen_convdiff(3, 6, 0.0, seed=1).to_scipy()
    assert abs(S - S.T).max() == 0.0


def test_generators_deterministic():
    a = cmamd.gen_convdiff(3, 10, 1.0, seed=7)
    b = cmamd.gen_convdiff(3, 10, 1.0, seed=7)
    c = cmamd.gen_convdiff(3, 10, 1.0, seed=8)
    assert np.array_equal(a.val, b.val) and not np.array_equal(a.val, c.val)
    u1, u2 = cmamd.gen_uniform(3, 1000), cmamd.gen_uniform(3, 1000)
    assert np.array_equal(u1, u2) and u1.min() >= -1 and u1.max() <= 1 and abs(u1.mean()) < 0.1


def test_stencil27_structure():
    N, lo, hi = 20, 200, 600
    A = cmamd.gen_stencil27(N, 0.05, lo, hi, seed=9)
    S = A.to_scipy()
    lens = np.diff(A.ptr)
    base = (3 * N - 2) ** 3
    extra = A.nnz - base
    heavy = lens > 27
    assert heavy.sum() > 0 and abs(heavy.mean() - 0.05) < 0.02
    assert np.all(lens[heavy] <= 27 + hi) and np.all(lens[heavy] >= 8 + lo)
    assert extra > heavy.sum() * lo // 2
    d = S.diagonal()
    offabs = np.asarray(abs(S).sum(axis=1)).ravel() - np.abs(d)
    assert np.allclose(d, 1.05 * offabs, rtol=1e-12)
    for i in range(0, A.n, 97):
        c = A.idx[A.ptr[i]:A.ptr[i + 1]]
        assert np.all(np.diff(c) > 0)


# --------------------------------------------------------------- Matrix Market

def test_read_sherman2_matches_golden(tmp_path):
    z = np.load(os.path.join(GOLDEN, "sherman2.npz"))
    d = {k: z[k] for k in ("ptr", "idx", "val")}
    # write the golden matrix back out in coordinate form and read it
    p = tmp_path / "s.mtx"
    n = len(d["ptr"]) - 1
    with open(p, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% test\n")
        f.write(f"{n} {n} {len(d['val'])}\n")
        for i in range(n):
            for k in range(d["ptr"][i], d["ptr"][i + 1]):
                f.write("%d %d %.17g\n" % (i + 1, d['idx'][k] + 1, d['val'][k]))
    A, sym = cmamd.read_mtx(str(p))
    assert not sym
    assert np.array_equal(A.ptr, d["ptr"]) and np.array_equal(A.idx, d["idx"]) and np.array_equal(A.val, d["val"])


@pytest.mark.skipif(not (os.path.exists(REF_SHERMAN) and oracle.ref_available("afiim")),
                    reason="reference tree or its reader build not present")
def test_read_matches_reference_reader():
    A, _ = cmamd.read_mtx(REF_SHERMAN)
    R = oracle.ref_read_coeff_matrix(REF_SHERMAN)
    assert np.array_equal(A.ptr, R.ptr) and np.array_equal(A.idx, R.idx) and np.array_equal(A.val, R.val)
    b = cmamd.read_mtx_vector(REF_SHERMAN.replace("sherman2.mtx", "sherman2_rhs1.mtx"))
    assert np.array_equal(b, oracle.ref_read_rhs_vector(REF_SHERMAN.replace("sherman2.mtx", "sherman2_rhs1.mtx")))


def test_read_symmetric_modes(tmp_path):
    p = tmp_path / "sym.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n1 1 4.0\n2 1 -1.0\n"
                 "3 2 -2.0\n3 3 5.0\n")
    A, sym = cmamd.read_mtx(str(p), symmetric_mode=0)
    assert sym
    assert np.array_equal(A.to_scipy().toarray(), np.array([[4, -1, 0], [-1, 0, -2], [0, -2, 5.0]]))
    U, _ = cmamd.read_mtx(str(p), symmetric_mode=1)
    assert np.array_equal(U.to_scipy().toarray(), np.array([[4, -1, 0], [0, 0, -2], [0, 0, 5.0]]))


def test_read_rejects_unsupported(tmp_path):
    p = tmp_path / "pat.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 1\n")
    with pytest.raises(cmamd.EngineError):
        cmamd.read_mtx(str(p))
    with pytest.raises(cmamd.EngineError):
        cmamd.read_mtx(str(tmp_path / "missing.mtx"))


# --------------------------------------------------------------------- C ABI

def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cm_\w+)\s*\(", txt)))


@pytest.mark.parametrize("header,lib", [("cm_solve.h", "solve"), ("cm_serve.h", "solve"), ("cm_host.h", "host")])
def test_library_exports_every_declared_symbol(header, lib):
    L = cmamd.solve_lib() if lib == "solve" else cmamd.host_lib()
    names = _declared(header)
    assert len(names) >= 5
    for name in names:
        assert hasattr(L, name), f"{name} declared in {header} but not exported"


def test_solve_entry_points_fail_loudly_without_gpu():
    if cmamd.device_count() > 0:
        pytest.skip("a GPU is visible")
    A = cmamd.gen_convdiff(2, 6, 1.0, seed=1)
    F = cmamd.afiim(A, 0.1)
    with pytest.raises(cmamd.EngineError) as e:
        cmamd.afiim_epbicgstab(A, np.ones(A.n), F)
    assert e.value.code == cmamd.CM_ERR_NODEV
    with pytest.raises(cmamd.EngineError) as e:
        cmamd.System("afiim", A, F)
    assert e.value.code == cmamd.CM_ERR_NODEV
    with pytest.raises(cmamd.EngineError) as e:
        cmamd.Server(slots=2, precision=32)
    assert e.value.code == cmamd.CM_ERR_NODEV
