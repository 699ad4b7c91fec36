"""Error behaviour at the C ABI on a GPU: malformed inputs are rejected
before any solve kernel reads them, with the documented codes (CM_ERR_ARG =
-2, CM_ERR_UNSUPPORTED = -3) and a message; the engine stays usable
afterwards.  Row pointers are checked on the host; column ranges on the host
where the host reads the columns itself (symmetric expansion, transposes,
level schedules), else by a check kernel on the device before the system is
returned.
The reference itself has no error returns (its only failure is
std::bad_alloc); the drop-in C++ header turns these codes into
std::runtime_error."""
import numpy as np
import pytest

import cmamd

pytestmark = pytest.mark.gpu


def _small():
    A = cmamd.gen_convdiff(2, 20, 0.5, seed=3)
    return A, cmamd.afiim(A, 0.1)


def test_bad_csr_rejected():
    A, F = _small()
    bad = cmamd.Csr(A.n, A.ptr.copy(), A.idx.copy(), A.val.copy())
    bad.idx[5] = A.n  # column out of range
    with pytest.raises(cmamd.EngineError) as e:
        cmamd.System("afiim", bad, F)
    assert e.value.code == -2
    bad = cmamd.Csr(A.n, A.ptr.copy(), A.idx.copy(), A.val.copy())
    bad.ptr[3], bad.ptr[4] = bad.ptr[4], bad.ptr[3]  # row pointer not monotone
    with pytest.raises(cmamd.EngineError) as e:
        cmamd.System("afiim", bad, F)
    assert e.value.code == -2
    with pytest.raises(cmamd.EngineError):
        cmamd.System("acifim", A, cmamd.acifim(A, 0.01, 0.1), mode="Q")
    # the engine still works
    s = cmamd.System("afiim", A, F)
    s.set_rhs(cmamd.csr_mv_host(A, np.ones(A.n)))
    s.set_x(None)
    assert s.solve(tol=1e-8, nmax=200)[2] == 0
    s.close()


def _bad_col(M, value, at=5):
    bad = cmamd.Csr(M.n, M.ptr.copy(), M.idx.copy(), M.val.copy())
    bad.idx[at] = value
    return bad


def test_bad_columns_in_every_matrix_rejected():
    A, F = _small()
    b = cmamd.csr_mv_host(A, np.ones(A.n))
    cases = [
        lambda: cmamd.System("afiim", _bad_col(A, -1), F),
        lambda: cmamd.System("afiim", A, cmamd.Factors(_bad_col(F.G, A.n), F.idv, F.H)),
        lambda: cmamd.System("afiim", A, cmamd.Factors(F.G, F.idv, _bad_col(F.H, 10 * A.n))),
        lambda: cmamd.afiim_epbicgstab(_bad_col(A, A.n + 3), b, F, nmax=50),
    ]
    Fc = cmamd.acifim(A, 0.01, 0.1)
    for mode, nr in (("C", 2), ("I", 1), ("J", 1)):  # L checked on the device (C) or on the host (I, J)
        Fb = cmamd.Factors(Fc.G, Fc.idv, Fc.H, _bad_col(Fc.L, A.n), Fc.U)
        cases.append(lambda Fb=Fb, mode=mode, nr=nr: cmamd.System("acifim", A, Fb, nr=nr, mode=mode))
    # symmetric PDFIIM: the stored triangle is expanded on the host
    Fp = cmamd.pdfiim(A, 1, 1, 1, 0.05, -1, "I", "B")
    cases.append(lambda: cmamd.System("pdfiim", _bad_col(A, A.n), Fp, a_symmetric=True))
    for make in cases:
        with pytest.raises(cmamd.EngineError) as e:
            make()
        assert e.value.code == -2 and "out of range" in str(e.value)
    s = cmamd.System("afiim", A, F)
    s.set_rhs(b)
    s.set_x(None)
    assert s.solve(tol=1e-8, nmax=200)[2] == 0
    s.close()


def test_block_and_batch_limits():
    A, F = _small()
    s = cmamd.System("afiim", A, F)
    with pytest.raises(cmamd.EngineError) as e:
        s.block_set(np.ones((A.n, 17)))
    assert e.value.code == -2
    s.block_set(np.ones((A.n, 2)))
    with pytest.raises(cmamd.EngineError):  # mixed precision not enabled
        s.block_solve_mixed(32)
    Ac = cmamd.gen_convdiff(2, 20, 2.0, seed=4)
    sc = cmamd.System("acifim", Ac, cmamd.acifim(Ac, 0.01, 0.1), nr=2, mode="C")
    with pytest.raises(cmamd.EngineError) as e:
        cmamd.batch_solve([(sc, np.ones(Ac.n), None, 1e-8, 100)])
    assert e.value.code == -3
    sc.close()
    s.close()


def test_server_rejects():
    A, F = _small()
    s = cmamd.System("afiim", A, F)
    srv = cmamd.Server(slots=1, precision=32)
    sid = srv.add_system(s)
    with pytest.raises(cmamd.EngineError):
        srv.add_system(s)  # registered twice
    with pytest.raises(cmamd.EngineError):
        srv.submit(sid + 1, np.ones(A.n))
    with pytest.raises(cmamd.EngineError):
        srv.submit(sid, np.ones((A.n, 17)))
    r = srv.wait(srv.submit(sid, cmamd.csr_mv_host(A, np.ones(A.n))))
    assert r.info[0] == 0
    srv.close()
    s.close()


def test_spmv_format_errors():
    """Format selection refuses what a copy does not have, with CM_ERR_ARG,
    and leaves the current format in place; the raw C entry point refuses an
    unknown format number and a matrix index out of range."""
    A, F = _small()
    s = cmamd.System("afiim", A, F)
    before = s.spmv_formats()
    for bits, which, fmt in ((64, "L", "dia"),        # AFIIM has no L factor
                             (64, "A", "staged16"),   # no 16-bit staged form in FP64
                             (32, "A", "dia"),        # mixed precision not enabled
                             (16, "G", "slot")):
        with pytest.raises(cmamd.EngineError) as e:
            s.spmv_format(bits, which, fmt)
        assert e.value.code == -2
    assert s.spmv_formats() == before
    lib = cmamd.solve_lib()
    assert lib.cm_system_spmv_format(s._h, 64, 0, 9) == -2
    assert lib.cm_system_spmv_format(s._h, 64, 7, 0) == -2
    assert lib.cm_system_spmv_format(s._h, 48, 0, 0) == -2
    # still solves, in a format that exists
    s.spmv_format(64, "A", "dia")
    s.set_rhs(cmamd.csr_mv_host(A, np.ones(A.n)))
    s.set_x(None)
    assert s.solve(tol=1e-8, nmax=200)[2] == 0
    s.close()


def test_block_resident_and_dia_switch_errors():
    """block_set without right-hand sides needs a resident block of the same
    width; switching the diagonal block products off and on keeps solves
    right."""
    A, F = _small()
    s = cmamd.System("afiim", A, F)
    lib = cmamd.solve_lib()
    assert lib.cm_system_block_set(s._h, 3, None, None) == -2  # nothing resident yet
    B = np.stack([cmamd.csr_mv_host(A, cmamd.gen_uniform(80 + j, A.n)) for j in range(3)], axis=1)
    s.block_set(B)
    assert lib.cm_system_block_set(s._h, 2, None, None) == -2  # other width
    for on in (False, True, False):
        s.block_dia(on)
        s.block_set(None)
        rv, it, info = s.block_solve(tol=1e-8, nmax=300)
        assert np.all(info == 0)
        X = s.block_get()
        for j in range(3):
            assert np.linalg.norm(B[:, j] - cmamd.csr_mv_host(A, X[:, j])) / np.linalg.norm(B[:, j]) <= 1e-6
    s.close()
