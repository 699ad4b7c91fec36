"""Error behaviour at the C ABI on a GPU: malformed inputs are rejected on the
host, before any kernel runs, with the documented codes (CM_ERR_ARG = -2,
CM_ERR_UNSUPPORTED = -3) and a message; the engine stays usable afterwards.
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
