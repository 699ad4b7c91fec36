"""Every exit branch of the reference loop on the GPU, in the three solve
paths (single system, multi-RHS block, persistent batch), against the oracle
and against the reference's own rules:

* divergence, info 1 (AFIIM/src/bicgstab.cpp:219-234): systems whose residual
  grows on every pass (tests/exit_cases.py) stop when msteps reaches
  maxmsteps -- 10, 12 and 20 for AFIIM / PDFIIM at n = 64, 600, 1024;
  ACIFIM allows 40 (ACIFIM/src/bicgstab.cpp:132).  AFIIM / PDFIIM exits are
  held to the oracle's (iter, info) exactly.  ACIFIM's 40-pass run amplifies
  rounding until two implementations' traces part (the oracle's own traces at
  different summation orders do too), so on the GPU its exit is held to the
  reference rule applied to the GPU's own trace, and must come after the
  AFIIM limit of the same system;
* ACIFIM's non-finite-alpha guard (ACIFIM/src/bicgstab.cpp:208-213): A = 2I,
  M = I/2, b = 1 makes s = 0 exactly, omega = 0/0, alpha NaN on pass 2;
  AFIIM / PDFIIM have no guard and run to NMAX on NaNs;
* NMAX = 0 (info 3, iter 0, resval 1) and NMAX < 0 (info left as the caller
  set it) (AFIIM/src/bicgstab.cpp:150-151,284-288);
* stagnation below rounding (info 2) and NMAX reached (info 3).
"""
import numpy as np
import pytest

import cmamd
import exit_cases
import oracle

pytestmark = pytest.mark.gpu


def _exact_diag(n=16):
    # every operation exact: alpha = 1, s = 0, omega = 0/0
    I = exit_cases.identity(n)
    A = cmamd.Csr(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.full(n, 2.0))
    return A, np.ones(n), cmamd.Factors(I, np.full(n, 0.5), I, I, I)


def _single(variant, A, F, b, nmax, tol=1e-8, info_in=0, trace_cap=128):
    s = cmamd.System(variant, A, F)
    s.set_rhs(b)
    s.set_x(None)
    out = s.solve(tol=tol, nmax=nmax, trace_cap=trace_cap, info_in=info_in)
    s.close()
    return out


@pytest.mark.parametrize("name", sorted(exit_cases.DIVERGING))
@pytest.mark.parametrize("variant", ["afiim", "pdfiim"])
def test_divergence_single(name, variant):
    A, b, F, d = exit_cases.diverging(name)
    limit = exit_cases.maxmsteps(variant, A.n)
    print(f"divergence single {variant} {name} n={A.n} limit={limit}", flush=True)
    rv, it, info, tr = _single(variant, A, F, b, 200)
    xo, rvo, ito, info_o, tro = oracle.bicgstab(variant, A, b, F, tol=1e-8, nmax=200, trace_cap=128)
    assert (it, info) == (ito, info_o) == (limit, 1)
    assert abs(rv - rvo) <= 1e-8 * rvo
    m = min(len(tr), len(tro))
    assert m == limit + 1 and np.max(np.abs(tr[:m] - tro[:m]) / tro[:m]) < 1e-8


def test_divergence_single_acifim():
    for name in ("d64", "d600", "d1024"):
        A, b, F, d = exit_cases.diverging(name)
        print(f"divergence single acifim {name} n={A.n}", flush=True)
        rv, it, info, tr = _single("acifim", A, F, b, 200, trace_cap=256)
        # the first passes are the oracle's; the exit follows the rule on the
        # GPU's own trace with ACIFIM's limit, never AFIIM's
        xo, rvo, ito, info_o, tro = oracle.bicgstab("acifim", A, b, F, tol=1e-8, nmax=200, trace_cap=16)
        assert np.max(np.abs(tr[:12] - tro[:12]) / tro[:12]) < 1e-8
        assert it > exit_cases.maxmsteps("afiim", A.n)
        if info == 1:
            assert exit_cases.divergence_exit(tr, 40) == (it, pytest.approx(rv, rel=1e-12))
        else:
            with pytest.raises(ValueError):
                exit_cases.divergence_exit(tr, 40)
            assert (it, info) == (200, 3)


def test_divergence_block_columns():
    # one system, three right-hand sides: b, 2b (the same trajectory scaled
    # exactly) and a convergent one (b = A M-range vector); each column keeps
    # its own exit
    A, b, F, d = exit_cases.diverging("d64")
    b3 = np.zeros(A.n)
    b3[4:] = 1.0  # no null component: converges
    B = np.stack([b, 2.0 * b, b3], axis=1)
    s = cmamd.System("afiim", A, F)
    s.block_set(B)
    rv, it, info = s.block_solve(tol=1e-8, nmax=200)
    s.close()
    for j in range(3):
        xo, rvo, ito, info_o, _ = oracle.bicgstab("afiim", A, B[:, j], F, tol=1e-8, nmax=200)
        assert info[j] == info_o, j
        if j < 2:  # diverging: the exit pass is exact; converging: the 2% bar
            assert it[j] == ito and abs(rv[j] - rvo) <= 1e-8 * rvo, j
        else:
            assert abs(it[j] - ito) <= max(1, int(np.ceil(0.02 * ito))), j
    assert (it[0], info[0]) == (it[1], info[1]) == (10, 1)
    assert info[2] == 0


def test_divergence_and_guards_batched():
    cases = []
    for name, variant in (("d64", "afiim"), ("d600", "pdfiim"), ("d1024", "afiim")):
        A, b, F, d = exit_cases.diverging(name)
        cases.append((variant, A, F, b, 200))
    Ae, be, Fe = _exact_diag()
    cases.append(("acifim", Ae, Fe, be, 7))   # alpha guard
    cases.append(("afiim", Ae, Fe, be, 7))    # NaN run to NMAX
    systems = [cmamd.System(v, A, F) for v, A, F, b, nm in cases]
    items = [(h, b, None, 1e-8, nm) for h, (v, A, F, b, nm) in zip(systems, cases)]
    res = cmamd.batch_solve(items)
    for h in systems:
        h.close()
    for r, (v, A, F, b, nm) in zip(res, cases):
        xo, rvo, ito, info_o, _ = oracle.bicgstab(v, A, b, F, tol=1e-8, nmax=nm)
        assert (r.iter, r.info) == (ito, info_o), v
        if np.isnan(rvo):
            assert np.isnan(r.resval)
        else:
            assert abs(r.resval - rvo) <= 1e-8 * rvo
    assert [(r.iter, r.info) for r in res] == [(10, 1), (12, 1), (20, 1), (7, 3), (7, 3)]
    assert res[3].resval == 1.0


@pytest.mark.parametrize("variant", ["afiim", "pdfiim", "acifim"])
def test_alpha_guard_and_nan_single(variant):
    A, b, F = _exact_diag()
    rv, it, info, tr = _single(variant, A, F, b, 7)
    xo, rvo, ito, info_o, _ = oracle.bicgstab(variant, A, b, F, nmax=7)
    assert (it, info) == (ito, info_o) == (7, 3)
    if variant == "acifim":
        assert rv == rvo == 1.0  # stopped by the guard on pass 2, rc still ||r_0||
    else:
        assert np.isnan(rv) and np.isnan(rvo)


def test_alpha_guard_block():
    A, b, F = _exact_diag()
    s = cmamd.System("acifim", A, F)
    s.block_set(np.stack([b, 2.0 * b], axis=1))
    rv, it, info = s.block_solve(tol=1e-8, nmax=7)
    s.close()
    assert list(it) == [7, 7] and list(info) == [3, 3] and list(rv) == [1.0, 1.0]


def test_nmax_zero_and_negative_all_paths():
    A = cmamd.gen_convdiff(2, 30, 0.5, seed=1)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(2, A.n))
    F = cmamd.afiim(A, 0.1)
    for nmax, want in ((0, 3), (-1, 77)):
        xo, rvo, ito, info_o, _ = oracle.bicgstab("afiim", A, b, F, nmax=nmax, info_in=77)
        assert (ito, info_o, rvo) == (0, want, 1.0)
        rv, it, info, _ = _single("afiim", A, F, b, nmax, info_in=77)
        assert (it, info, rv) == (0, want, 1.0), nmax
        s = cmamd.System("afiim", A, F)
        s.block_set(np.stack([b, -b], axis=1))
        rvb, itb, infob = s.block_solve(nmax=nmax, info_in=77)
        assert list(itb) == [0, 0] and list(infob) == [want, want] and list(rvb) == [1.0, 1.0], nmax
        r = cmamd.batch_solve([(s, b, None, 1e-8, nmax)], info_in=77)[0]
        assert (r.iter, r.info, r.resval) == (0, want, 1.0), nmax
        assert np.all(r.x == 0.0)
        s.close()


def test_stagnation_and_nmax_exact_codes():
    A = cmamd.gen_convdiff(2, 100, 0.5, seed=1)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(2, A.n))
    F = cmamd.afiim(A, 0.1)
    xo, rvo, ito, info_o, _ = oracle.bicgstab("afiim", A, b, F, tol=1e-30, nmax=500)
    assert info_o == 2
    rv, it, info, _ = _single("afiim", A, F, b, 500, tol=1e-30)
    assert info == 2
    # stagnation is decided at the rounding floor, so the pass may move by a
    # few where the two summation orders differ by an ulp
    assert abs(it - ito) <= max(3, int(np.ceil(0.05 * ito)))
    rv, it, info, _ = _single("afiim", A, F, b, 5)
    assert (it, info) == (5, 3)
