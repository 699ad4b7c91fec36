"""Pins of the oracle (oracle/bicgstab_oracle.c, the CPU restatement of the
reference solve phase) against what the mathematics and independent libraries
fix -- never against a retyped copy of its own formulas:

* SpMV (general / symmetric-upper / transposed) against dense products in
  extended precision;
* triangular solves against scipy.linalg.solve_triangular;
* ACIFIM apply(): mode E against the dense product G diag(IDv) H, mode C
  against the closed form of nr Richardson sweeps
  (I - (I - H L)^nr) L^{-1} b, then (I - (I - G U)^nr) U^{-1} t2, modes I/J
  against triangular solves;
* the Krylov recurrence against an independent textbook right-preconditioned
  BiCGSTAB (van der Vorst 1992) written with numpy and dense M;
* the exit codes against hand-derived cases (NMAX reached; NMAX <= 0; the
  unguarded omega = 0/0 breakdown of an exact preconditioner, where AFIIM and
  PDFIIM run to NMAX on NaNs and ACIFIM's finite-alpha guard stops at pass 2;
  stagnation below rounding; divergence on systems whose singular
  preconditioned operator makes the residual grow on every pass, with the exit
  pass derived from the textbook trace and the reference's counting rule, for
  the AFIIM / PDFIIM limit at three sizes and the ACIFIM limit).
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import cmamd
import exit_cases
import oracle


def _rand_csr(n, seed, density=0.1, diag=1.0):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=rng, format="csr")
    M = M + sp.diags(np.full(n, diag))
    return cmamd.Csr.from_scipy(M)


def _ld_mv(m, x):
    d = m.to_scipy().toarray().astype(np.longdouble)
    return (d @ x.astype(np.longdouble)).astype(np.float64)


def test_csr_mv_general():
    A = _rand_csr(300, 1)
    x = np.random.default_rng(2).uniform(-1, 1, 300)
    y = oracle.csr_mv(A, x)
    ref = _ld_mv(A, x)
    assert np.linalg.norm(y - ref) <= 1e-14 * np.linalg.norm(ref)


def test_csr_mv_identity_and_permutation_exact():
    n = 50
    perm = np.random.default_rng(3).permutation(n)
    P = cmamd.Csr.from_scipy(sp.csr_matrix((np.ones(n), (np.arange(n), perm)), shape=(n, n)))
    x = np.random.default_rng(4).uniform(-1, 1, n)
    assert np.array_equal(oracle.csr_mv(P, x), x[perm])


def test_csr_mv_stencil_row_sums():
    # A @ 1 = row sums; for the convection-diffusion generator every interior
    # row sums to exactly 0 (the diagonal is the sum of the off-diagonal
    # magnitudes), boundary rows to a positive number
    A = cmamd.gen_convdiff(2, 20, 1.0, seed=3)
    y = oracle.csr_mv(A, np.ones(A.n))
    S = A.to_scipy()
    ref = np.asarray(S.sum(axis=1)).ravel()
    assert np.allclose(y, ref, rtol=0, atol=1e-12)
    assert np.all(y > -1e-12)


def test_csr_mv_symmetric_upper_and_transpose():
    A = _rand_csr(200, 5)
    Sfull = A.to_scipy() + A.to_scipy().T
    Up = cmamd.Csr.from_scipy(sp.triu(Sfull).tocsr())
    x = np.random.default_rng(6).uniform(-1, 1, 200)
    assert np.allclose(oracle.csr_mv(Up, x, "sym_upper"), Sfull @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(oracle.csr_mv(A, x, "trans"), A.to_scipy().T @ x, rtol=1e-13, atol=1e-13)


def _tri_factors(n, seed):
    rng = np.random.default_rng(seed)
    L = np.tril(rng.uniform(-0.3, 0.3, (n, n)) * (rng.random((n, n)) < 0.2), -1) + np.eye(n)
    U = np.triu(rng.uniform(-0.3, 0.3, (n, n)) * (rng.random((n, n)) < 0.2), 1) + np.eye(n)
    G = np.triu(rng.uniform(-0.3, 0.3, (n, n)) * (rng.random((n, n)) < 0.2), 1) + np.eye(n)
    H = np.tril(rng.uniform(-0.3, 0.3, (n, n)) * (rng.random((n, n)) < 0.2), -1) + np.eye(n)
    idv = rng.uniform(0.5, 2.0, n)
    c = lambda M: cmamd.Csr.from_scipy(sp.csr_matrix(M))  # noqa: E731
    return cmamd.Factors(c(G), idv, c(H), c(L), c(U)), G, H, L, U, idv


def _with_stray(T, lower, seed):
    # the triangular factor plus entries on the wrong side of the diagonal,
    # which the TRIANGULAR descriptor makes mkl_sparse_d_trsv ignore
    rng = np.random.default_rng(seed)
    n = T.shape[0]
    S = (np.triu if lower else np.tril)(rng.uniform(-5, 5, (n, n)) * (rng.random((n, n)) < 0.05),
                                         1 if lower else -1)
    return cmamd.Csr.from_scipy(sp.csr_matrix(T + S))


def test_trsv_ignores_entries_outside_the_triangle():
    F, G, H, L, U, idv = _tri_factors(80, 3)
    b = np.random.default_rng(4).uniform(-1, 1, 80)
    Ls, Us = _with_stray(L, True, 5), _with_stray(U, False, 6)
    assert Ls.nnz > F.L.nnz and Us.nnz > F.U.nnz
    assert np.allclose(oracle.trsv(Ls, b, lower=True), sla.solve_triangular(L, b, lower=True), rtol=1e-13)
    assert np.allclose(oracle.trsv(Us, b, lower=False), sla.solve_triangular(U, b, lower=False), rtol=1e-13)


def test_trsv_against_scipy():
    F, G, H, L, U, idv = _tri_factors(80, 1)
    b = np.random.default_rng(2).uniform(-1, 1, 80)
    assert np.allclose(oracle.trsv(F.L, b, lower=True), sla.solve_triangular(L, b, lower=True), rtol=1e-13)
    assert np.allclose(oracle.trsv(F.U, b, lower=False), sla.solve_triangular(U, b, lower=False), rtol=1e-13)


def test_apply_mode_e_dense():
    F, G, H, L, U, idv = _tri_factors(80, 3)
    v = np.random.default_rng(4).uniform(-1, 1, 80)
    assert np.allclose(oracle.apply(F, v, mode="E"), G @ (idv * (H @ v)), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("nr", [1, 2, 3, 5])
def test_apply_mode_c_closed_form(nr):
    F, G, H, L, U, idv = _tri_factors(60, 5)
    b = np.random.default_rng(6).uniform(-1, 1, 60)
    I = np.eye(60)
    t1 = (I - np.linalg.matrix_power(I - H @ L, nr)) @ np.linalg.solve(L, b)
    t2 = idv * t1
    x = (I - np.linalg.matrix_power(I - G @ U, nr)) @ np.linalg.solve(U, t2)
    assert np.allclose(oracle.apply(F, b, nr=nr, mode="C"), x, rtol=1e-11, atol=1e-12)


def test_apply_modes_i_j():
    F, G, H, L, U, idv = _tri_factors(60, 7)
    b = np.random.default_rng(8).uniform(-1, 1, 60)
    xi = sla.solve_triangular(U, idv * sla.solve_triangular(L, b, lower=True), lower=False)
    xj = sla.solve_triangular(U, sla.solve_triangular(L, b, lower=True), lower=False)
    assert np.allclose(oracle.apply(F, b, mode="I"), xi, rtol=1e-12)
    assert np.allclose(oracle.apply(F, b, mode="J"), xj, rtol=1e-12)


def textbook_bicgstab(A, M, b, iters):
    """Right-preconditioned BiCGSTAB, van der Vorst (1992), dense numpy:
    returns ||r_k|| for k = 0..iters."""
    x = np.zeros_like(b)
    r = b - A @ x
    rh = r.copy()
    rho_old = alpha = omega = 1.0
    v = np.zeros_like(b)
    p = np.zeros_like(b)
    norms = [np.linalg.norm(r)]
    for _ in range(iters):
        rho = rh @ r
        beta = (rho / rho_old) * (alpha / omega)
        p = r + beta * (p - omega * v)
        ph = M @ p
        v = A @ ph
        alpha = rho / (rh @ v)
        s = r - alpha * v
        sh = M @ s
        t = A @ sh
        omega = (t @ s) / (t @ t)
        x = x + alpha * ph + omega * sh
        r = s - omega * t
        rho_old = rho
        norms.append(np.linalg.norm(r))
    return np.array(norms)


@pytest.mark.parametrize("variant", ["afiim", "pdfiim", "acifim"])
def test_bicgstab_recurrence_matches_textbook(variant):
    A = cmamd.gen_convdiff(2, 14, 2.0, seed=11)
    xs = cmamd.gen_uniform(12, A.n)
    b = cmamd.csr_mv_host(A, xs)
    F = cmamd.afiim(A, 0.3) if variant != "acifim" else cmamd.acifim(A, 0.05, 0.3)
    M = F.G.to_scipy().toarray() @ np.diag(F.idv) @ F.H.to_scipy().toarray()
    x, rv, it, info, tr = oracle.bicgstab(variant, A, b, F, tol=1e-10, nmax=200, trace_cap=200)
    ref = textbook_bicgstab(A.to_scipy().toarray(), M, b, len(tr) - 1)
    m = min(len(tr), 20)
    assert np.max(np.abs(tr[:m] - ref[:m]) / ref[:m]) < 1e-9
    assert info == 0
    assert np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b) < 1e-9


def test_bicgstab_pdfiim_symmetric_branch():
    # symmetric A stored as its upper triangle, M = G D^{-1} G^T
    A = cmamd.gen_convdiff(2, 12, 0.0, seed=13)  # Pe = 0: symmetric
    S = A.to_scipy()
    assert abs(S - S.T).max() < 1e-12
    Up = cmamd.Csr.from_scipy(sp.triu(S).tocsr())
    F = cmamd.afiim(A, 0.2)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(14, A.n))
    Fs = cmamd.Factors(F.G, F.idv, None)
    x, rv, it, info, tr = oracle.bicgstab("pdfiim", Up, b, Fs, nmax=200, issym=True, trace_cap=100)
    Gd = F.G.to_scipy().toarray()
    ref = textbook_bicgstab(S.toarray(), Gd @ np.diag(F.idv) @ Gd.T, b, len(tr) - 1)
    m = min(len(tr), 20)
    assert np.max(np.abs(tr[:m] - ref[:m]) / ref[:m]) < 1e-9
    assert info == 0


def _c1_small():
    A = cmamd.gen_convdiff(2, 30, 0.5, seed=1)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(2, A.n))
    return A, b, cmamd.afiim(A, 0.1)


def test_exit_nmax_reached():
    A, b, F = _c1_small()
    x, rv, it, info, _ = oracle.bicgstab("afiim", A, b, F, nmax=3)
    assert (it, info) == (3, 3)


def test_exit_nmax_nonpositive():
    A, b, F = _c1_small()
    x, rv, it, info, _ = oracle.bicgstab("afiim", A, b, F, nmax=0, info_in=77)
    assert (it, info, rv) == (0, 3, 1.0)
    x, rv, it, info, _ = oracle.bicgstab("afiim", A, b, F, nmax=-1, info_in=77)
    assert (it, info, rv) == (0, 77, 1.0)  # the reference leaves info untouched


def _exact_diag():
    # A = 2 I, M = I/2, b = ones(16): every product is exact, ||b||^2 = 16 is
    # a perfect square, so alpha = m1/(r0.v) = 1 exactly and s = 0 exactly
    n = 16
    A = cmamd.Csr(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.full(n, 2.0))
    I = cmamd.Csr(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n))
    F = cmamd.Factors(I, np.full(n, 0.5), I, I, I)
    return A, np.ones(n), F


def test_exit_exact_preconditioner_breakdown():
    # M = A^{-1}: alpha = 1, s = 0, omega = 0/0 -> NaN (not guarded)
    A, b, F = _exact_diag()
    x, rv, it, info, _ = oracle.bicgstab("afiim", A, b, F, nmax=7)
    assert (it, info) == (7, 3) and np.isnan(rv)
    # ACIFIM: pass 2 has alpha = NaN/NaN -> iter = NMAX, info = 3, rc still ||r0||
    x, rv, it, info, _ = oracle.bicgstab("acifim", A, b, F, nmax=7)
    assert (it, info, rv) == (7, 3, 1.0)


def test_exit_stagnation_below_rounding():
    A, b, F = _c1_small()
    x, rv, it, info, tr = oracle.bicgstab("afiim", A, b, F, tol=1e-30, nmax=500, trace_cap=500)
    assert info == 2
    assert it < 500 and rv < 1e-12


@pytest.mark.parametrize("name", sorted(exit_cases.DIVERGING))
@pytest.mark.parametrize("variant", ["afiim", "pdfiim"])
def test_exit_divergence_afiim_limit(name, variant):
    # n = 64, 600, 1024: maxmsteps 10, 12, 20 -- the exit pass follows from an
    # independent textbook trace and the reference's counting rule
    A, b, F, d = exit_cases.diverging(name)
    limit = exit_cases.maxmsteps(variant, A.n)
    assert limit == {"d64": 10, "d600": 12, "d1024": 20}[name]
    norms = exit_cases.textbook_norms(d, b, 60)
    it_e, rv_e = exit_cases.divergence_exit(norms, limit)
    assert it_e == limit  # the residual grows from the first pass on
    x, rv, it, info, tr = oracle.bicgstab(variant, A, b, F, tol=1e-8, nmax=200, trace_cap=100)
    assert (it, info) == (it_e, 1)
    assert abs(rv - rv_e) <= 1e-9 * rv_e
    m = len(tr)
    assert m == it + 1 and np.max(np.abs(tr - norms[:m]) / norms[:m]) < 1e-9


def test_exit_divergence_acifim_limit():
    # ACIFIM allows 40 growing passes (ACIFIM/src/bicgstab.cpp:132)
    A, b, F, d = exit_cases.diverging("d600")
    assert exit_cases.maxmsteps("acifim", A.n) == 40
    x, rv, it, info, tr = oracle.bicgstab("acifim", A, b, F, tol=1e-8, nmax=200, trace_cap=100)
    assert (it, info) == (40, 1)
    assert exit_cases.divergence_exit(tr, 40)[0] == 40
    assert exit_cases.divergence_exit(exit_cases.textbook_norms(d, b, 60), 40)[0] == 40
    # past the AFIIM / PDFIIM limit of the same system without stopping
    for name in ("d64", "d1024"):
        A, b, F, d = exit_cases.diverging(name)
        x, rv, it, info, tr = oracle.bicgstab("acifim", A, b, F, tol=1e-8, nmax=200)
        assert it > exit_cases.maxmsteps("afiim", A.n)


def test_thread_count_only_changes_rounding():
    A = cmamd.gen_convdiff(2, 60, 0.5, seed=1)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(2, A.n))
    F = cmamd.afiim(A, 0.1)
    oracle.set_threads(1)
    r1 = oracle.bicgstab("afiim", A, b, F, trace_cap=40)
    oracle.set_threads(4)
    r4 = oracle.bicgstab("afiim", A, b, F, trace_cap=40)
    oracle.set_threads(1)
    assert r1[2] == r4[2] and r1[3] == r4[3]
    assert np.max(np.abs(r1[4][:20] - r4[4][:20]) / r1[4][:20]) < 1e-10
