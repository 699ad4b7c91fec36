"""GPU parity of the multi-right-hand-side (block) solves, through the C ABI.

Each column of a block solve is the reference loop run on that column alone
(AFIIM/src/bicgstab.cpp:149-281): the oracle solves every column separately
and the block result is held to the single-solve bars -- same info code,
iteration count within 2% (at least 1), final relative residual within 1e-8
relative (plus a 1e-12 rounding floor) when the counts agree and the exit
falls in the first 20 passes (the span of the per-pass residual bar; later,
different summation orders legitimately drift further), solution within
1e-6.  Padding columns (k not a power of two), long rows (the wave-per-row
bin), the symmetric PDFIIM branch (G^T copy), nonzero initial guesses and a
zero right-hand side (a NaN column, which must not disturb the others) are
covered; mixed-precision block refinement is checked per column against the
true residual and the FP64 oracle solution.
"""
import numpy as np
import pytest

import cmamd
import oracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _rhs(A, k, seed):
    return np.stack([cmamd.csr_mv_host(A, cmamd.gen_uniform(seed + j, A.n)) for j in range(k)], axis=1)


def _check_columns(variant, A, F, B, X, rv, it, info, nmax, tol=1e-8, X0=None, issym=False, **kw):
    for j in range(B.shape[1]):
        x0 = None if X0 is None else X0[:, j]
        xo, rvo, ito, info_o, _ = oracle.bicgstab(variant, A, B[:, j], F, x0=x0, tol=tol, nmax=nmax, issym=issym,
                                                 **kw)
        assert info[j] == info_o, (j, info[j], info_o)
        assert abs(it[j] - ito) <= max(1, int(np.ceil(0.02 * ito))), (j, it[j], ito)
        if np.isnan(rvo):
            assert np.isnan(rv[j]), j
            continue
        if it[j] == ito and ito <= 20:  # the residual bar covers the first 20 passes
            assert abs(rv[j] - rvo) <= 1e-8 * rvo + 1e-12, (j, rv[j], rvo)
        if info_o == 0:
            assert _rel(X[:, j], xo) <= 1e-6, j


@pytest.mark.parametrize("k", [1, 3, 8, 16])
def test_block_afiim_columns(k):
    A = cmamd.gen_convdiff(2, 100, 0.5, seed=1)
    F = cmamd.afiim(A, 0.1)
    B = _rhs(A, k, 20)
    sys_ = cmamd.System("afiim", A, F)
    sys_.block_set(B)
    rv, it, info = sys_.block_solve(tol=1e-8, nmax=500)
    X = sys_.block_get()
    sys_.close()
    assert np.all(info == 0)
    _check_columns("afiim", A, F, B, X, rv, it, info, 500)


@pytest.mark.parametrize("kind,k", [("cd3", 1), ("cd3", 5), ("cd3", 16), ("long_rows", 5), ("long_rows", 16),
                                    ("long_rows_large", 5), ("long_rows_large", 16)])
def test_block_dia_matches_row_binned(kind, k):
    """Block products in the diagonal format (the default for stencil and
    factor matrices; with long rows, their remainders summed by a wave-per-row
    pass first) against the row-binned kernels (block_dia off): both meet the
    oracle checks in FP64, both FP32- and FP16-inner refinements reach the
    tolerance, and the two agree to 1e-6."""
    if kind == "cd3":
        A = cmamd.gen_convdiff(3, 20, 2.0, seed=17)
    elif kind == "long_rows":  # 27-point stencil, 5% of rows with 200-600 long-range couplings
        A = cmamd.gen_stencil27(14, 0.05, 200, 600, seed=9)
        assert (np.diff(A.ptr) > 64).sum() > 0
    else:  # the same at 52^3 rows (about 7000 long rows over 140608 columns)
        A = cmamd.gen_stencil27(52, 0.05, 200, 600, seed=9)
        assert (np.diff(A.ptr) > 64).sum() > 0
    F = cmamd.afiim(A, 0.1)
    B = _rhs(A, k, 70)
    out = {}
    for on in (True, False):
        sys_ = cmamd.System("afiim", A, F)
        sys_.enable_mixed(32)
        sys_.enable_mixed(16)
        sys_.block_dia(on)
        sys_.block_set(B)
        rv, it, info = sys_.block_solve(tol=1e-8, nmax=500)
        X = sys_.block_get()
        _check_columns("afiim", A, F, B, X, rv, it, info, 500)
        res = {64: X}
        for bits in (32, 16):
            sys_.block_set(B)
            rv, outer, inner, info = sys_.block_solve_mixed(bits, tol=1e-8, inner_tol=1e-4, nmax=500)
            Xm = sys_.block_get()
            assert np.all(info == 0)
            for j in range(k):
                assert np.linalg.norm(B[:, j] - cmamd.csr_mv_host(A, Xm[:, j])) / np.linalg.norm(B[:, j]) <= 1e-8
            res[bits] = Xm
        out[on] = res
        sys_.close()
    for bits in (64, 32, 16):
        for j in range(k):
            assert _rel(out[True][bits][:, j], out[False][bits][:, j]) <= 1e-6, (bits, j)


def test_block_resolve_with_resident_rhs():
    """block_set(None) keeps the resident right-hand sides and resets x: the
    second solve repeats the first bit for bit; a different width without B
    is refused."""
    A = cmamd.gen_convdiff(2, 60, 1.0, seed=23)
    F = cmamd.afiim(A, 0.1)
    B = _rhs(A, 4, 90)
    sys_ = cmamd.System("afiim", A, F)
    sys_.enable_mixed(32)
    sys_.block_set(B)
    r1 = sys_.block_solve_mixed(32, tol=1e-8, inner_tol=1e-4, nmax=500)
    X1 = sys_.block_get()
    sys_.block_set(None)
    r2 = sys_.block_solve_mixed(32, tol=1e-8, inner_tol=1e-4, nmax=500)
    X2 = sys_.block_get()
    assert np.array_equal(X1, X2) and np.array_equal(r1[0], r2[0]) and r1[1:3] == r2[1:3]
    sys_._bk = 5
    with pytest.raises(cmamd.EngineError):
        sys_.block_set(None)
    sys_.close()


def test_block_long_rows():
    # 27-point stencil with long-range couplings: 154 rows longer than the
    # short-row bin (up to 623 nonzeros)
    A = cmamd.gen_stencil27(14, 0.05, 200, 600, seed=9)
    assert (np.diff(A.ptr) > 64).sum() > 0
    F = cmamd.afiim(A, 0.1)
    B = _rhs(A, 5, 40)
    sys_ = cmamd.System("afiim", A, F)
    sys_.block_set(B)
    rv, it, info = sys_.block_solve(tol=1e-8, nmax=500)
    X = sys_.block_get()
    sys_.close()
    _check_columns("afiim", A, F, B, X, rv, it, info, 500)


def test_block_dia_switched_on_after_block_set():
    # the long-row buffer of the diagonal block products must exist when a
    # block workspace made with them off is used after switching them on
    A = cmamd.gen_stencil27(14, 0.05, 200, 600, seed=9)
    F = cmamd.afiim(A, 0.1)
    B = _rhs(A, 4, 60)
    sys_ = cmamd.System("afiim", A, F)
    sys_.block_dia(False)
    sys_.block_set(B)
    sys_.enable_mixed(32)
    sys_.block_dia(True)
    rv, it, info = sys_.block_solve(tol=1e-8, nmax=500)
    X = sys_.block_get()
    _check_columns("afiim", A, F, B, X, rv, it, info, 500)
    sys_.block_set(None)
    rvm, outer, inner, infom = sys_.block_solve_mixed(32, tol=1e-8)
    Xm = sys_.block_get()
    sys_.close()
    assert np.all(infom == 0)
    for j in range(B.shape[1]):
        assert np.linalg.norm(B[:, j] - cmamd.csr_mv_host(A, Xm[:, j])) / np.linalg.norm(B[:, j]) <= 1e-8


def test_block_pdfiim_symmetric():
    import scipy.sparse as sp
    A = cmamd.gen_convdiff(3, 14, 0.0, seed=13)
    Up = cmamd.Csr.from_scipy(sp.triu(A.to_scipy()).tocsr())
    F = cmamd.afiim(A, 0.1)
    Fs = cmamd.Factors(F.G, F.idv, None)
    B = _rhs(A, 5, 60)
    sys_ = cmamd.System("pdfiim", Up, Fs, a_symmetric=True)
    sys_.block_set(B)
    rv, it, info = sys_.block_solve(tol=1e-8, nmax=1200)
    X = sys_.block_get()
    sys_.close()
    _check_columns("pdfiim", Up, Fs, B, X, rv, it, info, 1200, issym=True)


def test_block_pdfiim_general():
    A = cmamd.gen_convdiff(3, 20, 1.0, seed=5)
    F = cmamd.pdfiim(A, 2, 2, 1, 0.05, -1, "I", "B")
    B = _rhs(A, 4, 80)
    sys_ = cmamd.System("pdfiim", A, F)
    sys_.block_set(B)
    rv, it, info = sys_.block_solve(tol=1e-8, nmax=1200)
    X = sys_.block_get()
    sys_.close()
    _check_columns("pdfiim", A, F, B, X, rv, it, info, 1200)


def test_block_acifim_mode_e_and_reject():
    A = cmamd.gen_convdiff(3, 16, 5.0, seed=8)
    F = cmamd.acifim(A, 0.01, 0.1)
    B = _rhs(A, 4, 100)
    sys_ = cmamd.System("acifim", A, F, nr=1, mode="E")
    sys_.block_set(B)
    rv, it, info = sys_.block_solve(tol=1e-10, nmax=2000)
    X = sys_.block_get()
    sys_.close()
    _check_columns("acifim", A, F, B, X, rv, it, info, 2000, tol=1e-10, nr=1, mode="E")
    sys_ = cmamd.System("acifim", A, F, nr=2, mode="C")
    with pytest.raises(cmamd.EngineError) as e:
        sys_.block_set(B)
    assert e.value.code == -3
    sys_.close()


def test_block_initial_guess_zero_rhs_and_nmax():
    A = cmamd.gen_convdiff(2, 100, 0.5, seed=1)
    F = cmamd.afiim(A, 0.1)
    B = _rhs(A, 6, 120)
    B[:, 2] = 0.0  # ||r_0|| = 0: alpha is 0/0, the column runs to NMAX on NaN
    X0 = np.zeros_like(B)
    X0[:, 4] = cmamd.gen_uniform(200, A.n)
    X0[:, 5] = 0.5 * cmamd.gen_uniform(201, A.n)
    sys_ = cmamd.System("afiim", A, F)
    sys_.block_set(B, X0)
    rv, it, info = sys_.block_solve(tol=1e-8, nmax=60)
    X = sys_.block_get()
    assert info[2] == 3 and it[2] == 60 and np.isnan(rv[2])
    _check_columns("afiim", A, F, B, X, rv, it, info, 60, X0=X0)
    # NMAX reached by every column; then a second solve on the same block
    sys_.block_set(B[:, [0, 1, 3]])
    rv, it, info = sys_.block_solve(tol=1e-8, nmax=2)
    assert np.all(it == 2) and np.all(info == 3)
    X = sys_.block_get()
    sys_.close()
    _check_columns("afiim", A, F, B[:, [0, 1, 3]], X, rv, it, info, 2)


@pytest.mark.parametrize("bits", [32, 16])
@pytest.mark.parametrize("case", ["cd2d_afiim", "cd3d_pe5_acifim"])
def test_block_mixed_precision(bits, case):
    if case == "cd2d_afiim":
        A = cmamd.gen_convdiff(2, 100, 0.5, seed=1)
        F = cmamd.afiim(A, 0.1)
        variant, tol = "afiim", 1e-8
    else:
        A = cmamd.gen_convdiff(3, 24, 5.0, seed=4)
        F = cmamd.acifim(A, 0.01, 0.1)
        variant, tol = "acifim", 1e-10
    B = _rhs(A, 5, 140)
    sys_ = cmamd.System(variant, A, F)
    sys_.enable_mixed(bits)
    sys_.block_set(B)
    rv, outer, inner, info = sys_.block_solve_mixed(bits, tol=tol, inner_tol=1e-4, nmax=2000)
    X = sys_.block_get()
    sys_.close()
    assert np.all(info == 0), (rv, outer, inner, info)
    assert 1 <= outer <= 6
    for j in range(B.shape[1]):
        b = B[:, j]
        true_res = np.linalg.norm(b - cmamd.csr_mv_host(A, X[:, j])) / np.linalg.norm(b)
        assert true_res <= tol and abs(true_res - rv[j]) <= 1e-3 * tol + 1e-2 * rv[j], j
        xo = oracle.bicgstab(variant, A, b, F, tol=tol, nmax=2000)[0]
        assert _rel(X[:, j], xo) <= 1e3 * tol, j


def test_block_mixed_matches_single_mixed():
    """Block refinement of k columns against k single mixed solves: the same
    outer exit per column and solutions within the refinement tolerance."""
    A = cmamd.gen_convdiff(2, 100, 0.5, seed=1)
    F = cmamd.afiim(A, 0.1)
    B = _rhs(A, 3, 160)
    sys_ = cmamd.System("afiim", A, F)
    sys_.enable_mixed(32)
    sys_.block_set(B)
    rv, outer, inner, info = sys_.block_solve_mixed(32, tol=1e-8, inner_tol=1e-4, nmax=500)
    X = sys_.block_get()
    for j in range(3):
        sys_.set_rhs(B[:, j])
        sys_.set_x(None)
        rv1, o1, in1, info1 = sys_.solve_mixed(32, tol=1e-8, inner_tol=1e-4, nmax=500)
        x1 = sys_.get_x()
        assert info1 == info[j] == 0
        assert _rel(X[:, j], x1) <= 1e-6
    sys_.close()


def test_block_graph_replays_match_eager():
    """Captured passes (hipGraph replays) against kernel-by-kernel launches:
    bit-identical per-column results, over several replays with eager work
    (a single-RHS solve and host synchronisations) between them."""
    A = cmamd.gen_convdiff(2, 100, 0.5, seed=1)
    F = cmamd.afiim(A, 0.1)
    B = _rhs(A, 4, 180)
    sys_ = cmamd.System("afiim", A, F)
    sys_.enable_mixed(32)
    ref = {}
    for graphs in (False, True, True, True):
        sys_.block_graphs(graphs)
        sys_.block_set(B)
        r64 = sys_.block_solve(tol=1e-8, nmax=500)
        X64 = sys_.block_get()
        sys_.block_set(B)
        rmx = sys_.block_solve_mixed(32, tol=1e-8, inner_tol=1e-4, nmax=500)
        Xmx = sys_.block_get()
        sys_.set_rhs(B[:, 0])
        sys_.set_x(None)
        sys_.solve(tol=1e-8, nmax=500)  # eager single-RHS work between replays
        got = (X64, r64, Xmx, rmx)
        if not ref:
            ref = got
            continue
        assert np.array_equal(got[0], ref[0])
        for a, b in zip(got[1], ref[1]):
            assert np.array_equal(a, b)
        assert np.array_equal(got[2], ref[2])
        assert np.array_equal(got[3][0], ref[3][0]) and got[3][1:3] == ref[3][1:3]
        assert np.array_equal(got[3][3], ref[3][3])
    sys_.close()


def test_block_autotuned_bins_keep_parity():
    A = cmamd.gen_stencil27(14, 0.05, 200, 600, seed=9)
    F = cmamd.afiim(A, 0.1)
    B = _rhs(A, 8, 220)
    sys_ = cmamd.System("afiim", A, F)
    sys_.enable_mixed(32)
    sys_.block_set(B)
    for bits in (64, 32):
        chosen = sys_.block_autotune(bits, reps=5)
        assert all(v in (16, 32, 64, 128) for v in chosen.values())
    rv, it, info = sys_.block_solve(tol=1e-8, nmax=500)
    X = sys_.block_get()
    _check_columns("afiim", A, F, B, X, rv, it, info, 500)
    sys_.block_set(B)
    rv, outer, inner, info = sys_.block_solve_mixed(32, tol=1e-8, inner_tol=1e-4, nmax=500)
    X = sys_.block_get()
    sys_.close()
    assert np.all(info == 0)
    for j in range(B.shape[1]):
        assert np.linalg.norm(B[:, j] - cmamd.csr_mv_host(A, X[:, j])) / np.linalg.norm(B[:, j]) <= 1e-8
