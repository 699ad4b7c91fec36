"""GPU parity of the FP64 solve path against the oracle (CPU restatement of the
reference solve loops), through the C ABI.

Bars: SpMV and M-apply within 1e-12 relative (normwise) of the oracle; the
per-pass residual norms within 1e-8 relative over the first 20 passes (down
to a rounding floor of 1e-12 ||r_0||, see _trace_close);
iteration counts within 2%; identical info codes; the true relative residual
of the returned x meets the tolerance.
"""
import numpy as np
import pytest

import cmamd
import oracle

pytestmark = pytest.mark.gpu

SPMV_TOL = 1e-12
TRACE_TOL = 1e-8


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _trace_close(tr, tro, rel=TRACE_TOL, floor=1e-12):
    """Per-pass residual norms agree within `rel` relative, down to a rounding
    floor of `floor` * ||r_0||: two FP64 implementations that sum in different
    orders differ by ~eps * ||r_0|| in ||r_k||, which is 1e-8 relative once the
    residual has dropped eight orders (tro[0] is ||r_0||)."""
    return bool(np.all(np.abs(tr - tro) <= rel * tro + floor * tro[0]))


def _system(kind):
    if kind == "cd2d":
        return cmamd.gen_convdiff(2, 100, 0.5, seed=1)
    if kind == "cd3":
        return cmamd.gen_convdiff(3, 20, 1.0, seed=5)
    if kind == "irregular":
        return cmamd.gen_stencil27(14, 0.05, 200, 600, seed=9)
    raise KeyError(kind)


def _random_csr(n, seed, long_rows=(), empty_rows=()):
    """Random CSR with a dominant diagonal, a few very long rows (longer than
    one SpMV tile) and some empty rows."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        if i in empty_rows:
            continue
        k = 3000 if i in long_rows else rng.integers(1, 12)
        c = rng.choice(n, size=min(k, n), replace=False)
        rows += [i] * len(c)
        cols += list(c)
    m = sp.csr_matrix((rng.uniform(-1, 1, len(rows)), (rows, cols)), shape=(n, n))
    m.sum_duplicates()
    return cmamd.Csr.from_scipy(m)


@pytest.mark.parametrize("kind", ["cd2d", "cd3", "irregular"])
def test_spmv_matches_oracle(kind):
    A = _system(kind)
    F = cmamd.afiim(A, 0.1)
    sys_ = cmamd.System("afiim", A, F)
    x = cmamd.gen_uniform(3, A.n)
    for which, M in (("A", A), ("G", F.G), ("H", F.H)):
        y = sys_.spmv(which, x)
        assert _rel(y, oracle.csr_mv(M, x)) <= SPMV_TOL, which
    sys_.close()


def test_spmv_long_and_empty_rows():
    A = _random_csr(5000, 3, long_rows={7, 4000, 4999}, empty_rows={0, 1, 2500})
    # identity factors: the system only needs a G and an H
    I = cmamd.Csr(A.n, np.arange(A.n + 1, dtype=np.int32), np.arange(A.n, dtype=np.int32), np.ones(A.n))
    F = cmamd.Factors(I, np.ones(A.n), I)
    sys_ = cmamd.System("afiim", A, F)
    x = np.random.default_rng(1).uniform(-1, 1, A.n)
    y = sys_.spmv("A", x)
    ref = oracle.csr_mv(A, x)
    assert _rel(y, ref) <= SPMV_TOL
    assert np.all(y[[0, 1, 2500]] == 0.0)
    sys_.close()


@pytest.mark.parametrize("kind", ["cd2d", "cd3"])
def test_apply_matches_oracle(kind):
    A = _system(kind)
    F = cmamd.afiim(A, 0.1)
    sys_ = cmamd.System("afiim", A, F)
    v = cmamd.gen_uniform(4, A.n)
    assert _rel(sys_.apply(v), oracle.apply(F, v, mode="E")) <= SPMV_TOL
    sys_.close()


def _compare_solve(variant, A, F, b, nmax, tol=1e-8, **kw):
    sys_ = cmamd.System(variant, A, F, **kw)
    sys_.set_rhs(b)
    sys_.set_x(None)
    rv, it, info, tr = sys_.solve(tol=tol, nmax=nmax, trace_cap=64)
    x = sys_.get_x()
    sys_.close()
    okw = {k: v for k, v in kw.items() if k in ("nr", "mode")}
    xo, rvo, ito, info_o, tro = oracle.bicgstab(variant, A, b, F, tol=tol, nmax=nmax, trace_cap=64, **okw)
    m = min(20, len(tr), len(tro))
    assert m > 0
    assert _trace_close(tr[:m], tro[:m])
    assert info == info_o
    assert abs(it - ito) <= max(1, int(np.ceil(0.02 * ito)))
    return x, rv, it, info, xo


def test_solve_c1_afiim():
    A = _system("cd2d")
    xs = cmamd.gen_uniform(2, A.n)
    b = cmamd.csr_mv_host(A, xs)
    F = cmamd.afiim(A, 0.1)
    x, rv, it, info, xo = _compare_solve("afiim", A, F, b, 500)
    assert info == 0
    assert np.linalg.norm(b - cmamd.csr_mv_host(A, x)) / np.linalg.norm(b) <= 1e-7
    assert _rel(x, xo) <= 1e-6


def test_solve_3d_pdfiim():
    A = _system("cd3")
    xs = cmamd.gen_uniform(6, A.n)
    b = cmamd.csr_mv_host(A, xs)
    F = cmamd.pdfiim(A, 2, 2, 1, 0.05, -1, "I", "B")
    x, rv, it, info, xo = _compare_solve("pdfiim", A, F, b, 1200)
    assert info == 0


@pytest.mark.parametrize("mode,nr", [("E", 1), ("C", 2), ("C", 3)])
def test_solve_3d_acifim(mode, nr):
    A = cmamd.gen_convdiff(3, 16, 5.0, seed=8)
    xs = cmamd.gen_uniform(7, A.n)
    b = cmamd.csr_mv_host(A, xs)
    F = cmamd.acifim(A, 0.01, 0.1)
    sys_ = cmamd.System("acifim", A, F, nr=nr, mode=mode)
    v = cmamd.gen_uniform(8, A.n)
    assert _rel(sys_.apply(v), oracle.apply(F, v, nr=nr, mode=mode)) <= SPMV_TOL
    sys_.close()
    _compare_solve("acifim", A, F, b, 2000, nr=nr, mode=mode)


def test_forced_exits():
    A = _system("cd2d")
    xs = cmamd.gen_uniform(2, A.n)
    b = cmamd.csr_mv_host(A, xs)
    F = cmamd.afiim(A, 0.1)
    # NMAX reached
    x, rv, it, info, _ = _compare_solve("afiim", A, F, b, 2)
    assert (it, info) == (2, 3)
    # tolerance below rounding: the loop ends by stagnation (info 2), as the
    # oracle's does
    x, rv, it, info, _ = _compare_solve("afiim", A, F, b, 500, tol=1e-30)
    assert info == 2


def test_dropin_entry_points():
    A = _system("cd2d")
    xs = cmamd.gen_uniform(2, A.n)
    b = cmamd.csr_mv_host(A, xs)
    F = cmamd.afiim(A, 0.1)
    r = cmamd.afiim_epbicgstab(A, b, F, nmax=500)
    xo, rvo, ito, info_o, _ = oracle.bicgstab("afiim", A, b, F, nmax=500)
    assert r.info == info_o == 0
    assert abs(r.iter - ito) <= 1
    assert _rel(r.x, xo) <= 1e-6


@pytest.mark.parametrize("mode", ["I", "J"])
def test_acifim_triangular_modes(mode):
    A = cmamd.gen_convdiff(3, 12, 5.0, seed=8)
    xs = cmamd.gen_uniform(7, A.n)
    b = cmamd.csr_mv_host(A, xs)
    F = cmamd.acifim(A, 0.01, 0.1)
    sys_ = cmamd.System("acifim", A, F, nr=1, mode=mode)
    v = cmamd.gen_uniform(8, A.n)
    assert _rel(sys_.apply(v), oracle.apply(F, v, mode=mode)) <= SPMV_TOL
    sys_.close()
    _compare_solve("acifim", A, F, b, 2000, nr=1, mode=mode)


@pytest.mark.parametrize("mode", ["I", "J"])
def test_triangular_modes_ignore_stray_entries(mode):
    # entries on the wrong side of L's / U's diagonal are not read, as the
    # TRIANGULAR descriptors of the reference (ACIFIM/src/bicgstab.cpp:170-176)
    # make mkl_sparse_d_trsv skip them
    import scipy.sparse as sp
    A = cmamd.gen_convdiff(3, 10, 5.0, seed=8)
    F = cmamd.acifim(A, 0.01, 0.1)
    rng = np.random.default_rng(9)

    def stray(T, lower):
        M = T.to_scipy().tocoo()
        n = T.n
        r = rng.integers(0, n, 400)
        c = rng.integers(0, n, 400)
        keep = (c > r) if lower else (c < r)
        E = sp.coo_matrix((rng.uniform(-3, 3, keep.sum()), (r[keep], c[keep])), shape=(n, n))
        out = (M + E).tocsr()
        out.sum_duplicates()
        out.sort_indices()
        return cmamd.Csr.from_scipy(out)

    Fs = cmamd.Factors(F.G, F.idv, F.H, stray(F.L, True), stray(F.U, False))
    v = cmamd.gen_uniform(8, A.n)
    sys_ = cmamd.System("acifim", A, Fs, nr=1, mode=mode)
    got = sys_.apply(v)
    sys_.close()
    assert _rel(got, oracle.apply(F, v, mode=mode)) <= SPMV_TOL
    assert _rel(got, oracle.apply(Fs, v, mode=mode)) <= SPMV_TOL


def test_pdfiim_symmetric_branch():
    import scipy.sparse as sp
    A = cmamd.gen_convdiff(3, 14, 0.0, seed=13)  # Pe = 0: symmetric
    Up = cmamd.Csr.from_scipy(sp.triu(A.to_scipy()).tocsr())
    F = cmamd.afiim(A, 0.1)
    Fs = cmamd.Factors(F.G, F.idv, None)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(14, A.n))
    sys_ = cmamd.System("pdfiim", Up, Fs, a_symmetric=True)
    x = cmamd.gen_uniform(15, A.n)
    assert _rel(sys_.spmv("A", x), oracle.csr_mv(Up, x, "sym_upper")) <= SPMV_TOL
    sys_.set_rhs(b)
    sys_.set_x(None)
    rv, it, info, tr = sys_.solve(tol=1e-8, nmax=1200, trace_cap=64)
    sys_.close()
    xo, rvo, ito, info_o, tro = oracle.bicgstab("pdfiim", Up, b, Fs, nmax=1200, issym=True, trace_cap=64)
    m = min(20, len(tr), len(tro))
    assert _trace_close(tr[:m], tro[:m])
    assert info == info_o == 0 and abs(it - ito) <= max(1, int(np.ceil(0.02 * ito)))


@pytest.mark.parametrize("bits", [32, 16])
@pytest.mark.parametrize("case", ["cd2d_afiim", "cd3d_pe5_acifim"])
def test_mixed_precision_refinement(bits, case):
    if case == "cd2d_afiim":
        A = cmamd.gen_convdiff(2, 100, 0.5, seed=1)
        F = cmamd.afiim(A, 0.1)
        variant, tol = "afiim", 1e-8
    else:
        A = cmamd.gen_convdiff(3, 24, 5.0, seed=4)
        F = cmamd.acifim(A, 0.01, 0.1)
        variant, tol = "acifim", 1e-10
    xs = cmamd.gen_uniform(5, A.n)
    b = cmamd.csr_mv_host(A, xs)
    sys_ = cmamd.System(variant, A, F)
    sys_.enable_mixed(bits)
    sys_.set_rhs(b)
    sys_.set_x(None)
    rv, outer, inner, info = sys_.solve_mixed(bits, tol=tol, inner_tol=1e-4, nmax=2000)
    x = sys_.get_x()
    sys_.close()
    true_res = np.linalg.norm(b - cmamd.csr_mv_host(A, x)) / np.linalg.norm(b)
    assert info == 0, (rv, outer, inner, info)
    assert true_res <= tol and abs(true_res - rv) <= 1e-3 * tol + 1e-2 * rv
    xo = oracle.bicgstab(variant, A, b, F, tol=tol, nmax=2000)[0]
    # error against the FP64 oracle solution: both are within tol of the
    # exact solution in the residual sense
    assert _rel(x, xo) <= 1e3 * tol
    assert 1 <= outer <= 6


def test_autotuned_tiles_keep_parity():
    """The tuner re-cuts the SpMV row blocks per matrix; solves afterwards
    meet the same bars against the oracle (FP64 and mixed)."""
    A = cmamd.gen_convdiff(3, 24, 1.0, seed=21)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(22, A.n))
    F = cmamd.pdfiim(A, 2, 2, 1, 0.05, -1, "I", "B")
    sys_ = cmamd.System("pdfiim", A, F)
    sys_.enable_mixed(16)
    tiles = sys_.autotune(reps=5)
    assert tiles["A"] in (1024, 2048, 4096) and tiles["G"] in (1024, 2048, 4096)
    assert tiles["L"] == 0 and tiles["U"] == 0
    x = cmamd.gen_uniform(23, A.n)
    assert _rel(sys_.spmv("A", x), oracle.csr_mv(A, x)) <= SPMV_TOL
    assert _rel(sys_.apply(x), oracle.apply(F, x)) <= SPMV_TOL
    sys_.set_rhs(b)
    sys_.set_x(None)
    rv, it, info, tr = sys_.solve(tol=1e-8, nmax=1200, trace_cap=64)
    xo, rvo, ito, info_o, tro = oracle.bicgstab("pdfiim", A, b, F, tol=1e-8, nmax=1200, trace_cap=64)
    m = min(20, len(tr), len(tro))
    assert _trace_close(tr[:m], tro[:m])
    assert info == info_o == 0 and abs(it - ito) <= max(1, int(np.ceil(0.02 * ito)))
    sys_.set_x(None)
    rv, outer, inner, info = sys_.solve_mixed(16, tol=1e-8, inner_tol=1e-4, nmax=1200)
    xm = sys_.get_x()
    sys_.close()
    assert info == 0
    assert np.linalg.norm(b - cmamd.csr_mv_host(A, xm)) / np.linalg.norm(b) <= 1e-8


def test_autotuned_mode_c_sweeps_keep_parity():
    """ACIFIM apply mode C with two sweeps multiplies by L and U, so the
    tuner times their SpMV formats too (the diagonal format, on a stencil's
    factors); the solve afterwards meets the oracle's trace, iterations and
    exit, and the FP32-inner refinement reaches the tolerance."""
    A = cmamd.gen_convdiff(3, 24, 5.0, seed=31)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(32, A.n))
    F = cmamd.acifim(A, 0.01, 0.1)
    sys_ = cmamd.System("acifim", A, F, nr=2, mode="C")
    sys_.enable_mixed(32)
    sys_.autotune(reps=3)
    fm = sys_.spmv_formats()
    assert fm[64]["L"] != "staged" or fm[64]["U"] != "staged", fm
    x = cmamd.gen_uniform(33, A.n)
    assert _rel(sys_.apply(x), oracle.apply(F, x, nr=2, mode="C")) <= SPMV_TOL
    sys_.set_rhs(b)
    sys_.set_x(None)
    rv, it, info, tr = sys_.solve(tol=1e-10, nmax=2000, trace_cap=64)
    xo, rvo, ito, info_o, tro = oracle.bicgstab("acifim", A, b, F, tol=1e-10, nmax=2000, nr=2, mode="C",
                                                trace_cap=64)
    m = min(20, len(tr), len(tro))
    assert _trace_close(tr[:m], tro[:m])
    assert info == info_o == 0 and abs(it - ito) <= max(1, int(np.ceil(0.02 * ito)))
    sys_.set_x(None)
    rv, outer, inner, info = sys_.solve_mixed(32, tol=1e-10, inner_tol=1e-4, nmax=2000)
    xm = sys_.get_x()
    sys_.close()
    assert info == 0
    assert np.linalg.norm(b - cmamd.csr_mv_host(A, xm)) / np.linalg.norm(b) <= 1e-10


def _slot_case(kind):
    import scipy.sparse as sp
    if kind == "afiim":
        A = _system("cd2d")
        return "afiim", A, A, cmamd.afiim(A, 0.1), {}, False
    if kind == "pdfiim_sym":
        A = cmamd.gen_convdiff(3, 14, 0.0, seed=13)  # Pe = 0: symmetric
        Up = cmamd.Csr.from_scipy(sp.triu(A.to_scipy()).tocsr())
        F = cmamd.afiim(A, 0.1)
        return "pdfiim", A, Up, cmamd.Factors(F.G, F.idv, None), {"a_symmetric": True}, True
    A = cmamd.gen_convdiff(3, 16, 5.0, seed=8)
    return "acifim", A, A, cmamd.acifim(A, 0.01, 0.1), {"nr": 2, "mode": "C"}, False


@pytest.mark.parametrize("fmt", ["slot", "slot16", "dia", "dia2"])
@pytest.mark.parametrize("kind", ["afiim", "pdfiim_sym", "acifim"])
def test_slot_major_format_keeps_parity(kind, fmt):
    """Every SpMV matrix (A, G, H, G^T) in every precision switched to the
    slot-major format (one row per thread, rows ordered by length in blocks
    of 256) or to the diagonal format (dense diagonals plus a CSR
    remainder): the FP64 products, preconditioner apply and residual trace
    meet the oracle bars above, and the mixed solves reach the tolerance."""
    prec, A, Aup, F, kw, sym = _slot_case(kind)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(32, A.n))
    sys_ = cmamd.System(prec, Aup, F, **kw)
    sys_.enable_mixed(32)
    sys_.enable_mixed(16)
    switched = 0
    for bits in (64, 32, 16):
        for which in ("A", "G", "H", "GT"):
            try:
                sys_.spmv_format(bits, which, fmt)
                switched += 1
            except cmamd.EngineError:
                pass  # absent matrix (H of the symmetric branch, G^T elsewhere)
    f = sys_.spmv_formats()
    assert f[64]["A"] == fmt and f[32]["A"] == fmt and f[16]["A"] == fmt and f[64]["G"] == fmt
    assert switched >= 9
    x = cmamd.gen_uniform(33, A.n)
    if sym:
        assert _rel(sys_.spmv("A", x), oracle.csr_mv(Aup, x, "sym_upper")) <= SPMV_TOL
    else:
        assert _rel(sys_.spmv("A", x), oracle.csr_mv(A, x)) <= SPMV_TOL
    okw = {k: v for k, v in kw.items() if k in ("nr", "mode")}
    # symmetric branch: M = G D^-1 G^T, so the oracle's apply takes H = G^T
    Fo = cmamd.Factors(F.G, F.idv, cmamd.Csr.from_scipy(F.G.to_scipy().T.tocsr())) if sym else F
    assert _rel(sys_.apply(x), oracle.apply(Fo, x, **okw)) <= SPMV_TOL
    sys_.set_rhs(b)
    sys_.set_x(None)
    rv, it, info, tr = sys_.solve(tol=1e-8, nmax=2000, trace_cap=64)
    xo, rvo, ito, info_o, tro = oracle.bicgstab(prec, Aup, b, F, tol=1e-8, nmax=2000, trace_cap=64,
                                                **({"issym": True} if sym else okw))
    m = min(20, len(tr), len(tro))
    assert _trace_close(tr[:m], tro[:m])
    assert info == info_o == 0 and abs(it - ito) <= max(1, int(np.ceil(0.02 * ito)))
    for bits in (32, 16):
        sys_.set_x(None)
        rv, outer, inner, info = sys_.solve_mixed(bits, tol=1e-8, inner_tol=1e-4, nmax=2000)
        xm = sys_.get_x()
        assert info == 0
        assert np.linalg.norm(b - cmamd.csr_mv_host(A, xm)) / np.linalg.norm(b) <= 1e-8
    # back to the staged kernels: the same products
    for which in ("A", "G"):
        sys_.spmv_format(64, which, "staged")
    assert _rel(sys_.spmv("A", x), oracle.csr_mv(Aup, x, "sym_upper") if sym else oracle.csr_mv(A, x)) <= SPMV_TOL
    sys_.close()


def test_dia_format_odd_row_count():
    """Diagonal format on a matrix with an odd row count (the FP16 copies
    take two rows per thread; the last thread has one): FP64 products and
    apply against the oracle, and FP32 / FP16 refinement to the tolerance."""
    A = cmamd.gen_convdiff(2, 77, 1.0, seed=41)  # n = 5929
    F = cmamd.afiim(A, 0.1)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(42, A.n))
    sys_ = cmamd.System("afiim", A, F)
    sys_.enable_mixed(32)
    sys_.enable_mixed(16)
    for bits in (64, 32, 16):
        for which in ("A", "G"):
            sys_.spmv_format(bits, which, "dia" if bits == 64 else "dia2")
    x = cmamd.gen_uniform(43, A.n)
    assert _rel(sys_.spmv("A", x), oracle.csr_mv(A, x)) <= SPMV_TOL
    assert _rel(sys_.apply(x), oracle.apply(F, x)) <= SPMV_TOL
    for bits in (32, 16):
        sys_.set_rhs(b)
        sys_.set_x(None)
        rv, outer, inner, info = sys_.solve_mixed(bits, tol=1e-8, inner_tol=1e-4, nmax=2000)
        xm = sys_.get_x()
        assert info == 0
        assert np.linalg.norm(b - cmamd.csr_mv_host(A, xm)) / np.linalg.norm(b) <= 1e-8
    sys_.close()


def test_slot_major_format_refuses_long_rows():
    """Rows longer than 64 nonzeros keep the staged format; the autotuner
    then never offers the slot-major one for that matrix."""
    A = _system("irregular")
    F = cmamd.afiim(A, 0.1)
    sys_ = cmamd.System("afiim", A, F)
    with pytest.raises(cmamd.EngineError):
        sys_.spmv_format(64, "A", "slot")
    sys_.autotune(reps=2)
    assert sys_.spmv_formats()[64]["A"] not in ("slot", "slot16")
    sys_.close()


def test_sampled_profiling_counts_and_results():
    """Sampled profiling brackets only the first passes of each solve and
    replays the rest from graphs: the counted launches match, and the
    solution is bit-identical to an unprofiled solve."""
    A = _system("cd2d")
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(2, A.n))
    F = cmamd.afiim(A, 0.1)
    sys_ = cmamd.System("afiim", A, F)
    sys_.set_rhs(b)
    sys_.set_x(None)
    rv0, it0, info0, _ = sys_.solve(tol=1e-8, nmax=500)
    x0 = sys_.get_x()
    for _ in range(2):
        sys_.profile(True, cats=("F1",), sample=5)
        sys_.set_x(None)
        rv, it, info, _ = sys_.solve(tol=1e-8, nmax=500)
        st = sys_.kernel_stats()
        assert (rv, it, info) == (rv0, it0, info0)
        assert np.array_equal(sys_.get_x(), x0)
        assert st["F1"][0] == 2 * 5 and st["A"][0] == 0  # two factor applies per pass, 5 passes
    sys_.profile(False)
    sys_.close()


def test_algorithmic_bytes_do_not_depend_on_format():
    """The per-launch algorithmic bytes of the H products (with D^{-1}) are
    the CSR figure nnz(H)*(8 + 4) + 4*(n+1) + 3 FP64 n-vectors in every
    storage format; the format's own bytes differ from it (row-block table
    added for the staged kernel, diagonal zeros in and diagonal column
    indices out for the diagonal one)."""
    A = _system("cd2d")
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(2, A.n))
    F = cmamd.afiim(A, 0.1)
    sys_ = cmamd.System("afiim", A, F)
    sys_.set_rhs(b)
    n, nnz_h = A.n, F.H.nnz
    alg = nnz_h * 12.0 + 4.0 * (n + 1) + 3 * 8.0 * n
    seen = {}
    for fmt in ("staged", "dia", "slot"):
        sys_.spmv_format(64, "H", fmt)
        sys_.profile(True, cats=("F1",))
        sys_.set_x(None)
        sys_.solve(tol=1e-8, nmax=500)
        nl, _, by = sys_.kernel_stats()["F1"]
        fb = sys_.kernel_format_bytes()["F1"]
        assert nl > 0
        assert by / nl == pytest.approx(alg, rel=1e-12)
        seen[fmt] = fb / nl
        assert seen[fmt] != pytest.approx(alg, rel=1e-9)
    assert seen["staged"] > alg  # CSR plus the row-block table
    sys_.profile(False)
    sys_.close()


def test_graph_replays_match_eager_single():
    """Single-RHS passes replayed from captured hipGraphs against kernel-by-
    kernel launches: bit-identical traces and solutions over several replays,
    with other eager work (a profiled solve, a mixed solve) in between."""
    A = _system("cd2d")
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(2, A.n))
    F = cmamd.afiim(A, 0.1)
    sys_ = cmamd.System("afiim", A, F)
    sys_.enable_mixed(32)
    sys_.set_rhs(b)
    ref = None
    for graphs in (False, True, True, True):
        sys_.graphs(graphs)
        sys_.set_x(None)
        rv, it, info, tr = sys_.solve(tol=1e-8, nmax=500, trace_cap=64)
        x = sys_.get_x()
        sys_.set_x(None)
        mx = sys_.solve_mixed(32, tol=1e-8, inner_tol=1e-4, nmax=500)
        xm = sys_.get_x()
        sys_.profile(True)  # eager, event-bracketed solve between replays
        sys_.set_x(None)
        sys_.solve(tol=1e-8, nmax=500)
        sys_.profile(False)
        got = (x, rv, it, info, tr, xm, mx)
        if ref is None:
            ref = got
            continue
        assert np.array_equal(got[0], ref[0]) and got[1:4] == ref[1:4] and np.array_equal(got[4], ref[4])
        assert np.array_equal(got[5], ref[5]) and got[6] == ref[6]
    sys_.close()


def test_mixed_with_wide_column_span():
    """Low-precision products use block-relative 16-bit column indices when a
    row block's columns span less than 65536; a matrix with couplings 70,000
    columns apart keeps 32-bit indices for the blocks' matrix.  Both must give
    the same refinement result quality."""
    import scipy.sparse as sp
    base = cmamd.gen_convdiff(2, 290, 1.0, seed=61).to_scipy().tolil()  # n = 84100
    n = base.shape[0]
    rng = np.random.default_rng(5)
    for i in rng.choice(n - 70000, size=40, replace=False):
        base[i, i + 70000] = -0.01
        base[i + 70000, i] = -0.01
    A = cmamd.Csr.from_scipy(sp.csr_matrix(base))
    F = cmamd.afiim(A, 0.1)
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(62, n))
    for bits in (32, 16):
        sys_ = cmamd.System("afiim", A, F)
        sys_.enable_mixed(bits)
        sys_.set_rhs(b)
        sys_.set_x(None)
        rv, outer, inner, info = sys_.solve_mixed(bits, tol=1e-8, inner_tol=1e-4, nmax=2000)
        x = sys_.get_x()
        sys_.close()
        assert info == 0
        assert np.linalg.norm(b - cmamd.csr_mv_host(A, x)) / np.linalg.norm(b) <= 1e-8
        xo = oracle.bicgstab("afiim", A, b, F, tol=1e-8, nmax=2000)[0]
        assert _rel(x, xo) <= 1e-4
