"""The diagonal index built on the device (build_dia_device in the engine)
against the selection rule restated here in numpy: the diagonals held by at
least 60% of the rows, the 32 most populated (ties: the smaller offset), in
offset order, every other entry in the CSR remainder.  The matrices cover
stencils (all entries on diagonals), a factor with a remainder, a partly
filled diagonal on each side of the 60% bar, a remainder too dense to stage
through LDS, more than 32 qualifying diagonals, nothing qualifying, and the
two cases the device leaves to the
host build (a sampled row longer than 64 entries, a remainder row longer
than 64), whose index must follow the same rule.  Products through the
index, in both diagonal variants, match the host CSR product and each other
bit for bit, and a one-shot solve (which picks the
diagonal format from this index without timing) matches the oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

import cmamd
import oracle

pytestmark = pytest.mark.gpu


def _rule(A):
    """(offsets, remainder entries) by the selection rule."""
    rows = np.repeat(np.arange(A.n), np.diff(A.ptr))
    off = A.idx.astype(np.int64) - rows
    u, c = np.unique(off, return_counts=True)
    keep = c >= 0.6 * A.n
    u, c = u[keep], c[keep]
    order = np.lexsort((u, -c))[:32]
    sel = np.sort(u[order])
    return sel, int(A.nnz - np.isin(off, sel).sum())


def _csr(M):
    M = sp.csr_matrix(M)
    M.sum_duplicates()
    M.sort_indices()
    return cmamd.Csr.from_scipy(M)


def _with_diag(n, rng, diags, scatter=0):
    """Full main diagonal, diags = {offset: fill}, plus scattered entries."""
    rows, cols = [np.arange(n)], [np.arange(n)]
    for o, fill in diags.items():
        i = np.arange(max(0, -o), min(n, n - o))
        i = i[rng.random(len(i)) < fill]
        rows.append(i)
        cols.append(i + o)
    if scatter:
        rows.append(rng.integers(0, n, scatter))
        cols.append(rng.integers(0, n, scatter))
    r, c = np.concatenate(rows), np.concatenate(cols)
    v = rng.uniform(-1, 1, len(r))
    v[: n] = 50.0
    return _csr(sp.coo_matrix((v, (r, c)), shape=(n, n)))


def _cases():
    rng = np.random.default_rng(11)
    A3 = cmamd.gen_convdiff(3, 20, 1.0, seed=2)
    F3 = cmamd.pdfiim(A3, 2, 2, 1, 0.05, -1, "I", "B")
    n = 5000
    band = _csr(sp.diags([np.full(n - abs(o), 1.0 + 0.01 * o) for o in range(-20, 20)], list(range(-20, 20)),
                         shape=(n, n)) + sp.eye(n) * 100)
    long_sample = _with_diag(n, rng, {1: 1.0, -1: 1.0})  # row 0 is sampled: give it 100 entries
    M = long_sample.to_scipy().tolil()
    M[0, rng.choice(np.arange(2, n), 100, replace=False)] = 0.5
    long_sample = _csr(M)
    long_rem = _with_diag(n, rng, {1: 1.0, -1: 1.0})  # a remainder row of 80 entries off the sample grid
    M = long_rem.to_scipy().tolil()
    M[7, rng.choice(np.arange(10, n), 80, replace=False)] = 0.5
    long_rem = _csr(M)
    return {
        "conv3d": A3,
        # 32^3 rows: evenly spaced samples (every 32nd row) would all sit on the x = 0 face, which
        # lacks the -1 neighbour
        "conv3d_pow2": cmamd.gen_convdiff(3, 32, 1.0, seed=5),
        "conv2d": cmamd.gen_convdiff(2, 50, 0.5, seed=3),
        "pdfiim_H": F3.H,
        "partial_fill": _with_diag(n, rng, {1: 1.0, -1: 1.0, 5: 0.7, -7: 0.55, 300: 0.62}, scatter=3000),
        # ~8 remainder entries per row: a workgroup's remainder range exceeds what the product stages
        # through LDS, so its rows walk their own entries
        "dense_remainder": _with_diag(n, rng, {1: 1.0, -1: 1.0}, scatter=8 * n),
        "band40": band,
        "nothing": _csr(sp.random(n, n, 4e-3, random_state=5)),
        "long_sample_row": long_sample,
        "long_remainder_row": long_rem,
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_device_dia_index_follows_rule(name):
    M = _cases()[name]
    want_off, want_rem = _rule(M)
    idv = np.ones(M.n)
    s = cmamd.System("afiim", M, cmamd.Factors(M, idv, M))
    off, nrem = s.dia_info("A")
    if name == "nothing":
        assert len(want_off) == 0
    assert np.array_equal(off, want_off), (off, want_off)
    if len(off):
        assert nrem == want_rem
        x = cmamd.gen_uniform(4, M.n)
        ref = cmamd.csr_mv_host(M, x)
        ys = []
        for fmt in ("dia", "dia2"):  # one row per thread; two rows per thread, 9 diagonals in flight
            s.spmv_format(64, "A", fmt)
            ys.append(s.spmv("A", x))
            assert np.linalg.norm(ys[-1] - ref) <= 1e-13 * np.linalg.norm(ref), fmt
        # each row's sum runs in the same order in both variants
        assert np.array_equal(ys[0], ys[1])
    s.close()


def test_oneshot_solve_in_diagonal_format_matches_oracle():
    A = cmamd.gen_convdiff(3, 24, 1.0, seed=2)
    F = cmamd.pdfiim(A, 2, 2, 1, 0.05, -1, "I", "B")
    xs = cmamd.gen_uniform(3, A.n)
    b = cmamd.csr_mv_host(A, xs)
    r = cmamd.pdfiim_epbicgstab(A, b, F, tol=1e-8, nmax=1200)
    xo, _, it, info, _ = oracle.bicgstab("pdfiim", A, b, F, tol=1e-8, nmax=1200)
    assert r.info == info == 0
    assert abs(r.iter - it) <= 1
    assert np.linalg.norm(r.x - xo) <= 1e-8 * np.linalg.norm(xo)
