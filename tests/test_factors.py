"""Host factor producers (libcmhost.so) against the reference's own builds.

AFIIM and ACIFIM: bit-exact against the reference's afiim()/acifim()
(compiled unmodified into oracle/_ref, when present) and against the committed
SHA-256 digests those builds produced (tests/golden/digests.json, always).
PDFIIM (its reference construction needs MKL headers and is not built here):
pinned by what the mathematics fixes -- full patterns reproduce A^{-1}
exactly, one-element patterns follow a closed form -- plus layout invariants
and thread-count independence.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import cmamd
import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _digest_csr(m):
    h = hashlib.sha256()
    for a in (m.ptr, m.idx, m.val):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _load_digests():
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)


def _sherman2():
    d = np.load(os.path.join(GOLDEN, "sherman2.npz"))
    return cmamd.Csr(len(d["ptr"]) - 1, d["ptr"], d["idx"], d["val"]), d["rhs"]


def _inputs():
    sh, _ = _sherman2()
    return {
        "sherman2": sh,
        "convdiff2d_100_pe0.5_s1": cmamd.gen_convdiff(2, 100, 0.5, seed=1),
        "convdiff3d_16_pe5_s8": cmamd.gen_convdiff(3, 16, 5.0, seed=8),
        "stencil27_10_s9": cmamd.gen_stencil27(10, 0.05, 200, 600, seed=9),
    }


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


def _csr_equal(a, b):
    return _same(a.ptr, b.ptr) and _same(a.idx, b.idx) and _same(a.val, b.val)


DIG = _load_digests()


def test_generated_inputs_match_digests():
    ins = _inputs()
    for name, m in ins.items():
        assert _digest_csr(m) == DIG["inputs"][name], name


@pytest.mark.parametrize("key", sorted(DIG["afiim"]))
def test_afiim_matches_reference_digests(key):
    name, dtol = key.split("|")
    A = _inputs()[name]
    F = cmamd.afiim(A, float(dtol))
    want = DIG["afiim"][key]
    assert _digest_csr(F.G) == want["G"]
    assert _digest_csr(F.H) == want["H"]
    assert hashlib.sha256(F.idv.tobytes()).hexdigest() == want["idv"]


@pytest.mark.parametrize("key", sorted(DIG["acifim"]))
def test_acifim_matches_reference_digests(key):
    name, d1, d2, el = key.split("|")
    A = _inputs()[name]
    F = cmamd.acifim(A, float(d1), float(d2), element_limit=int(el))
    want = DIG["acifim"][key]
    for k in "GHLU":
        assert _digest_csr(getattr(F, k)) == want[k], k
    assert hashlib.sha256(F.idv.tobytes()).hexdigest() == want["idv"]


def test_full_factor_fixtures():
    sh, _ = _sherman2()
    d = np.load(os.path.join(GOLDEN, "afiim_sherman2.npz"))
    F = cmamd.afiim(sh, float(d["dtol"]))
    assert _same(F.G.ptr, d["gp"]) and _same(F.G.idx, d["gi"]) and _same(F.G.val, d["gv"])
    assert _same(F.H.ptr, d["hp"]) and _same(F.H.idx, d["hi"]) and _same(F.H.val, d["hv"])
    assert _same(F.idv, d["idv"])
    d = np.load(os.path.join(GOLDEN, "acifim_sherman2.npz"))
    F = cmamd.acifim(sh, float(d["dtol1"]), float(d["dtol2"]))
    for k, p in (("G", "g"), ("H", "h"), ("L", "l"), ("U", "u")):
        m = getattr(F, k)
        assert _same(m.ptr, d[p + "p"]) and _same(m.idx, d[p + "i"]) and _same(m.val, d[p + "v"]), k


def _random_system(n, seed, density=0.08):
    rng = np.random.default_rng(seed)
    M = (rng.random((n, n)) < density) * rng.uniform(-1, 1, (n, n))
    M += np.diag(np.abs(M).sum(1) + 1.0)
    import scipy.sparse as sp
    return cmamd.Csr.from_scipy(sp.csr_matrix(M))


@pytest.mark.skipif(not oracle.ref_available("afiim"), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("dtol", [0.0, 0.01, 0.2])
def test_afiim_bitexact_against_reference_build(seed, dtol):
    A = _random_system(150, seed)
    F = cmamd.afiim(A, dtol)
    R = oracle.ref_afiim(A, dtol)
    assert _csr_equal(F.G, R.G) and _csr_equal(F.H, R.H) and _same(F.idv, R.idv)


@pytest.mark.skipif(not oracle.ref_available("acifim"), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("d1,d2,el", [(0.0, 0.0, 200), (0.01, 0.05, 200), (0.05, 0.05, 3)])
def test_acifim_bitexact_against_reference_build(seed, d1, d2, el):
    A = _random_system(150, seed)
    F = cmamd.acifim(A, d1, d2, element_limit=el)
    rc, R = oracle.ref_acifim(A, d1, d2, element_limit=el)
    assert rc == 0
    for k in "GHLU":
        assert _csr_equal(getattr(F, k), getattr(R, k)), k
    assert _same(F.idv, R.idv)


def test_acifim_nan_schur_returns_error():
    # a NaN in A reaches the Schur complement of row 1: acifim returns 1
    A = cmamd.Csr(3, np.array([0, 2, 4, 5], np.int32), np.array([0, 1, 0, 1, 2], np.int32),
                  np.array([2.0, 1.0, np.nan, 3.0, 1.0]))
    with pytest.raises(cmamd.EngineError):
        cmamd.acifim(A, 0.01, 0.01)
    if oracle.ref_available("acifim"):
        rc, _ = oracle.ref_acifim(A, 0.01, 0.01)
        assert rc == 1


def _dense(m):
    return m.to_scipy().toarray()


def _check_layout(F, lu=False):
    n = F.G.n
    for M, upper, diag_first in ((F.G, True, True), (F.H, False, False)) + (
            ((F.U, True, True), (F.L, False, False)) if lu else ()):
        for i in range(n):
            cols = M.idx[M.ptr[i]:M.ptr[i + 1]]
            vals = M.val[M.ptr[i]:M.ptr[i + 1]]
            assert np.all(np.diff(cols) > 0)
            if diag_first:
                assert cols[0] == i and vals[0] == 1.0
            else:
                assert cols[-1] == i and vals[-1] == 1.0
            assert np.all(cols >= i) if upper else np.all(cols <= i)


def test_factor_layout_invariants():
    A = cmamd.gen_convdiff(2, 12, 1.0, seed=4)
    _check_layout(cmamd.afiim(A, 0.05))
    _check_layout(cmamd.acifim(A, 0.01, 0.05), lu=True)
    _check_layout(cmamd.pdfiim(A, 2, 2, 1, 0.05))


def test_afiim_zero_tolerance_is_exact_inverse():
    # With no dropping the bordering recursion is exact: G D^{-1} H = A^{-1}.
    A = _random_system(40, 7, density=0.3)
    F = cmamd.afiim(A, 0.0)
    M = _dense(F.G) @ np.diag(F.idv) @ _dense(F.H)
    assert np.linalg.norm(M @ _dense(A) - np.eye(A.n)) < 1e-10


def test_pdfiim_full_pattern_is_exact_inverse():
    # Dense A: the first pattern of row i is already {0..i-1}; with ftol = 0
    # nothing is dropped, each small solve is exact and M = A^{-1}.
    rng = np.random.default_rng(3)
    n = 12
    M = rng.uniform(-1, 1, (n, n)) + np.diag(np.full(n, 2.0 * n))
    import scipy.sparse as sp
    A = cmamd.Csr.from_scipy(sp.csr_matrix(M))
    F = cmamd.pdfiim(A, 1, 1, 1, 0.0, -1, "I", "B")
    Minv = _dense(F.G) @ np.diag(F.idv) @ _dense(F.H)
    assert np.linalg.norm(Minv @ M - np.eye(n)) < 1e-10


def test_pdfiim_tridiagonal_closed_form():
    # Tridiagonal A, one pass: the pattern of row i is {i-1}, so
    # g = -A[i-1,i]/A[i-1,i-1], h = -A[i,i-1]/A[i-1,i-1] and
    # D(i) = 1/(h (A[i-1,i-1] g + A[i-1,i]) + A[i,i-1] g + A[i,i]).
    n = 9
    rng = np.random.default_rng(5)
    lo, di, up = rng.uniform(-1, 1, n - 1), rng.uniform(3, 4, n), rng.uniform(-1, 1, n - 1)
    T = np.diag(di) + np.diag(lo, -1) + np.diag(up, 1)
    import scipy.sparse as sp
    A = cmamd.Csr.from_scipy(sp.csr_matrix(T))
    F = cmamd.pdfiim(A, 1, 1, 1, 0.0, -1, "I", "B")
    G, H = _dense(F.G), _dense(F.H)
    assert F.idv[0] == 1.0 / T[0, 0]
    for i in range(1, n):
        g = -T[i - 1, i] / T[i - 1, i - 1]
        h = -T[i, i - 1] / T[i - 1, i - 1]
        s = h * (T[i - 1, i - 1] * g + T[i - 1, i]) + (T[i, i - 1] * g + T[i, i])
        assert G[i - 1, i] == pytest.approx(g, rel=1e-15)
        assert H[i, i - 1] == pytest.approx(h, rel=1e-15)
        assert F.idv[i] == pytest.approx(1.0 / s, rel=1e-14)


def test_pdfiim_thread_count_independent():
    A = cmamd.gen_convdiff(3, 10, 1.0, seed=2)
    F1 = cmamd.pdfiim(A, 2, 2, 1, 0.05, nthreads=1)
    F8 = cmamd.pdfiim(A, 2, 2, 1, 0.05, nthreads=8)
    assert _csr_equal(F1.G, F8.G) and _csr_equal(F1.H, F8.H) and _same(F1.idv, F8.idv)


def test_pdfiim_is_a_useful_preconditioner():
    A = cmamd.gen_convdiff(3, 12, 1.0, seed=2)
    F = cmamd.pdfiim(A, 2, 2, 1, 0.05)
    M = _dense(F.G) @ np.diag(F.idv) @ _dense(F.H)
    Ad = _dense(A)
    # the preconditioned operator is much closer to I than the Jacobi one
    e_m = np.linalg.norm(M @ Ad - np.eye(A.n)) / np.sqrt(A.n)
    e_j = np.linalg.norm(np.diag(1 / np.diag(Ad)) @ Ad - np.eye(A.n)) / np.sqrt(A.n)
    assert e_m < 0.5 * e_j


def test_pdfiim_rejects_undefined_modes():
    A = cmamd.gen_convdiff(2, 5, 1.0, seed=2)
    with pytest.raises(cmamd.EngineError):
        cmamd.pdfiim(A, 1, 2, 1, 0.05, patt="L")
    F = cmamd.pdfiim(A, 1, 1, 1, 0.05, patt="U")
    _check_layout(F)
