"""GPU parity of the persistent batched solver (many small systems of uneven
size in one launch, one workgroup per solve, device work queue).

Every item runs the reference loop (AFIIM/src/bicgstab.cpp:149-281) alone
in its workgroup, so each is held to the single-solve bars against the
oracle: same info code, iteration count within 2% (at least 1), final
relative residual within 1e-8 relative (+1e-12 floor) when the counts agree,
solution within 1e-6.  Systems: 2D/3D convection-diffusion of different
sizes (AFIIM), ACIFIM mode E, the symmetric PDFIIM branch (G^T), one system
repeated with several right-hand sides, nonzero initial guesses, NMAX exits.
The result of an item must not depend on the rest of the batch (bitwise).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import cmamd
import oracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _mix():
    out = []
    for d, N, pe, seed in ((2, 40, 0.5, 71), (3, 12, 3.0, 72), (2, 110, 1.0, 73), (3, 20, 0.2, 74)):
        A = cmamd.gen_convdiff(d, N, pe, seed=seed)
        F = cmamd.afiim(A, 0.1)
        out.append(("afiim", A, F, {}, cmamd.System("afiim", A, F)))
    A = cmamd.gen_convdiff(3, 14, 5.0, seed=75)
    F = cmamd.acifim(A, 0.01, 0.1)
    out.append(("acifim", A, F, {"nr": 1, "mode": "E"}, cmamd.System("acifim", A, F, nr=1, mode="E")))
    A = cmamd.gen_convdiff(3, 13, 0.0, seed=76)
    Up = cmamd.Csr.from_scipy(sp.triu(A.to_scipy()).tocsr())
    Fs = cmamd.Factors(cmamd.afiim(A, 0.1).G, cmamd.afiim(A, 0.1).idv, None)
    out.append(("pdfiim", Up, Fs, {"issym": True}, cmamd.System("pdfiim", Up, Fs, a_symmetric=True)))
    return out


def _full(variant, A, kw):
    if kw.get("issym"):
        M = A.to_scipy()
        return cmamd.Csr.from_scipy((M + sp.triu(M, 1).T).tocsr())
    return A


def test_batch_matches_oracle():
    mix = _mix()
    rng = np.random.default_rng(3)
    items, meta = [], []
    for q, (variant, A, F, kw, h) in enumerate(mix):
        Af = _full(variant, A, kw)
        for j in range(3 if q == 2 else 1):
            b = cmamd.csr_mv_host(Af, cmamd.gen_uniform(300 + 10 * q + j, A.n))
            x0 = None if j < 2 else rng.uniform(-1, 1, A.n)
            nmax = 2000 if variant == "acifim" else 1000
            items.append((h, b, x0, 1e-8, nmax))
            meta.append((variant, A, F, kw, b, x0, nmax))
    # an NMAX exit
    variant, A, F, kw, h = mix[0]
    b = cmamd.csr_mv_host(A, cmamd.gen_uniform(399, A.n))
    items.append((h, b, None, 1e-8, 3))
    meta.append((variant, A, F, kw, b, None, 3))
    res = cmamd.batch_solve(items)
    for r, (variant, A, F, kw, b, x0, nmax) in zip(res, meta):
        okw = {k: v for k, v in kw.items() if k in ("nr", "mode", "issym")}
        xo, rvo, ito, info_o, _ = oracle.bicgstab(variant, A, b, F, x0=x0, tol=1e-8, nmax=nmax, **okw)
        assert r.info == info_o, (variant, A.n, r.info, info_o)
        assert abs(r.iter - ito) <= max(1, int(np.ceil(0.02 * ito))), (variant, A.n, r.iter, ito)
        if r.iter == ito and ito <= 20:
            assert abs(r.resval - rvo) <= 1e-8 * rvo + 1e-12
        if info_o == 0:
            assert _rel(r.x, xo) <= 1e-6
    assert res[-1].iter == 3 and res[-1].info == 3
    # the same item alone gives the same bits
    alone = cmamd.batch_solve([items[3]])[0]
    assert np.array_equal(alone.x, res[3].x) and alone.iter == res[3].iter
    for *_, h in mix:
        h.close()


def test_batch_true_residual_restart():
    mix = _mix()[:4]
    items, meta = [], []
    for q, (variant, A, F, kw, h) in enumerate(mix):
        b = cmamd.csr_mv_host(A, cmamd.gen_uniform(500 + q, A.n))
        items.append((h, b, None, 1e-8, 1000))
        meta.append((A, b))
    res = cmamd.batch_solve(items, restart_true=True)
    for r, (A, b) in zip(res, meta):
        assert r.info == 0
        assert np.linalg.norm(b - cmamd.csr_mv_host(A, r.x)) / np.linalg.norm(b) <= 1e-8
    for *_, h in mix:
        h.close()
