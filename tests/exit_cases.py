"""Hand-built systems that drive the reference loop into each of its exits
(shared by the oracle tests and the GPU tests; not a test module).

Divergence (info 1, AFIIM/src/bicgstab.cpp:219-234): A = I, G = H = I and
IDv = diag(d) with d = (0, .., 0, 1^p, 2^p, .., (n-z)^p), so M A = diag(d) is
singular: the z leading components of r can never be reduced, and with
b = (a, .., a, 1, .., 1) the BiCG coefficients keep feeding them back, so the
residual grows from the first pass on, by several percent per pass (far above
rounding).  The loop must stop when msteps reaches maxmsteps: n / 50 clipped
to [10, 20] for AFIIM / PDFIIM (AFIIM/src/bicgstab.cpp:114,
PDFIIM/src/bicgstab.cpp:106), and 40 for ACIFIM (ACIFIM/src/bicgstab.cpp:132)
on the same system.
"""
import numpy as np

import cmamd

# (n, p, a, z): each grows on every pass through pass 41 (textbook trace)
DIVERGING = {"d64": (64, 1.5, 16.0, 4), "d600": (600, 0.5, 32.0, 4), "d1024": (1024, 1.5, 32.0, 4)}


def identity(n):
    return cmamd.Csr(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n))


def diverging(name):
    """(A, b, F, d) of one diverging system: A = I, M = diag(d)."""
    n, p, a, z = DIVERGING[name]
    d = np.r_[np.zeros(z), np.arange(1, n - z + 1, dtype=np.float64) ** p]
    b = np.ones(n)
    b[:z] = a
    I = identity(n)
    return I, b, cmamd.Factors(I, d, I, I, I), d


def maxmsteps(variant, n):
    """The reference's divergence limit, from its own text (see module doc)."""
    if variant == "acifim":
        return min(max(n, 2000) // 50, 40)
    return min(max(n, 500) // 50, 20)


def textbook_norms(d, b, passes):
    """||r_k||, k = 0..passes, of right-preconditioned BiCGSTAB (van der Vorst
    1992) on A M = diag(d), written independently of the oracle."""
    r = b.astype(np.float64).copy()
    rh = r.copy()
    rho_old = alpha = omega = 1.0
    v = np.zeros_like(r)
    p = np.zeros_like(r)
    out = [np.linalg.norm(r)]
    for _ in range(passes):
        rho = rh @ r
        beta = (rho / rho_old) * (alpha / omega)
        p = r + beta * (p - omega * v)
        v = d * p
        alpha = rho / (rh @ v)
        s = r - alpha * v
        t = d * s
        omega = (t @ s) / (t @ t)
        r = s - omega * t
        rho_old = rho
        out.append(np.linalg.norm(r))
    return np.array(out)


def divergence_exit(norms, limit):
    """(iter, resval) at which the reference loop leaves with info 1: in pass
    k (iter = k - 1 when it starts) it compares the lagged ||r_{k-1}|| with
    ||r_{k-2}|| (||r_0|| in pass 1) and stops once `limit` passes in a row
    grew (AFIIM/src/bicgstab.cpp:210-234)."""
    m = 0
    for k in range(1, len(norms) + 1):
        prc = norms[k - 2] if k >= 2 else norms[0]
        rc = norms[k - 1]
        m = m + 1 if prc < rc else 0
        if m >= limit:
            return k - 1, rc / norms[0]
    raise ValueError("no divergence exit within the trace")
