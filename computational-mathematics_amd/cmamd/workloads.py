"""Workload definitions shared by the benchmark and the full-size tests.

Each entry of WORKLOADS names one configuration of the solve phase: the
synthetic operator (seeded generators of libcmhost.so), the preconditioner
built on the host, the right-hand side b = A x* and the reference loop's
tolerance and NMAX.  build() makes the system; serve_systems() /
serve_requests() make the serving mix (resident systems and the Poisson
request stream).
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import Csr, Factors, acifim, afiim, csr_mv_host, gen_convdiff, gen_stencil27, gen_uniform, pdfiim

WORKLOADS = {
    "cd3d-128": dict(
        desc="3D 7-point convection-diffusion 128^3 (n=2097152, nnz=14581760), log-normal diffusion, "
             "velocity (1,1,1), cell Peclet 1; PDFIIM(prefill=2, lfill=2, postfill=1, eps=0.05); "
             "b=A x*, x*~U[-1,1], x0=0; plain FP64 BiCGSTAB to rel-res 1e-8",
        dim=3, N=128, pe=1.0, seed=2, xseed=3, prec="pdfiim", args=(2, 2, 1, 0.05), tol=1e-8, nmax=1200),
    "cd3d-128-pe5": dict(
        desc="3D 7-point convection-diffusion 128^3 (n=2097152, nnz=14581760), cell Peclet 5; "
             "ACIFIM(eps1=0.01, eps2=0.1), apply mode E; b=A x*, x0=0; plain FP64 BiCGSTAB to rel-res 1e-10",
        dim=3, N=128, pe=5.0, seed=4, xseed=5, prec="acifim", args=(0.01, 0.1), tol=1e-10, nmax=2000),
    "cd3d-128-pe5-c2": dict(
        desc="3D 7-point convection-diffusion 128^3 (n=2097152, nnz=14581760), cell Peclet 5; "
             "ACIFIM(eps1=0.01, eps2=0.1), apply mode C with nr=2 (one Richardson correction with L and "
             "one with U around the approximate inverses); b=A x*, x0=0; plain FP64 BiCGSTAB to rel-res 1e-10",
        dim=3, N=128, pe=5.0, seed=4, xseed=5, prec="acifim", args=(0.01, 0.1), tol=1e-10, nmax=2000,
        mode="C", nr=2),
    "cd3d-128-pe5-i": dict(
        desc="3D 7-point convection-diffusion 128^3 (n=2097152, nnz=14581760), cell Peclet 5; "
             "ACIFIM(eps1=0.01, eps2=0.1), apply mode I (triangular solves with L and U, D^-1 between); "
             "b=A x*, x0=0; plain FP64 BiCGSTAB to rel-res 1e-10",
        dim=3, N=128, pe=5.0, seed=4, xseed=5, prec="acifim", args=(0.01, 0.1), tol=1e-10, nmax=2000,
        mode="I"),
    "cd3d-128-pe5-j": dict(
        desc="3D 7-point convection-diffusion 128^3 (n=2097152, nnz=14581760), cell Peclet 5; "
             "ACIFIM(eps1=0.01, eps2=0.1), apply mode J (triangular solves with L and U); "
             "b=A x*, x0=0; plain FP64 BiCGSTAB to rel-res 1e-10",
        dim=3, N=128, pe=5.0, seed=4, xseed=5, prec="acifim", args=(0.01, 0.1), tol=1e-10, nmax=2000,
        mode="J"),
    "st27-100-rhs16": dict(
        desc="3D 27-point stencil 100^3 (n=1000000), off-diagonals -(1+0.3 U[-1,1]), diagonal 1.05 x row "
             "absolute sum, 5% of rows with 200-600 long-range couplings -0.01 U[0,1] (nnz~46.3M); AFIIM "
             "eps=0.05; block of 16 RHS b_j=A x*_j, x0=0; FP32 inner block BiCGSTAB (FP32 A and factors, "
             "to 1e-4) inside per-column FP64 refinement to true rel-res 1e-8",
        gen="stencil27", N=100, frac=0.05, lo=200, hi=600, seed=11, xseed=12, prec="afiim", args=(0.05,),
        tol=1e-8, nmax=500, k=16),
    "serve-mix": dict(
        desc="serving: Poisson arrivals, 2D 5-point / 3D 7-point convection-diffusion systems (50/50), n "
             "log-uniform in [1e3, 5e5], log-normal diffusion, Peclet U[0,5], AFIIM eps=0.1 resident; 1-8 RHS "
             "per request; FP32-inner / FP64-refinement to true rel-res 1e-8 with continuous batching, vs the "
             "plain FP64 one-at-a-time GPU path on the same requests",
        prec="afiim", tol=1e-8, nmax=1000),
    "cd3d-128-sym": dict(
        desc="3D 7-point diffusion 128^3 (n=2097152), log-normal diffusion, no convection (symmetric): A "
             "held as its upper triangle with mtype SYMMETRIC, PDFIIM(2,2,1,0.05) built on that stored "
             "triangle as the reference driver does, solve on the symmetric branch (full A on the device, "
             "M = G D^-1 G^T through an explicit G^T copy); b=A x*, x0=0; plain FP64 BiCGSTAB to rel-res 1e-8",
        dim=3, N=128, pe=0.0, seed=6, xseed=7, prec="pdfiim", sym=True, args=(2, 2, 1, 0.05), tol=1e-8,
        nmax=1200),
    "cd2d-100": dict(
        desc="2D 5-point convection-diffusion 100^2 (n=10000, nnz=49600), cell Peclet 0.5; AFIIM eps=0.1; "
             "b=A x*, x0=0; plain FP64 BiCGSTAB to rel-res 1e-8",
        dim=2, N=100, pe=0.5, seed=1, xseed=2, prec="afiim", args=(0.1,), tol=1e-8, nmax=500),
}

def upper_triangle(A):
    """The upper triangle (diagonal included) of a CSR matrix: the stored
    form of a symmetric matrix (mtype SYMMETRIC)."""
    rows = np.repeat(np.arange(A.n, dtype=np.int32), np.diff(A.ptr))
    keep = A.idx >= rows
    ptr = np.zeros(A.n + 1, np.int32)
    np.cumsum(np.bincount(rows[keep], minlength=A.n), out=ptr[1:])
    return Csr(A.n, ptr, np.ascontiguousarray(A.idx[keep]), np.ascontiguousarray(A.val[keep]))


def build(w):
    """(A, x*, b, factors, host build seconds) of workload w.  For a
    symmetric workload A is the stored upper triangle, b = A_full x*, and the
    factors are built on the stored triangle with H left out (the solve uses
    G^T)."""
    if w.get("gen") == "stencil27":
        A = gen_stencil27(w["N"], w["frac"], w["lo"], w["hi"], seed=w["seed"])
    else:
        A = gen_convdiff(w["dim"], w["N"], w["pe"], seed=w["seed"])
    xs = gen_uniform(w["xseed"], A.n)
    b = csr_mv_host(A, xs)
    if w.get("sym"):
        A = upper_triangle(A)
    t0 = time.perf_counter()
    if w["prec"] == "pdfiim":
        F = pdfiim(A, *w["args"], -1, "I", "B")
        if w.get("sym"):
            F = Factors(F.G, F.idv, None)
    elif w["prec"] == "afiim":
        F = afiim(A, *w["args"])
    else:
        F = acifim(A, *w["args"])
    t_build = time.perf_counter() - t0
    return A, xs, b, F, t_build


def serve_systems(nsys, seed):
    """nsys systems: kind 2D/3D 50/50, n log-uniform in [1e3, 5e5], N =
    round(n^(1/d)), Peclet U[0,5], log-normal diffusion; AFIIM eps=0.1 built
    on host threads."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    specs = []
    for s in range(nsys):
        d = 2 if rng.random() < 0.5 else 3
        n = float(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        N = max(4, int(round(n ** (1.0 / d))))
        specs.append((d, N, float(rng.uniform(0.0, 5.0)), 1000 + s))

    def one(sp):
        d, N, pe, sd = sp
        A = gen_convdiff(d, N, pe, seed=sd)
        return A, afiim(A, 0.1)
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(one, specs))



def serve_requests(systems, nreq, rate, seed):
    """The request stream over `systems`: per system a pool of 8 right-hand
    sides b = A x* (pools[s] = (X*, B)), Poisson arrival times at `rate`
    requests/s, and per request (system, columns of its pool) with 1-8
    columns, all drawn from one generator seeded with `seed`."""
    rng = np.random.default_rng(seed)
    pools = []
    for s, (A, _) in enumerate(systems):
        XS = np.stack([gen_uniform(50000 + 8 * s + j, A.n) for j in range(8)], axis=1)
        pools.append((XS, np.stack([csr_mv_host(A, XS[:, j]) for j in range(8)], axis=1)))
    arrive = np.cumsum(rng.exponential(1.0 / rate, nreq))
    reqs = []
    for _ in range(nreq):
        s = int(rng.integers(len(systems)))
        k = int(rng.integers(1, 9))
        reqs.append((s, np.sort(rng.choice(8, size=k, replace=False))))
    return pools, arrive, reqs
