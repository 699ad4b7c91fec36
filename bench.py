#!/usr/bin/env python3
"""Benchmark of the MI355X preconditioned-BiCGSTAB solve engine.

Default workload (one GPU): 3D 7-point convection-diffusion on a 128^3 grid
(n = 2,097,152, nnz = 14,581,760), log-normal diffusion, velocity (1,1,1) at
cell Peclet 1, PDFIIM factors (prefill 2, lfill 2, postfill 1, eps 0.05,
inf-norm filter, both parts) built on the host, b = A x* with x* ~ U[-1,1],
x0 = 0, plain FP64 BiCGSTAB to relative residual 1e-8 (NMAX 1200).

A step is one complete solve (setup + every BiCGSTAB pass to the exit) with A,
the factors and b already resident in HBM.  value = solves/s over the K timed
steps; latency percentiles are per solve.  The dominant kernel's roofline is
measured live: every launch in the timed solves is bracketed by HIP events on
the engine's own stream, and achieved = algorithmic bytes / measured time.

Prints ONE JSON line.  --gpus N with N > 1 is refused (one-GPU build).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "computational-mathematics_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "PBiCGSTAB solves/s + p99 latency at rel-res 1e-8; SpMV/M-apply HBM GB/s vs peak"

WORKLOADS = {
    "c2": dict(
        desc="3D 7-point convection-diffusion 128^3 (n=2097152, nnz=14581760), log-normal diffusion, "
             "velocity (1,1,1), cell Peclet 1; PDFIIM(prefill=2, lfill=2, postfill=1, eps=0.05); "
             "b=A x*, x*~U[-1,1], x0=0; plain FP64 BiCGSTAB to rel-res 1e-8",
        dim=3, N=128, pe=1.0, seed=2, xseed=3, prec="pdfiim", args=(2, 2, 1, 0.05), tol=1e-8, nmax=1200),
    "c3": dict(
        desc="3D 7-point convection-diffusion 128^3 (n=2097152, nnz=14581760), cell Peclet 5; "
             "ACIFIM(eps1=0.01, eps2=0.1), apply mode E; b=A x*, x0=0; plain FP64 BiCGSTAB to rel-res 1e-10",
        dim=3, N=128, pe=5.0, seed=4, xseed=5, prec="acifim", args=(0.01, 0.1), tol=1e-10, nmax=2000),
    "c1": dict(
        desc="2D 5-point convection-diffusion 100^2 (n=10000, nnz=49600), cell Peclet 0.5; AFIIM eps=0.1; "
             "b=A x*, x0=0; plain FP64 BiCGSTAB to rel-res 1e-8",
        dim=2, N=100, pe=0.5, seed=1, xseed=2, prec="afiim", args=(0.1,), tol=1e-8, nmax=500),
}

CAT_NAMES = {"A": "A SpMV + fused dot(s)", "F1": "factor SpMV H (D^-1 fused)", "F2": "factor SpMV G",
             "LU": "L/U SpMV (mode C)", "VEC": "BiCGSTAB vector updates", "SETUP": "setup r=b-Ax",
             "FIN": "one-workgroup reduction finalize"}
CAT_KERNELS = {"A": "spmv_kernel<double, double, cm::EpiAlpha|EpiOmega>",
               "F1": "spmv_kernel<double, double, cm::EpiScale<double, double>>",
               "F2": "spmv_kernel<double, double, cm::EpiStore<double>>"}


def build_workload(w):
    import cmamd
    A = cmamd.gen_convdiff(w["dim"], w["N"], w["pe"], seed=w["seed"])
    xs = cmamd.gen_uniform(w["xseed"], A.n)
    b = cmamd.csr_mv_host(A, xs)
    t0 = time.perf_counter()
    if w["prec"] == "pdfiim":
        F = cmamd.pdfiim(A, *w["args"], -1, "I", "B")
    elif w["prec"] == "afiim":
        F = cmamd.afiim(A, *w["args"])
    else:
        F = cmamd.acifim(A, *w["args"])
    t_build = time.perf_counter() - t0
    return A, xs, b, F, t_build


def cpu_baseline(w, A, b, F, seconds):
    """The oracle's FP64 loop (C restatement of the reference solve) on the
    host cores: full solves until `seconds` have elapsed (at least one)."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, 64))
    oracle.set_threads(threads)
    times, its = [], []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        _, _, it, info, _ = oracle.bicgstab(w["prec"], A, b, F, tol=w["tol"], nmax=w["nmax"])
        times.append(time.perf_counter() - t0)
        its.append(it)
        if time.perf_counter() - t_start >= seconds or len(times) >= 50:
            break
    med = float(np.median(times))
    return {"value": 1.0 / med, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} full solve(s) of this workload ({its[0]} iterations each) by the oracle "
                      f"(C restatement of the reference loop, OpenMP, {threads} threads); median "
                      f"{med * 1e3:.1f} ms/solve"}


def mixed_leg(sys_, bits, w, A, b, x64, args):
    """Mixed precision on the same resident system: FP64 refinement around an
    inner BiCGSTAB on FP32 A, FP32 or FP16 (per-row scaled) factors and FP32
    vectors, to the same true relative residual as the FP64 path."""
    import cmamd
    sys_.enable_mixed(bits)
    for _ in range(max(1, args.warmup)):
        sys_.set_x(None)
        sys_.solve_mixed(bits, tol=w["tol"], inner_tol=1e-4, nmax=w["nmax"])
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        sys_.set_x(None)
        rv, outer, inner, info = sys_.solve_mixed(bits, tol=w["tol"], inner_tol=1e-4, nmax=w["nmax"])
        lat.append(time.perf_counter() - ts)
    elapsed = time.perf_counter() - t0
    x = sys_.get_x()
    true_res = float(np.linalg.norm(b - cmamd.csr_mv_host(A, x)) / np.linalg.norm(b))
    return {"value": round(args.steps / elapsed, 3), "unit": "solves/s",
            "ms_per_step": round(1e3 * elapsed / args.steps, 3),
            "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                           "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
            "outer_steps": outer, "inner_passes": inner, "info": info, "true_rel_residual": true_res,
            "rel_error_vs_fp64_solution": float(np.linalg.norm(x - x64) / np.linalg.norm(x64)),
            "inner": f"BiCGSTAB on FP32 A, FP{bits} factors{' (per-row scaled)' if bits == 16 else ''}, "
                     f"FP32 vectors, to 1e-4; FP64 refinement of b - A x"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision legs")
    ap.add_argument("--kernel-timing", choices=("timed-region", "separate"), default="timed-region",
                    help="bracket the dominant kernel's launches with HIP events inside the timed region "
                         "(default), or time the region without events and time that kernel in extra solves")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy of 1 GiB (x10), the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE hardware counters in rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8,
                          "launches": 10}))
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import cmamd
    if cmamd.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    w = WORKLOADS[args.workload]
    A, xs, b, F, t_build = build_workload(w)
    print(f"[bench] workload {args.workload}: n={A.n} nnz(A)={A.nnz} nnz(G)={F.G.nnz} "
          f"nnz(H)={F.H.nnz} factors built in {t_build:.1f}s", file=sys.stderr, flush=True)

    sys_ = cmamd.System(w["prec"], A, F)
    sys_.set_rhs(b)
    for _ in range(args.warmup):
        sys_.set_x(None)
        sys_.solve(tol=w["tol"], nmax=w["nmax"])
    # per-kernel breakdown: two extra solves with every launch bracketed
    sys_.profile(True)
    for _ in range(2):
        sys_.set_x(None)
        sys_.solve(tol=w["tol"], nmax=w["nmax"])
    stats = sys_.kernel_stats()
    dom = max((c for c in ("A", "F1", "F2", "LU") if stats[c][0] > 0), key=lambda c: stats[c][1])
    # timed region: only the dominant kernel's launches carry events
    sys_.profile(args.kernel_timing == "timed-region", cats=(dom,))
    lat = []
    iters = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        sys_.set_x(None)
        rv, it, info, _ = sys_.solve(tol=w["tol"], nmax=w["nmax"])
        lat.append(time.perf_counter() - ts)
        iters.append(it)
    elapsed = time.perf_counter() - t0
    if args.kernel_timing == "separate":
        sys_.profile(True, cats=(dom,))
        for _ in range(3):
            sys_.set_x(None)
            sys_.solve(tol=w["tol"], nmax=w["nmax"])
    dstats = sys_.kernel_stats()[dom]
    sys_.profile(False)
    x = sys_.get_x()
    true_res = float(np.linalg.norm(b - cmamd.csr_mv_host(A, x)) / np.linalg.norm(b))
    fwd_err = float(np.linalg.norm(x - xs) / np.linalg.norm(xs))
    mixed = {}
    if not args.no_mixed:
        for bits in (32, 16):
            mixed[f"fp{bits}_factors"] = mixed_leg(sys_, bits, w, A, b, x, args)
    sys_.close()

    value = args.steps / elapsed
    kernels = {}
    for cat, (n_l, ms, by) in stats.items():
        if n_l == 0:
            continue
        kernels[cat] = {"what": CAT_NAMES[cat], "launches": n_l, "avg_us": 1e3 * ms / n_l,
                        "bytes_per_launch": by / n_l, "achieved_GBs": by / ms / 1e6,
                        "share_of_gpu_time": None}
    tot_ms = sum(stats[c][1] for c in kernels)
    for c in kernels:
        kernels[c]["share_of_gpu_time"] = stats[c][1] / tot_ms
    n_l, ms, by = dstats
    d_avg_us = 1e3 * ms / n_l
    d_gbs = by / ms / 1e6
    roofline = {"bound": "hbm", "kernel": CAT_KERNELS.get(dom, dom), "achieved": round(d_gbs, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(d_gbs / HBM_PEAK_GBS, 4),
                "traffic": None, "bytes_per_launch": by / n_l, "avg_us": round(d_avg_us, 2),
                "launches_timed": n_l}
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        if dom in tr:
            roofline["traffic"] = tr[dom]

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["desc"], "n": A.n, "nnz_A": A.nnz, "nnz_G": F.G.nnz, "nnz_H": F.H.nnz,
                   "tol": w["tol"], "nmax": w["nmax"]},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "iterations": int(np.median(iters)), "info": info, "resval": rv, "true_rel_residual": true_res,
        "forward_error_vs_xstar": fwd_err,
        "mixed_precision": mixed,
        "roofline": roofline, "kernels": kernels,
        "kernels_note": "per-kernel table from 2 extra solves with every launch bracketed by HIP events; "
                        "roofline from the timed region, where only the dominant kernel is bracketed",
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, A, b, F, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
