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
SAMPLE_PASSES = 8  # passes per timed solve whose dominant-kernel launches are bracketed by HIP events
DROPIN_CALLS = 3  # one-shot drop-in calls timed per single-system workload
DOM_LAUNCHES_PER_PASS_MAX = 16  # above this, the dominant kernel is timed outside the timed region
METRIC = "PBiCGSTAB solves/s + p99 latency at rel-res 1e-8; SpMV/M-apply HBM GB/s vs peak"

from cmamd.workloads import WORKLOADS, build as build_workload, serve_requests, serve_systems  # noqa: E402

CAT_NAMES = {"A": "A SpMV + fused dot(s)", "F1": "factor SpMV H (D^-1 fused)", "F2": "factor SpMV G",
             "LU": "L/U SpMV (mode C) or triangular solve (modes I/J)", "VEC": "BiCGSTAB vector updates", "SETUP": "setup r=b-Ax",
             "FIN": "one-workgroup reduction finalize"}
CAT_KERNELS = {"A": "cm::spmv_kernel<2048, double, double, double, cm::EpiAlpha|EpiOmega<double> >",
               "F1": "cm::spmv_kernel<2048, double, double, double, cm::EpiScale<double, double> >",
               "F2": "cm::spmv_kernel<2048, double, double, double, cm::EpiStore<double> >"}


# Kernel name fragment -> category of the FP64 solve's per-kernel table.
PMC_CATEGORIES = (("EpiScale<double", "F1"), ("EpiStore<double", "F2"), ("EpiAlpha<double", "A"),
                  ("EpiOmega<double", "A"), ("k_s_update<double", "VEC"), ("k_xr_update<double", "VEC"),
                  ("k_p_update<double", "VEC"))


def pmc_summary(fetch_db, write_db, copy_fetch_db, copy_write_db, bench_json, workload):
    """HBM bytes per dispatch from two rocprofv3 --pmc runs of the bench (FETCH_SIZE, WRITE_SIZE, KiB
    per dispatch) scaled by what the same counters read for a plain 1 GiB device copy, per kernel of
    the FP64 solve; next to the bytes the storage format in use touches once (compulsory) and the
    CSR algorithmic bytes, both from the bench's own table.  Writes
    profiles/traffic_<workload>.json (read by the bench for roofline.traffic) and prints the table."""
    import sqlite3

    def per_dispatch(db, solves_only=False):
        con = sqlite3.connect(db)
        rows = con.execute("select dispatch_id, kernel_name, value from counters_collection").fetchall()
        con.close()
        if solves_only:
            # the autotuner's timing launches come before the first solve: keep the dispatches from
            # the first BiCGSTAB vector update on
            first = min((d for d, k, _ in rows if "k_s_update" in k), default=0)
            rows = [r for r in rows if r[0] >= first]
        return rows

    gib = float(2 ** 30)
    scale = {}
    for name, db in (("fetch", copy_fetch_db), ("write", copy_write_db)):
        v = [val for _, k, val in per_dispatch(db) if "copyBuffer" in k and val * 1024.0 > 0.25 * gib]
        if not v:
            raise SystemExit(f"no 1 GiB copy dispatch in {db}")
        scale[name] = gib / (1024.0 * float(np.median(v)))
    with open(bench_json) as f:
        bj = json.loads(f.read().strip().splitlines()[-1])
    kern = bj.get("kernels", {})
    algo = {c: kern[c]["bytes_per_launch"] for c in ("A", "F1", "F2") if c in kern}
    fmtb = {c: kern[c].get("format_bytes_per_launch", kern[c]["bytes_per_launch"]) for c in algo}
    out, rows = {}, []
    sums = {}
    for name, db in (("fetch", fetch_db), ("write", write_db)):
        by_kernel = {}
        for _, k, val in per_dispatch(db, solves_only=True):
            for frag, cat in PMC_CATEGORIES:
                if frag in k:
                    by_kernel.setdefault((frag, cat), []).append(1024.0 * val * scale[name])
        for key, v in by_kernel.items():
            v = np.asarray(v)
            v = v[v > 0.5 * v.max()]  # executed dispatches: launches after a solve's exit read ~0
            sums.setdefault(key, {})[name] = (float(v.mean()), len(v))
    for (frag, cat), d in sums.items():
        if "fetch" not in d or "write" not in d:
            continue
        t = d["fetch"][0] + d["write"][0]
        rows.append((frag.split("<")[0], cat, d["fetch"][1], d["fetch"][0], d["write"][0], t, fmtb.get(cat),
                     algo.get(cat)))
        if cat in algo:
            out.setdefault(cat, []).append(t)
    tr = {c: float(np.mean(v)) for c, v in out.items()}
    tr["_note"] = (f"HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE (x{scale['fetch']:.3f}) + WRITE_SIZE "
                   f"(x{scale['write']:.3f}), scaled by a 1 GiB device copy")
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    with open(path, "w") as f:
        json.dump(tr, f, indent=1)
    print(f"scale factors: fetch x{scale['fetch']:.3f}, write x{scale['write']:.3f}")
    print("| kernel | category | dispatches | fetch MB | write MB | traffic MB | compulsory MB | traffic / compulsory "
          "| CSR algorithmic MB |")
    print("|---|---|---|---|---|---|---|---|---|")
    for k, cat, nd, fe, wr, t, fb, a in rows:
        ratio = f"{t / fb:.3f}" if fb else "-"
        fm = f"{fb / 1e6:.1f}" if fb else "-"
        am = f"{a / 1e6:.1f}" if a else "-"
        print(f"| {k} | {cat} | {nd} | {fe / 1e6:.1f} | {wr / 1e6:.1f} | {t / 1e6:.1f} | {fm} | {ratio} | {am} |")
    print(f"wrote {path}")


def pmc_block_summary(fetch_db, write_db, copy_fetch_db, copy_write_db, workload):
    """HBM bytes per dispatch of the 16-column FP32 block kernels from two rocprofv3 --pmc runs of the
    block bench (FETCH_SIZE, WRITE_SIZE), scaled by a plain 1 GiB device copy as in pmc_summary.  The A
    product the block roofline times is the diagonal kernel (spmm_dia, BEpiStore) with the long-row
    dispatches right before it: the segment sums (spmm_short, BEpiStore) and their tile reduction
    (k_long_tiles), or one wave-per-row pass (spmm_long) when the remainders are not tiled.  Only the
    bench's SpMM timing of A launches that sequence.  Writes profiles/traffic_<workload>.json ({"A":
    bytes per A product}) and prints the table."""
    import re
    import sqlite3

    def rows_of(db):
        con = sqlite3.connect(db)
        rows = con.execute("select dispatch_id, kernel_name, value from counters_collection "
                           "order by dispatch_id").fetchall()
        con.close()
        return rows

    gib = float(2 ** 30)
    scale = {}
    for name, db in (("fetch", copy_fetch_db), ("write", copy_write_db)):
        v = [val for _, k, val in rows_of(db) if "copyBuffer" in k and val * 1024.0 > 0.25 * gib]
        if not v:
            raise SystemExit(f"no 1 GiB copy dispatch in {db}")
        scale[name] = gib / (1024.0 * float(np.median(v)))
    dia_k = "spmm_dia<float, float, float, 16, cm::BEpiStore<float>"  # either remainder form
    long_k = "spmm_long<float, float, float, 16, cm::BEpiStore<float> >"
    seg_k, tiles_k = "spmm_short<float, float, float, 16, cm::BEpiStore<float> >", "k_long_tiles<float, float>"
    per = {}
    pairs = {}
    for name, db in (("fetch", fetch_db), ("write", write_db)):
        rows = rows_of(db)
        for _, k, val in rows:
            if ("16" in k and ("spmm_" in k or "k_block_" in k) and "float" in k) or tiles_k in k:
                per.setdefault(re.sub(r"\(.*$", "", k), {}).setdefault(name, []).append(1024.0 * val * scale[name])
        a_long, a_dia = [], []
        for i in range(2, len(rows)):
            if dia_k not in rows[i][1]:
                continue
            if tiles_k in rows[i - 1][1] and seg_k in rows[i - 2][1]:
                lv = rows[i - 1][2] + rows[i - 2][2]
            elif long_k in rows[i - 1][1]:
                lv = rows[i - 1][2]
            else:
                continue
            a_long.append(1024.0 * lv * scale[name])
            a_dia.append(1024.0 * rows[i][2] * scale[name])
        pairs[name] = (a_long, a_dia)
    if not pairs["fetch"][0] or not pairs["write"][0]:
        raise SystemExit("no A product (long-row dispatches, then spmm_dia, BEpiStore) in the databases")
    lf, df = (float(np.mean(v)) for v in pairs["fetch"])
    lw, dw = (float(np.mean(v)) for v in pairs["write"])
    tr = {"A": lf + df + lw + dw,
          "_note": f"HBM bytes per 16-column FP32 A product (long-row dispatches + spmm_dia) from rocprofv3 --pmc FETCH_SIZE "
                   f"(x{scale['fetch']:.3f}) + WRITE_SIZE (x{scale['write']:.3f}), scaled by a 1 GiB device copy; "
                   f"{len(pairs['fetch'][0])} products"}
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    with open(path, "w") as f:
        json.dump(tr, f, indent=1)
    print(f"scale factors: fetch x{scale['fetch']:.3f}, write x{scale['write']:.3f}")
    print("| kernel | dispatches | fetch MB | write MB | traffic MB |")
    print("|---|---|---|---|---|")
    for k, d in sorted(per.items(), key=lambda kv: -np.mean(kv[1].get("fetch", [0]))):
        if "fetch" not in d or "write" not in d:
            continue
        fe, wr = float(np.mean(d["fetch"])), float(np.mean(d["write"]))
        print(f"| `{k}` | {len(d['fetch'])} | {fe / 1e6:.1f} | {wr / 1e6:.1f} | {(fe + wr) / 1e6:.1f} |")
    print(f"A product ({len(pairs['fetch'][0])} timed): long-row pass {(lf + lw) / 1e6:.1f} MB "
          f"(fetch {lf / 1e6:.1f}), diagonal kernel {(df + dw) / 1e6:.1f} MB (fetch {df / 1e6:.1f}), "
          f"total {tr['A'] / 1e6:.1f} MB")
    print(f"wrote {path}")


def trace_summary(db, thr_us=5.0, top=25):
    """Per-kernel durations from a rocprofv3 --kernel-trace result database, as a markdown table:
    launches, share of GPU time, mean duration, and the launches longer than thr_us (no-op launches
    after a solve's exit read the device state and return at once)."""
    import re
    import sqlite3
    con = sqlite3.connect(db)
    cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
    name_col = "kernel_name" if "kernel_name" in cols else "name"
    rows = con.execute(f"select {name_col}, start, end from kernels").fetchall()
    con.close()
    agg = {}
    for n, t0, t1 in rows:
        agg.setdefault(n, []).append((t1 - t0) / 1e3)
    tot = sum(sum(v) for v in agg.values())
    out = []
    for n, v in agg.items():
        ex = [d for d in v if d > thr_us]
        out.append((sum(v), re.sub(r"\(.*$", "", n), len(v), sum(v) / tot, sum(v) / len(v), len(ex),
                    sum(ex) / len(ex) if ex else 0.0))
    out.sort(reverse=True)
    print("| kernel | launches | share of GPU time | avg all (us) | executed | avg executed (us) |")
    print("|---|---|---|---|---|---|")
    for t, n, c, sh, a, ne, ae in out[:top]:
        print(f"| `{n}` | {c} | {100 * sh:.1f}% | {a:.1f} | {ne} | {ae:.1f} |")


def cpu_baseline(w, A, b, F, seconds):
    """The oracle's FP64 loop (C restatement of the reference solve) on the
    host cores: full solves until `seconds` have elapsed (at least one)."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, 64))
    oracle.set_threads(threads)
    times, its = [], []
    x_or = None
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        x, _, it, info, _ = oracle.bicgstab(w["prec"], A, b, F, tol=w["tol"], nmax=w["nmax"],
                                            nr=w.get("nr", 1), mode=w.get("mode", "E"),
                                            issym=w.get("sym", False))
        times.append(time.perf_counter() - t0)
        its.append(it)
        if x_or is None:
            x_or = x
        if time.perf_counter() - t_start >= seconds or len(times) >= 50:
            break
    med = float(np.median(times))
    out = {"value": 1.0 / med, "unit": "solves/s", "cores": threads, "kind": "port",
           "sample": f"{len(times)} full solve(s) of this workload ({its[0]} iterations each) by the oracle "
                     f"(C restatement of the reference loop, OpenMP, {threads} threads); median "
                     f"{med * 1e3:.1f} ms/solve",
           "iterations": its[0]}
    if w["prec"] == "afiim":
        # AFIIM's own build links sequential MKL (AFIIM/Makefile:12): the same loop on one thread,
        # bounded to a few seconds
        oracle.set_threads(1)
        t1 = []
        t_start = time.perf_counter()
        while True:
            t0 = time.perf_counter()
            oracle.bicgstab(w["prec"], A, b, F, tol=w["tol"], nmax=w["nmax"])
            t1.append(time.perf_counter() - t0)
            if time.perf_counter() - t_start >= min(seconds, 4.0) or len(t1) >= 50:
                break
        oracle.set_threads(threads)
        m1 = float(np.median(t1))
        out["single_thread"] = {"value": 1.0 / m1, "unit": "solves/s", "cores": 1,
                                "sample": f"{len(t1)} full solve(s) on one thread; median {m1 * 1e3:.1f} ms/solve"}
    return out, x_or


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def mixed_leg(sys_, bits, w, A, b, x64, args):
    """Mixed precision on the same resident system: FP64 refinement around an
    inner BiCGSTAB on FP32 A, FP32 or FP16 (per-row scaled) factors and FP32
    vectors, to the same true relative residual as the FP64 path."""
    import cmamd
    sys_.enable_mixed(bits)
    itol = args.inner_tol
    for _ in range(max(1, args.warmup)):
        sys_.set_x(None)
        sys_.solve_mixed(bits, tol=w["tol"], inner_tol=itol, nmax=w["nmax"])
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        sys_.set_x(None)
        rv, outer, inner, info = sys_.solve_mixed(bits, tol=w["tol"], inner_tol=itol, nmax=w["nmax"])
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
                     f"FP32 vectors, to {itol:g}; FP64 refinement of b - A x"}


def _pct(v, q):
    return round(float(np.percentile(v, q)), 3)


def bench_block(args, w):
    """Block workload: one step = one block solve of k right-hand sides (FP32 inner /
    FP64 refinement per column).  Baselines on the same inputs: the plain
    FP64 GPU path one column at a time, and the FP64 block solve."""
    import cmamd
    A, _, _, F, t_build = build_workload(w)
    k = w["k"]
    XS = np.stack([cmamd.gen_uniform(w["xseed"] + j, A.n) for j in range(k)], axis=1)
    B = np.stack([cmamd.csr_mv_host(A, XS[:, j]) for j in range(k)], axis=1)
    print(f"[bench] workload st27-100-rhs16: n={A.n} nnz(A)={A.nnz} nnz(G)={F.G.nnz} nnz(H)={F.H.nnz} "
          f"factors built in {t_build:.1f}s", file=sys.stderr, flush=True)
    sys_ = cmamd.System(w["prec"], A, F)
    sys_.enable_mixed(32)
    tuned = None
    if not args.no_autotune:
        # single-RHS SpMV formats (the FP64 one-column-at-a-time baseline runs on them)
        tuned = sys_.autotune(reps=10)
    if args.no_graphs:
        sys_.block_graphs(False)
        sys_.graphs(False)
    steps = max(1, args.steps)

    def timed(fn, reps):
        lat = []
        for _ in range(reps):
            ts = time.perf_counter()
            out = fn()
            lat.append(time.perf_counter() - ts)
        return out, lat

    # mixed block (the measured path): B resident in HBM before the timed region; a step resets
    # x to zero and solves the k columns
    sys_.block_set(B)

    def mixed():
        sys_.block_set(None)
        return sys_.block_solve_mixed(32, tol=w["tol"], inner_tol=1e-4, nmax=w["nmax"])
    for _ in range(args.warmup):
        mixed()
    t0 = time.perf_counter()
    rmx, lat = timed(mixed, steps)
    elapsed = time.perf_counter() - t0
    value = k * steps / elapsed
    Xmx = sys_.block_get()
    # PCIe-inclusive step (not the value): upload B, solve, copy X back
    t0 = time.perf_counter()
    sys_.block_set(B)
    sys_.block_solve_mixed(32, tol=w["tol"], inner_tol=1e-4, nmax=w["nmax"])
    sys_.block_get()
    t_pcie = time.perf_counter() - t0
    # FP64 block and FP64 one column at a time (plain path), right-hand sides resident likewise
    def fp64_block():
        sys_.block_set(None)
        return sys_.block_solve(tol=w["tol"], nmax=w["nmax"])
    fp64_block()
    t0 = time.perf_counter()
    r64b, _ = timed(fp64_block, max(1, steps // 2))
    v64b = k * max(1, steps // 2) / (time.perf_counter() - t0)
    cols = min(k, 4)
    sys_.set_rhs(B[:, 0])
    sys_.set_x(None)
    sys_.solve(tol=w["tol"], nmax=w["nmax"])
    t_solve = 0.0
    X64 = np.empty((A.n, cols))
    its = []
    for j in range(cols):
        sys_.set_rhs(B[:, j])  # upload outside the timed part
        sys_.set_x(None)
        t0 = time.perf_counter()
        _, it, info, _ = sys_.solve(tol=w["tol"], nmax=w["nmax"])
        t_solve += time.perf_counter() - t0
        X64[:, j] = sys_.get_x()
        its.append(it)
    v64 = cols / t_solve
    # the dominant kernel: FP32 SpMM with A over the 16 columns
    sys_.block_set(B)
    us, by = sys_.block_time_spmm("A", 32, reps=20)
    us_g, by_g = sys_.block_time_spmm("G", 32, reps=20)
    us_h, by_h = sys_.block_time_spmm("H", 32, reps=20)
    formats = sys_.spmv_formats() if tuned else {}
    sys_.close()
    res = [float(np.linalg.norm(B[:, j] - cmamd.csr_mv_host(A, Xmx[:, j])) / np.linalg.norm(B[:, j]))
           for j in range(k)]
    err64 = [float(np.linalg.norm(Xmx[:, j] - X64[:, j]) / np.linalg.norm(X64[:, j])) for j in range(cols)]
    errx = [float(np.linalg.norm(Xmx[:, j] - XS[:, j]) / np.linalg.norm(XS[:, j])) for j in range(k)]
    # CSR algorithmic bytes of one FP32 block SpMM over k columns: the matrix
    # once (4-byte values and column indices, row pointers), X read, Y written
    def alg_spmm(nnz):
        return nnz * 8.0 + 4.0 * (A.n + 1) + 2.0 * k * 4.0 * A.n
    alg = alg_spmm(A.nnz)
    gbs = alg / us / 1e3
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 refinement", "data": "synthetic",
        "config": {"workload": w["desc"], "n": A.n, "nnz_A": A.nnz, "nnz_G": F.G.nnz, "nnz_H": F.H.nnz,
                   "rhs_per_step": k, "tol": w["tol"]},
        "latency_ms": {"p50": _pct(np.array(lat) * 1e3, 50), "p99": _pct(np.array(lat) * 1e3, 99)},
        "outer_steps": rmx[1], "inner_passes": rmx[2], "info": [int(i) for i in rmx[3]],
        "pcie_inclusive": {"ms_per_step": round(1e3 * t_pcie, 3),
                           "what": "upload B, block solve, copy X back (not the bench value)"},
        "true_rel_residual_max": max(res), "rel_error_vs_fp64_solution_max": max(err64),
        "forward_error_vs_xstar_max": max(errx),
        "fp64_one_at_a_time": {"value": round(v64, 3), "unit": "solves/s", "columns_timed": cols,
                               "iterations": its},
        "fp64_block": {"value": round(v64b, 3), "unit": "solves/s", "iterations": [int(i) for i in r64b[1]]},
        "speedup_vs_fp64_one_at_a_time": round(value / v64, 3),
        "roofline": {"bound": "hbm", "kernel": "A block SpMM over 16 FP32 columns: spmm_dia<float, float, float, 16> "
                                               "after the long rows' remainders as column-tiled segments "
                                               "(spmm_short, then k_long_tiles); row-binned spmm_short/spmm_long "
                                               "when the diagonal block format does not apply",
                     "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "bytes_per_launch": alg,
                     "avg_us": round(us, 2), "timing": "20 back-to-back SpMMs between HIP events, separate "
                                                     "from the timed solves",
                     "bytes_definition": "CSR algorithmic bytes nnz*8 + 4*(n+1) + 2*k*4*n, independent of "
                                         "the storage format",
                     "format_bytes_per_launch": by, "format_frac": round(by / us / 1e3 / HBM_PEAK_GBS, 4)},
        "factor_spmm": {"G": {"avg_us": round(us_g, 2), "GBs": round(alg_spmm(F.G.nnz) / us_g / 1e3, 1),
                              "format_GBs": round(by_g / us_g / 1e3, 1)},
                        "H": {"avg_us": round(us_h, 2), "GBs": round(alg_spmm(F.H.nnz) / us_h / 1e3, 1),
                              "format_GBs": round(by_h / us_h / 1e3, 1)}},
        "single_rhs_spmv_formats": {str(b_): f for b_, f in formats.items()} if tuned else None,
    }
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            out["roofline"]["traffic"] = json.load(f).get("A")
    if not args.no_cpu:
        out["cpu_baseline"], x_or = cpu_baseline(w, A, B[:, 0], F, args.cpu_seconds)
        out["rel_error_vs_oracle_column0"] = {"mixed_block": _rel(Xmx[:, 0], x_or), "fp64_single": _rel(X64[:, 0], x_or)}
    else:
        out["cpu_baseline"] = None
    return out


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_serve(args, w):
    """Serving workload: Poisson arrivals of solve requests over resident systems.  The
    served path (continuous batching, FP32 inner / FP64 refinement) and the
    plain FP64 one-at-a-time path replay the same request list on the same
    arrival clock; latency = arrival to solution on the host."""
    import threading
    import cmamd
    t0 = time.perf_counter()
    systems = serve_systems(args.serve_systems, args.serve_seed)
    t_build = time.perf_counter() - t0
    # per system a pool of 8 right-hand sides b = A x*; Poisson arrivals
    pools, arrive, reqs = serve_requests(systems, args.serve_requests, args.serve_rate, args.serve_seed)
    n_tot = sum(A.n for A, _ in systems)
    print(f"[bench] workload serve-mix: {len(systems)} systems (total n={n_tot}), built in {t_build:.1f}s; "
          f"{len(reqs)} requests at {args.serve_rate}/s", file=sys.stderr, flush=True)
    handles = [cmamd.System("afiim", A, F) for A, F in systems]
    if args.no_graphs:
        for h in handles:
            h.block_graphs(False)
            h.graphs(False)

    def check(i, X):
        """(worst true rel-res, worst error vs x*) of request i's columns."""
        s, cols = reqs[i]
        A = systems[s][0]
        wr, wx = 0.0, 0.0
        for j, c in enumerate(cols):
            b, xs = pools[s][1][:, c], pools[s][0][:, c]
            wr = max(wr, float(np.linalg.norm(b - cmamd.csr_mv_host(A, X[:, j])) / np.linalg.norm(b)))
            wx = max(wx, float(np.linalg.norm(X[:, j] - xs) / np.linalg.norm(xs)))
        return wr, wx

    SAMPLE = 10  # every 10th served solution is kept for the comparison with the FP64 path

    def replay(submit, wait):
        results = [None] * len(reqs)
        ids = [None] * len(reqs)

        def waiter():
            for i in range(len(reqs)):
                while ids[i] is None:
                    time.sleep(0.0005)
                results[i] = wait(ids[i])  # checked after the replay, off the serving clock
        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        for i, (s, cols) in enumerate(reqs):
            dt = start + arrive[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids[i] = submit(s, pools[s][1][:, cols])
        th.join()
        return results, time.perf_counter() - start

    # served: continuous batching, mixed precision
    srv = cmamd.Server(slots=args.serve_slots, precision=32, small_n=args.serve_small_n)
    sid = [srv.add_system(h) for h in handles]
    _progress(f"{len(handles)} systems uploaded and registered")
    t0 = time.perf_counter()
    for kw in (1, 2, 4, 8, 16):  # create every system's block workspaces once (not timed)
        ids = [srv.submit(sid[s], np.tile(pools[s][1], (1, 2))[:, :kw], tol=w["tol"], nmax=w["nmax"])
               for s in range(len(systems))]
        for r in ids:
            srv.wait(r)
    t_warm = time.perf_counter() - t0
    st0 = srv.stats()
    _progress(f"block workspaces of {len(systems)} systems created in {t_warm:.1f}s; replaying {len(reqs)} requests")
    served_raw, t_served = replay(lambda s, B: srv.submit(sid[s], B, tol=w["tol"], nmax=w["nmax"]), srv.wait)
    st = srv.stats()
    _progress(f"served in {t_served:.1f}s; checking every column against b on the host")
    served = []
    info_counts, failed = {}, []
    for i, r in enumerate(served_raw):
        if i and i % 2000 == 0:
            _progress(f"checked {i} requests")
        wr, wx = check(i, r.X)
        served.append((r.latency_ms, r.queue_ms, r.batch_cols, int(np.count_nonzero(r.info)), wr, wx,
                       r.X if i % SAMPLE == 0 else None, r.service_ms))
        for code in r.info:
            info_counts[int(code)] = info_counts.get(int(code), 0) + 1
        if np.any(r.info != 0) and len(failed) < 20:
            s, cols = reqs[i]
            failed.append({"request": i, "n": int(systems[s][0].n), "k": len(cols), "batch_cols": r.batch_cols,
                           "info": [int(c) for c in r.info], "resval": [float(v) for v in r.resval],
                           "outer": r.outer, "inner": r.inner, "true_rel_residual": wr})
    del served_raw
    _progress("checked; autotuning the single-RHS formats for the FP64 baseline")
    for key in ("batches", "columns", "persistent_batches", "persistent_columns", "evictions"):
        st[key] -= st0[key]
    srv.close()
    # plain FP64 one-at-a-time: each column a single solve on its resident
    # system, in arrival order, one at a time, on autotuned single-RHS SpMV formats
    t_tune = time.perf_counter()
    if not args.no_autotune:
        for h in handles:
            h.autotune(reps=3)
    t_tune = time.perf_counter() - t_tune
    _progress(f"autotuned in {t_tune:.1f}s; FP64 one-at-a-time baseline")
    lock = threading.Lock()
    queue = []
    cv = threading.Condition(lock)

    def fp64_worker(out):  # noqa: C901
        while True:
            with cv:
                while not queue:
                    cv.wait()
                item = queue.pop(0)
            if item is None:
                return
            i, s, B, t_arr = item
            h = handles[s]
            ts = time.perf_counter()
            X = np.empty_like(B)
            info = []
            for j in range(B.shape[1]):
                h.set_rhs(B[:, j])
                h.set_x(None)
                _, _, inf, _ = h.solve(tol=w["tol"], nmax=w["nmax"])
                X[:, j] = h.get_x()
                info.append(inf)
            te = time.perf_counter()
            out[i] = (X if i % SAMPLE == 0 else None, info, 1e3 * (ts - t_arr), 1e3 * (te - t_arr))
            if i and i % 2000 == 0:
                _progress(f"FP64 baseline: {i} requests solved")
    nb = len(reqs) if args.serve_fp64_requests <= 0 else min(len(reqs), args.serve_fp64_requests)
    base = [None] * nb
    th = threading.Thread(target=fp64_worker, args=(base,))
    th.start()
    start = time.perf_counter()
    for i, (s, cols) in enumerate(reqs[:nb]):
        dt = start + arrive[i] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        with cv:
            queue.append((i, s, pools[s][1][:, cols], time.perf_counter()))
            cv.notify()
    with cv:
        queue.append(None)
        cv.notify()
    th.join()
    t_base = time.perf_counter() - start
    ncols = sum(len(c) for _, c in reqs)
    lat_s = np.array([r[0] for r in served])
    lat_b = np.array([b[3] for b in base])
    # replay wall time: first arrival (t = 0) to the last solution on the
    # host, from the per-request latencies (the harvesting thread's own
    # checks do not count)
    t_served = float(np.max(arrive + lat_s / 1e3))
    t_base = float(np.max(arrive[:nb] + lat_b / 1e3))
    ncols_b = sum(len(c) for _, c in reqs[:nb])
    # accuracy: true residual and error vs x* of every served column; error
    # vs the FP64 path's solution on every 10th request
    worst_res = max(r[4] for r in served)
    worst_x = max(r[5] for r in served)
    worst_64 = 0.0
    for r, b in zip(served, base):
        if r[6] is None:
            continue
        for j in range(r[6].shape[1]):
            worst_64 = max(worst_64, float(np.linalg.norm(r[6][:, j] - b[0][:, j]) / np.linalg.norm(b[0][:, j])))
    for h in handles:
        h.close()
    value = ncols / t_served
    v64 = ncols_b / t_base
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1, "steps": len(reqs),
        "warmup": len(systems), "ms_per_step": round(1e3 * t_served / len(reqs), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 refinement", "data": "synthetic",
        "config": {"workload": w["desc"], "systems": len(systems), "requests": len(reqs),
                   "rate_per_s": args.serve_rate, "rhs_total": ncols, "slots": args.serve_slots, "tol": w["tol"]},
        "latency_ms": {"p50": _pct(lat_s, 50), "p99": _pct(lat_s, 99)},
        "queue_ms": {"p50": _pct([r[1] for r in served], 50), "p99": _pct([r[1] for r in served], 99)},
        "service_ms": {"p50": _pct([r[7] for r in served], 50), "p99": _pct([r[7] for r in served], 99)},
        "server": st, "small_n": args.serve_small_n, "warmup_s": round(t_warm, 1),
        "mean_batch_columns": round(float(np.mean([r[2] for r in served])), 2),
        "info_nonzero": int(sum(r[3] for r in served)), "info_counts": info_counts, "failed_sample": failed,
        "fp64_info_nonzero": int(sum(int(np.count_nonzero(b[1])) for b in base)),
        "true_rel_residual_max": worst_res, "forward_error_vs_xstar_max": worst_x,
        "rel_error_vs_fp64_solution_max": worst_64, "fp64_comparison_sample": f"every {SAMPLE}th request",
        "fp64_one_at_a_time": {"value": round(v64, 3), "unit": "solves/s",
                               "latency_ms": {"p50": _pct(lat_b, 50), "p99": _pct(lat_b, 99)},
                               "wall_s": round(t_base, 2), "requests": nb,
                               "spmv": "autotuned single-RHS formats" if not args.no_autotune else "staged",
                               "autotune_s": round(t_tune, 1)},
        "speedup_vs_fp64_one_at_a_time": round(value / v64, 3),
        "wall_s": round(t_served, 2),
        "roofline": None, "cpu_baseline": None,
        "note": "solves/s counts right-hand-side columns over the replay (first arrival to last completion); "
                "with arrivals slower than service it is arrival-bound and latency is the figure of merit",
    }
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cd3d-128", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision legs")
    ap.add_argument("--inner-tol", type=float, default=1e-4,
                    help="relative residual the mixed legs' low-precision inner solves stop at")
    ap.add_argument("--kernel-timing", choices=("timed-region", "separate"), default="timed-region",
                    help="bracket the dominant kernel's launches with HIP events inside the timed region "
                         "(default), or time the region without events and time that kernel in extra solves")
    ap.add_argument("--serve-systems", type=int, default=200)
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--serve-small-n", type=int, default=4096,
                    help="systems with at most this many rows go to the persistent batched kernel (0: never)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default SpMV row blocks and kernels")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch solve passes kernel by kernel instead of replaying captured hipGraphs "
                         "(for kernel tracers that do not follow graph launches)")
    ap.add_argument("--serve-fp64-requests", type=int, default=0,
                    help="replay only the first N requests on the FP64 one-at-a-time path (0: all)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy of 1 GiB (x10), the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE hardware counters in rocprofv3 --pmc runs")
    ap.add_argument("--pmc-summary", nargs=5, metavar=("FETCH_DB", "WRITE_DB", "COPY_FETCH_DB", "COPY_WRITE_DB", "BENCH_JSON"),
                    help="summarize rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE result databases of this workload's "
                         "bench and of --calibrate-copy into profiles/traffic_<workload>.json (no GPU use)")
    ap.add_argument("--legs", default=None,
                    help="comma-separated workloads run after the headline one in compact form and reported under "
                         "'legs' (default with the default workload: " + ",".join(DEFAULT_LEGS) + "; 'none' "
                         "to skip)")
    ap.add_argument("--trace-summary", metavar="TRACE_DB",
                    help="print the per-kernel table of a rocprofv3 --kernel-trace result database (no GPU use)")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
    if args.pmc_summary:
        if args.workload == "st27-100-rhs16":
            pmc_block_summary(*args.pmc_summary[:4], args.workload)
        else:
            pmc_summary(*args.pmc_summary, args.workload)
        return
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
    out = run_workload(args, args.workload)
    legs = args.legs if args.legs is not None else (",".join(DEFAULT_LEGS) if args.workload == "cd3d-128" else "none")
    if legs != "none":
        out["legs"] = {}
        for name in legs.split(","):
            if name == args.workload:
                continue
            _progress(f"leg {name}")
            t0 = time.perf_counter()
            leg = run_workload(_leg_args(args, name), name)
            out["legs"][name] = _leg_summary(leg, time.perf_counter() - t0)
    print(json.dumps(out), flush=True)


# The other configurations, run after the headline workload in compact form
# (fewer steps and requests, short CPU baselines) so that a default run
# reports every configuration in a few minutes.
DEFAULT_LEGS = ("cd3d-128-pe5", "st27-100-rhs16", "serve-mix", "cd2d-100", "cd3d-128-sym", "cd3d-128-pe5-c2",
                "cd3d-128-pe5-i", "cd3d-128-pe5-j")
LEG_SETTINGS = {
    "cd3d-128-pe5": dict(steps=5, warmup=1, cpu_seconds=3.0),
    "cd3d-128-pe5-c2": dict(steps=3, warmup=1, cpu_seconds=3.0),
    "cd3d-128-pe5-i": dict(steps=3, warmup=1, cpu_seconds=3.0),
    "cd3d-128-pe5-j": dict(steps=3, warmup=1, cpu_seconds=3.0),
    "cd3d-128-sym": dict(steps=5, warmup=1, cpu_seconds=3.0),
    "st27-100-rhs16": dict(steps=5, warmup=1, cpu_seconds=3.0),
    "serve-mix": dict(serve_requests=2000, serve_fp64_requests=300),
    "cd2d-100": dict(steps=100, warmup=5, cpu_seconds=3.0),
}


def _leg_args(args, name):
    a = argparse.Namespace(**vars(args))
    a.workload = name
    for k, v in LEG_SETTINGS.get(name, {}).items():
        setattr(a, k, v)
    return a


def _leg_summary(out, wall_s):
    """The figures of a workload's full line that a compact leg reports."""
    keep = ("value", "unit", "ms_per_step", "steps", "dtype", "latency_ms", "queue_ms", "service_ms",
            "iterations", "info", "true_rel_residual", "true_rel_residual_max", "forward_error_vs_xstar",
            "forward_error_vs_xstar_max", "rel_error_vs_oracle", "rel_error_vs_oracle_column0",
            "rel_error_vs_fp64_solution_max", "outer_steps", "inner_passes", "fp64_one_at_a_time", "fp64_block",
            "speedup_vs_fp64_one_at_a_time", "pcie_inclusive", "info_nonzero", "mean_batch_columns")
    leg = {k: out[k] for k in keep if k in out}
    leg["workload"] = out["config"]["workload"]
    if out.get("mixed_precision"):
        leg["mixed_precision"] = {k: {q: v[q] for q in ("value", "outer_steps", "inner_passes", "info",
                                                           "true_rel_residual", "rel_error_vs_fp64_solution")}
                                  for k, v in out["mixed_precision"].items()}
        leg["mixed_speedup_vs_fp64"] = {k: round(v["value"] / out["value"], 3)
                                        for k, v in out["mixed_precision"].items()}
    if out.get("roofline"):
        r = out["roofline"]
        leg["roofline"] = {k: r[k] for k in ("bound", "kernel", "achieved", "peak", "unit", "frac", "traffic",
                                              "avg_us", "format_frac") if k in r}
    if out.get("cpu_baseline"):
        c = out["cpu_baseline"]
        leg["cpu_baseline"] = {k: c[k] for k in ("value", "unit", "cores", "kind", "sample", "single_thread")
                               if k in c}
    leg["wall_s"] = round(wall_s, 1)
    return leg


def run_workload(args, name):
    w = WORKLOADS[name]
    if name == "st27-100-rhs16":
        return bench_block(args, w)
    if name == "serve-mix":
        return bench_serve(args, w)
    return bench_single(args, w)


def bench_single(args, w):
    """Single-system workload: one step = one full FP64 solve on the resident
    system, then the mixed-precision legs on the same system."""
    import cmamd
    A, xs, b, F, t_build = build_workload(w)
    sym = bool(w.get("sym", False))  # A is the stored upper triangle; residuals use the full matrix
    A_full = A
    if sym:
        import scipy.sparse as sps
        U = A.to_scipy()
        A_full = cmamd.Csr.from_scipy(U + U.T - sps.diags(U.diagonal()))
    nnz_h = F.H.nnz if F.H is not None else 0
    print(f"[bench] workload {args.workload}: n={A.n} nnz(A)={A_full.nnz} nnz(G)={F.G.nnz} "
          f"nnz(H)={nnz_h} factors built in {t_build:.1f}s", file=sys.stderr, flush=True)

    sys_ = cmamd.System(w["prec"], A, F, a_symmetric=sym, nr=w.get("nr", 1), mode=w.get("mode", "E"))
    if args.no_graphs:
        sys_.graphs(False)
    tuned = None
    if not args.no_autotune:
        # per-matrix SpMV row-block size, and for the low-precision copies the
        # 16-bit-index product against the 32-bit one, chosen by timing
        if not args.no_mixed:
            sys_.enable_mixed(32)
            sys_.enable_mixed(16)
        tuned = sys_.autotune(reps=10)
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
    fstats = sys_.kernel_format_bytes()
    dom = max((c for c in ("A", "F1", "F2", "LU") if stats[c][0] > 0), key=lambda c: stats[c][1])
    # timed region: only the dominant kernel's launches carry events, and only in the first
    # SAMPLE_PASSES passes of each solve (launched kernel by kernel); the other passes replay from
    # the captured graphs, as in an unprofiled solve.  A dominant category of many launches per pass
    # (the level-scheduled triangular solves: thousands per pass) is timed in separate solves
    # instead: events around each of its launches would slow the timed solves themselves.
    profiled_passes = max(1, stats["A"][0] // 2)  # passes of the two profiled solves (two A products each)
    kernel_timing = args.kernel_timing
    if stats[dom][0] / profiled_passes > DOM_LAUNCHES_PER_PASS_MAX:
        kernel_timing = "separate"
    sys_.profile(kernel_timing == "timed-region", cats=(dom,), sample=SAMPLE_PASSES)
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
    if kernel_timing == "separate":
        sys_.profile(True, cats=(dom,))
        for _ in range(3):
            sys_.set_x(None)
            sys_.solve(tol=w["tol"], nmax=w["nmax"])
    dstats = sys_.kernel_stats()[dom]
    dfby = sys_.kernel_format_bytes()[dom]
    sys_.profile(False)
    x = sys_.get_x()
    true_res = float(np.linalg.norm(b - cmamd.csr_mv_host(A_full, x)) / np.linalg.norm(b))
    fwd_err = float(np.linalg.norm(x - xs) / np.linalg.norm(xs))
    # PCIe-inclusive rate: calls of the reference-signature drop-in entry
    # point, each of which uploads A, the factors and b, solves, and copies x
    # back -- the reference's callers make one such call per right-hand side
    t_calls = []
    for _ in range(DROPIN_CALLS):
        t0 = time.perf_counter()
        if w["prec"] == "pdfiim":
            dr = cmamd.pdfiim_epbicgstab(A, b, F, tol=w["tol"], nmax=w["nmax"], a_symmetric=sym)
        elif w["prec"] == "afiim":
            dr = cmamd.afiim_epbicgstab(A, b, F, tol=w["tol"], nmax=w["nmax"])
        else:
            dr = cmamd.acifim_pbicgstab(A, b, F, nr=w.get("nr", 1), mode=w.get("mode", "E"), tol=w["tol"],
                                        nmax=w["nmax"])
        t_calls.append(time.perf_counter() - t0)
    t_dropin = float(np.median(t_calls))
    mixed = {}
    xs_mixed = {}
    if not args.no_mixed:
        for bits in (32, 16):
            mixed[f"fp{bits}_factors"] = mixed_leg(sys_, bits, w, A_full, b, x, args)
            xs_mixed[f"fp{bits}_factors"] = sys_.get_x()
    formats = sys_.spmv_formats() if tuned else {}
    # remainder entries per row of each matrix's diagonal index: the diagonal kernel stages the
    # remainder through LDS from DIA_REM_LDS_MIN (0.1) per row up, a template argument of its name
    rem_per_row = {}
    for mat in ("A", "G", "H", "GT", "L", "U"):
        try:
            rem_per_row[mat] = sys_.dia_info(mat)[1] / A.n
        except cmamd.EngineError:  # no such matrix in this system
            pass
    sys_.close()

    value = args.steps / elapsed
    kernels = {}
    for cat, (n_l, ms, by) in stats.items():
        if n_l == 0:
            continue
        kernels[cat] = {"what": CAT_NAMES[cat], "launches": n_l, "avg_us": 1e3 * ms / n_l,
                        "bytes_per_launch": by / n_l, "achieved_GBs": by / ms / 1e6,
                        "format_bytes_per_launch": fstats[cat] / n_l, "share_of_gpu_time": None}
    tot_ms = sum(stats[c][1] for c in kernels)
    for c in kernels:
        kernels[c]["share_of_gpu_time"] = stats[c][1] / tot_ms
    n_l, ms, by = dstats
    d_avg_us = 1e3 * ms / n_l
    d_gbs = by / ms / 1e6
    kname = CAT_KERNELS.get(dom, dom)
    mode = w.get("mode", "E")
    if dom == "LU":  # mode C's L / U products, or modes I / J's level-scheduled triangular solves
        if mode in ("I", "J"):
            kname = "cm::k_trsv_level<double, double> (one launch per dependency level)"
        else:
            fmt_l = formats.get(64, {}).get("L", "staged") if tuned else "staged"
            if fmt_l in ("dia", "dia2"):
                lds = "true" if rem_per_row.get("L", 0.0) >= 0.1 else "false"
                shape = "8, 1" if fmt_l == "dia" else "9, 2"
                kname = f"cm::spmv_dia<double, double, double, cm::EpiSubFrom<double>, {shape}, {lds}>"
            else:
                tile = (tuned or {}).get("L") or 2048
                kname = f"cm::spmv_kernel<{tile}, double, double, double, cm::EpiSubFrom<double> >"
    elif mode == "C" and w.get("nr", 1) > 1 and dom in ("F1", "F2"):
        kname = kname.split(", cm::Epi")[0] + ", cm::EpiStore|EpiAccum|EpiAccumScale<double> >"
    if tuned and dom != "LU":
        mat = {"A": "A", "F1": "GT" if tuned["GT"] else "H", "F2": "G"}.get(dom)
        fmt = formats.get(64, {}).get(mat, "staged") if mat else "staged"
        if fmt.startswith("slot"):  # slot-major kernel: spmv_slot<V, X, Acc, Epi, 16-bit offsets>
            kname = (kname.replace("cm::spmv_kernel<2048, ", "cm::spmv_slot<")[:-2] +
                     (", true>" if fmt == "slot16" else ", false>"))
        elif fmt in ("dia", "dia2"):
            # diagonals + CSR remainder: spmv_dia<V, X, Acc, Epi, CHUNK, rows per thread, remainder via LDS>;
            # the second variant has 9 diagonals in flight and 2 rows per thread
            lds = "true" if rem_per_row.get(mat, 0.0) >= 0.1 else "false"
            shape = "8, 1" if fmt == "dia" else "9, 2"
            kname = kname.replace("cm::spmv_kernel<2048, ", "cm::spmv_dia<")[:-2] + f", {shape}, {lds}>"
        else:
            tile = {"A": tuned["A"], "F1": tuned["GT"] or tuned["H"], "F2": tuned["G"]}.get(dom)
            if tile:
                kname = kname.replace("<2048,", f"<{tile},")
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(d_gbs, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(d_gbs / HBM_PEAK_GBS, 4),
                "traffic": None, "bytes_per_launch": by / n_l, "avg_us": round(d_avg_us, 2),
                "launches_timed": n_l,
                "bytes_definition": "CSR algorithmic bytes nnz*(value bytes + 4) + 4*(n+1) plus the vectors "
                                    "read and written, independent of the storage format",
                "format_bytes_per_launch": dfby / n_l,
                "format_frac": round(dfby / ms / 1e6 / HBM_PEAK_GBS, 4)}
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
        "config": {"workload": w["desc"], "n": A.n, "nnz_A": A_full.nnz, "nnz_G": F.G.nnz, "nnz_H": nnz_h,
                   "tol": w["tol"], "nmax": w["nmax"], "symmetric": sym},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "iterations": int(np.median(iters)), "info": info, "resval": rv, "true_rel_residual": true_res,
        "forward_error_vs_xstar": fwd_err,
        "mixed_precision": mixed, "autotune_tiles": tuned,
        "spmv_formats": {str(k): v for k, v in formats.items()} if tuned else None,
        "pcie_inclusive": {"ms_per_solve": round(1e3 * t_dropin, 1), "iterations": dr.iter, "info": dr.info,
                           "calls_ms": [round(1e3 * t, 1) for t in t_calls],
                           "what": f"median of {DROPIN_CALLS} back-to-back calls of the drop-in entry point with "
                                   "host arrays, each uploading and checking A, the factors and b, solving and "
                                   "copying x back (not the bench value)"},
        "roofline": roofline, "kernels": kernels,
        "kernels_note": "per-kernel table from 2 extra solves with every launch bracketed by HIP events; " + (
            "roofline from the timed region, where only the dominant kernel is bracketed, in the "
            f"first {SAMPLE_PASSES} passes of each solve (the rest replay from hipGraphs)"
            if kernel_timing == "timed-region" else
            "roofline from 3 solves after the timed region with the dominant kernel's launches bracketed "
            "(the timed solves carry no events)"),
    }
    if not args.no_cpu:
        out["cpu_baseline"], x_or = cpu_baseline(w, A, b, F, args.cpu_seconds)
        # full-size parity: the GPU solutions against the oracle's FP64 solution of the same system
        out["rel_error_vs_oracle"] = {"fp64": _rel(x, x_or), **{k: _rel(v, x_or) for k, v in xs_mixed.items()}}
    else:
        out["cpu_baseline"] = None
    return out


if __name__ == "__main__":
    main()
