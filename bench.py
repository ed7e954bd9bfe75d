"""Benchmark: COCR time-to-1e-8 per (frequency x polarization) system, plus SpMV HBM GB/s.

Workload (the mixed-precision configuration): a 100x100x60-cell mesh (1,844,260 complex edge
unknowns), 50 m core cells, 8 padding cells per side growing x1.4, 6 air cells at 1e-8 S/m;
seeded log10(rho) ~ U[0,3] smoothed by a 3-cell Gaussian; 8 log-spaced frequencies from 1e-3 to
1e2 Hz x 2 polarizations = 16 systems sharing K and S (A = K + i*omega*S).

A "step" solves all 16 systems to the reference's stop test (MRS-smoothed relative residual
< 1e-8) with the batched FP64 solver: one launch sequence over all systems, the matrix read once
per iteration for the whole batch, systems retiring independently.
value = wall seconds of one step / 16 (inputs resident in HBM, timed between device
synchronisations). Reported beside it:
  fp64_single    the plain FP64 GPU solver, one system at a time (the comparison baseline)
  mixed_batched  FP32 inner COCR + FP64 refinement on the true residual, FP64 finish for the
                 systems whose refinement stalls
The roofline is that of the dominant kernel per iteration (the longer of the batched FP64
update pass and the batched FP64 SpMV), timed with HIP events on the solver's stream; the SpMV,
the update pass and the x/p pass (every 8th or 16th iteration) are each reported with their HBM
traffic from rocprofv3 counters. The CPU baseline is the oracle (FP64 CPU restatement of the
reference's COCR_FP) run to the stop test on one system of the same workload, on the host's
cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "COCR time-to-1e-8 per freq×pol (s) + SpMV HBM GB/s; mixed/batched vs FP64 GPU"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


_hip = None


def sync():
    """Device-wide synchronisation (hipDeviceSynchronize) around the timed region."""
    global _hip
    if _hip is None:
        import ctypes

        _hip = ctypes.CDLL("libamdhip64.so")
    _hip.hipDeviceSynchronize()


def build_workload(args):
    from mtcalc_with_dfp_cocr_amd import model as M

    t0 = time.time()
    m = M.random_model(args.nx, args.ny, args.nz, seed=args.seed, h=50.0, name="bench")
    freqs = M.frequencies(1e-3, 1e2, args.nfreq)
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(m)
    ma = M.mesh_arrays(m)
    prs = np.stack([M.assemble_rhs(m, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    log(f"workload {m.shape_cells} cells, nb={p.nb}, {len(cases)} systems, "
        f"host assembly {time.time() - t0:.1f}s")
    return M, m, p, (kd, kg, sd, sg), ma, cases, prs, om


def batched_only(m, freqs, args, label):
    """Batched FP64 solve of all (frequency x polarization) systems of one model, autotuned, in
    the device row order --order-block sets; returns a summary (warm-up solves, then `steps`
    timed solves), with the mixed-precision form beside it (unless --no-mixed) and, for a batch
    of at most 16 systems, the SpMV, update pass and x/p pass rates on it."""
    from mtcalc_with_dfp_cocr_amd import model as M
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    t0 = time.time()
    cases = [(f, a) for f in freqs for a in (1, 0)]
    p, kd, kg, sd, sg = M.assemble_ks(m)
    ma = M.mesh_arrays(m)
    prs = np.stack([M.assemble_rhs(m, f, a) for f, a in cases])
    om = np.array([M.omega_of(f) for f, _ in cases])
    t_host = time.time() - t0
    log(f"{label}: {m.shape_cells} cells, nb={p.nb}, {len(cases)} systems, host {t_host:.1f}s")
    S = Solver(0)
    if args.order_block:
        S.set_mesh_order(ma, args.order_block)
    S.set_matrix_ks(p, kd, kg, sd, sg)
    S.set_gradient(ma)
    # batches of at most max_batch systems, neighbouring frequencies together (they converge at
    # similar rates), one after the other on the same handle
    w = args.max_batch if args.max_batch > 0 else len(cases)
    groups = [slice(i, min(i + w, len(cases))) for i in range(0, len(cases), w)]
    tune = {}
    for g in groups:
        S.set_rhs(prs[g], om[g])
        tune[f"{g.stop - g.start} systems"] = S.autotune("fp64", args.tune_reps)
    times = []
    st = None
    for k in range(args.warmup + args.steps):
        dt, st, rc = 0.0, [], 0
        for g in groups:
            S.set_rhs(prs[g], om[g])
            d, r, q = timed(S, "fp64", args)
            dt += d
            rc = rc or r
            st += q
        if k >= args.warmup:
            times.append(dt)
        log(f"{label}: {'warmup' if k < args.warmup else 'step'} {dt:.3f}s rc={rc}")
    info = S.info()
    out = {"model": label, "cells": list(m.shape_cells), "nb": int(p.nb),
           "systems": len(cases), "s_per_step": float(np.mean(times)),
           "s_per_system": float(np.mean(times)) / len(cases),
           "ok": all(q["req"] == 0 for q in st), "iters": [q["iter"] for q in st],
           "true_rel_res_max": max(q["true_rel_res"] for q in st),
           "device_bytes": info["device_bytes"], "launch": tune, "host_setup_s": t_host}
    if not args.no_mixed:
        # the mixed-precision form (FP32 inner COCR, FP64 refinement and finish), once
        dt, ms_st = 0.0, []
        for g in groups:
            S.set_rhs(prs[g], om[g])
            d, r, q = timed(S, "mixed", args)
            dt += d
            ms_st += q
        out["mixed"] = {"s_per_system": dt / len(cases), "ok": all(q["req"] == 0 for q in ms_st),
                        "true_rel_res_max": max(q["true_rel_res"] for q in ms_st),
                        "fp64_finish": sum(q["flags"] & 1 for q in ms_st),
                        "device_bytes": S.info()["device_bytes"]}
        log(f"{label}: mixed {out['mixed']}")
    if len(groups) == 1 and len(cases) <= 16:
        # kernel rates on this batch, as the solver launches them (HIP events)
        S.set_rhs(prs, om)
        gbs = lambda b, ms: b / (ms * 1e-3) / 1e9
        for name, (ms, by) in (("spmv_fp64_batched", S.bench_spmv("fp64", 10)),
                               ("update_fp64_batched", S.bench_update("fp64", 10)),
                               ("xp_fp64_batched", S.bench_flush(10))):
            out[name] = {"avg_ms": ms, "algorithmic_bytes": by, "achieved_gbs": gbs(by, ms),
                         "frac": gbs(by, ms) / HBM_PEAK_GBS}
        log(f"{label}: kernel rates { {k: out[k] for k in out if k.endswith('_batched')} }")
    S.close()
    return out


def run_config(args):
    """Configs 3 (four models x 32 systems) and 5 (256x256x128, 8 systems), and config 2 in
    batches of --max-batch systems: batched FP64, one JSON line in the benchmark's format."""
    from mtcalc_with_dfp_cocr_amd import model as M

    runs = []
    if args.config == 2:
        # the headline workload in batches of at most --max-batch systems (batch-width sweep)
        m = M.random_model(args.nx, args.ny, args.nz, seed=args.seed, h=50.0, name="bench")
        freqs = M.frequencies(1e-3, 1e2, args.nfreq)
        runs.append(batched_only(m, freqs, args, f"bench_batches_of_{args.max_batch}"))
        workload = (f"batched FP64 COCR, {args.nx}x{args.ny}x{args.nz} cells, {args.nfreq} freqs "
                    f"1e-3..1e2 Hz x 2 pol, in batches of at most {args.max_batch} systems")
    elif args.config == 3:
        shapes = [(32, 32, 32), (48, 48, 48), (64, 64, 48), (96, 96, 64)]
        freqs = M.frequencies(1e-3, 1e3, 16)
        for k, (nx, ny, nz) in enumerate(shapes):
            m = M.random_model(nx, ny, nz, seed=args.seed + k, h=50.0, name=f"c3_{nx}x{ny}x{nz}")
            runs.append(batched_only(m, freqs, args, m.name))
        workload = ("4 seeded models 32^3, 48^3, 64x64x48, 96x96x64; 16 freqs 1e-3..1e3 Hz x 2 pol"
                    " = 32 systems per model, one batch per model")
    else:
        m = M.ellipsoid_model(256, 256, 128, seed=args.seed)
        freqs = M.frequencies(1e-2, 1e1, 4)
        runs.append(batched_only(m, freqs, args, m.name))
        workload = ("256x256x128 cells, layered background + 10 ellipsoids, 4 freqs 1e-2..1e1 Hz"
                    " x 2 pol = 8 systems, one batch")
    total_sys = sum(r["systems"] for r in runs)
    total_s = sum(r["s_per_step"] for r in runs)
    out = {"metric": METRIC, "value": total_s / total_sys, "unit": "s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * total_s,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 (complex128)", "data": "synthetic (seeded conductivity models)",
           "config": {"workload": workload, "config": args.config, "parallelism": "single GPU",
                      "order_block": args.order_block},
           "runs": runs}
    print(json.dumps(out), flush=True)


def host_cores():
    """(nproc, cores this process may use, cores granted to it): os.cpu_count(), the affinity
    mask's size, and that capped by OMP_NUM_THREADS where the environment sets it (the GPU box
    grants a share of its host cores that way)."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    granted = aff
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        granted = min(granted, int(omp))
    return nproc, aff, granted


def cpu_solve(M, m, ma, case, threads):
    """The oracle (FP64 CPU restatement of COCR_FP) run to the reference's stop test on one
    system: (x, stats); stats["seconds"] is its kit interval (DFP set-up + iterations)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    f, a = case
    s = M.assemble(m, f, a)
    x, st, _ = O.pcocr(s, ma, eps=1e-8, maxiter=20000, nthreads=threads)
    return x, st


# SpMV kernel of each layout name the autotuner reports
SPMV_KERNELS = {"ell": "k_spmv", "rowpack": "k_spmv_rg", "rowblock": "k_spmv_blk",
                "stencil": "k_spmv_stc", "stencil_lds": "k_spmv_stl"}


def measure_traffic(args, sp):
    """HBM bytes per launch of the batched FP64 SpMV, the fused update pass and the x/p pass:
    two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE) over this script's --kernels-only
    mode, each scaled by what the counter reads for the 1 GiB copy kernel k_copy (known read and
    write bytes). Returns {"spmv": {...}, "update": {...}, "flush": {...}}; None when rocprofv3
    is unavailable or a pass fails."""
    import csv
    import re
    import shutil
    import signal
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    pats = {"spmv": re.compile(r"k_spmv(_rg|_blk|_stc|_stl)?<double, %d, 0>" % sp),
            "update": re.compile(r"k_xr_fused<double, %d, true(, true)?>" % sp),
            "flush": re.compile(r"k_flush_xp<%d(, \d+)?>" % sp)}
    res = {k: {} for k in pats}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="mtcocr_pmc_", dir="/tmp")
        cmd = [exe, "--pmc", ctr, "-d", d, "-o", "run", "--output-format", "csv", "--",
               sys.executable, os.path.abspath(__file__), "--kernels-only", "--no-mixed",
               "--spmv-reps", "5", "--tune-reps", "2", "--order-block",
               str(args.order_block), "--nx", str(args.nx), "--ny",
               str(args.ny), "--nz", str(args.nz), "--nfreq", str(args.nfreq), "--seed",
               str(args.seed)]
        log(f"traffic pass {ctr}")
        env = dict(os.environ, TMPDIR="/tmp")
        p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                             stderr=subprocess.DEVNULL, start_new_session=True)
        try:
            rc = p.wait(timeout=240)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            log(f"traffic pass {ctr} timed out")
            return None
        files = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs
                 if f.endswith("counter_collection.csv")]
        if rc != 0 or not files:
            log(f"traffic pass {ctr} failed (rc={rc})")
            return None
        rows = sorted(csv.DictReader(open(files[0])), key=lambda r: int(r["Dispatch_Id"]))
        copy = [float(r["Counter_Value"]) for r in rows if "k_copy" in r["Kernel_Name"]
                and r["Counter_Name"] == ctr]
        shutil.rmtree(d, ignore_errors=True)
        if not copy:
            return None
        factor = (1 << 30) / (float(np.mean(copy)) * 1024.0)
        for name, pat in pats.items():
            vals = [float(r["Counter_Value"]) for r in rows if pat.search(r["Kernel_Name"])
                    and r["Counter_Name"] == ctr]
            if len(vals) < 5:
                return None
            kib = float(np.mean(vals[-5:]))           # the timed launches at the tuned shape
            res[name][ctr] = {"kib_per_dispatch": kib, "calibration": factor,
                              "bytes": kib * 1024.0 * factor}
    return res


def impedance_agreement(M, m, cases, x_batched, x_single, x_mixed):
    """Largest relative difference of Zxy / Zyx at a 5x5 receiver grid (200 m spacing, centred)
    per frequency: batched FP64 vs one-at-a-time FP64, and mixed vs batched FP64."""
    from mtcalc_with_dfp_cocr_amd.receivers import ReceiverOperator, impedance

    g = np.arange(-400.0, 401.0, 200.0)
    rx, ry = np.meshgrid(g, g, indexing="ij")
    R = ReceiverOperator(m, np.stack([rx.ravel(), ry.ravel(), np.zeros(rx.size)], axis=1))
    freqs = sorted({c[0] for c in cases})
    idx = {c: k for k, c in enumerate(cases)}
    to_c = lambda v: v[0::2] + 1j * v[1::2]

    def Z(xs, f):
        Ex, Hx = R.fields(to_c(xs[idx[(f, 1)]]), f, 1)
        Ey, Hy = R.fields(to_c(xs[idx[(f, 0)]]), f, 0)
        return impedance(Ex, Hx, Ey, Hy)

    def rel(a, b):
        return float(max(np.max(np.abs(a[:, i, j] - b[:, i, j]) / np.abs(b[:, i, j]))
                         for i, j in ((0, 1), (1, 0))))

    out = {"receivers": "5x5, 200 m spacing", "freqs": [float(f) for f in freqs]}
    have = [f for f in freqs if idx[(f, 1)] in x_single and idx[(f, 0)] in x_single]
    if have:
        out["batched_vs_single_fp64"] = [rel(Z(x_batched, f), Z(x_single, f)) for f in have]
    if x_mixed is not None:
        out["mixed_vs_fp64"] = [rel(Z(x_mixed, f), Z(x_batched, f)) for f in freqs]
    return out


def timed(S, prec, args):
    sync()
    t0 = time.perf_counter()
    rc, st = S.solve_resident(args.eps, args.maxiter, prec)
    sync()
    return time.perf_counter() - t0, rc, st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=100)
    ap.add_argument("--ny", type=int, default=100)
    ap.add_argument("--nz", type=int, default=60)
    ap.add_argument("--nfreq", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--maxiter", type=int, default=20000)
    ap.add_argument("--fp64-systems", type=int, default=-1,
                    help="how many systems the one-at-a-time FP64 comparison solves (-1: all)")
    ap.add_argument("--cpu-system", type=int, default=13,
                    help="system the CPU baseline runs to convergence (13: 19.3 Hz, Y)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision run")
    ap.add_argument("--spmv-reps", type=int, default=20)
    ap.add_argument("--order-block", type=int, default=-1,
                    help="device row order: -1 the grid order of the box mesh (enables the "
                         "stencil SpMV), b > 0 pencils of b x b nodes, 0 the file order")
    ap.add_argument("--tune-reps", type=int, default=10,
                    help="launches per candidate shape in the autotuner")
    ap.add_argument("--config", type=int, default=2, choices=(2, 3, 5),
                    help="2: the headline (default); 3: four models x 32 systems; 5: 256x256x128")
    ap.add_argument("--max-batch", type=int, default=0,
                    help="solve in batches of at most this many systems (0: all); with config 2, "
                         "only the batched FP64 solve runs (a batch-width sweep)")
    ap.add_argument("--adjoint-cells", type=int, default=4,
                    help="cells of the config-4 adjoint vs finite-difference check (0: skip)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes that measure the kernels' HBM "
                         "traffic")
    ap.add_argument("--kernels-only", action="store_true",
                    help="profiling mode: autotune, then only the batched SpMV, update pass and "
                         "x/p pass (--spmv-reps launches each) and the 1 GiB copy calibration "
                         "kernel; prints one line")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(systems are independent and would shard across GPUs with no communication)",
              file=sys.stderr)
        sys.exit(2)

    if args.config != 2 or args.max_batch > 0:
        run_config(args)
        return

    M, m, p, ks, ma, cases, prs, om = build_workload(args)
    from mtcalc_with_dfp_cocr_amd.cocr import Solver

    nsys = len(cases)
    S = Solver(0)
    if args.order_block:
        S.set_mesh_order(ma, args.order_block)
    S.set_matrix_ks(p, *ks)
    # BuildGMatrix (host) + its upload: part of the reference's kit interval (pcocr.cpp:61-72),
    # done once per mesh here and shared by every system; reported beside value
    t0 = time.perf_counter()
    S.set_gradient(ma)
    t_grad = time.perf_counter() - t0
    log(f"device layout: {S.info()}, gradient set-up {t_grad:.3f}s")

    # ---- launch shapes: autotuned on the batch before anything is timed -------------------
    S.set_rhs(prs, om)
    tune = {"batched_fp64": S.autotune("fp64", args.tune_reps)}
    if not args.no_mixed:
        tune["batched_fp32vec"] = S.autotune("mixed", args.tune_reps)
    S.set_rhs(prs[:1], om[:1])
    tune["single_fp64"] = S.autotune("fp64", args.tune_reps)
    log(f"autotuned launch shapes: {tune}")

    if args.kernels_only:
        S.set_rhs(prs, om)
        ms, by = S.bench_spmv("fp64", args.spmv_reps)
        ums, uby = S.bench_update("fp64", args.spmv_reps)
        fms, fby = S.bench_flush(args.spmv_reps)
        cms = S.bench_copy(1 << 30, args.spmv_reps)
        print(json.dumps({"spmv_fp64_batched_ms": ms, "spmv_algorithmic_bytes": by,
                          "spmv_gbs": by / (ms * 1e-3) / 1e9,
                          "update_fp64_batched_ms": ums, "update_algorithmic_bytes": uby,
                          "update_gbs": uby / (ums * 1e-3) / 1e9,
                          "xp_fp64_batched_ms": fms, "xp_algorithmic_bytes": fby,
                          "xp_gbs": fby / (fms * 1e-3) / 1e9, "copy_bytes": 1 << 30,
                          "copy_ms": cms, "copy_gbs": 2 * (1 << 30) / (cms * 1e-3) / 1e9,
                          "launch": tune}), flush=True)
        S.close()
        return

    # ---- headline: FP64, all systems batched --------------------------------------------
    for w in range(args.warmup):
        S.set_rhs(prs, om)
        dt, rc, st = timed(S, "fp64", args)
        log(f"warmup {w}: {dt:.3f}s rc={rc}")
    times = []
    stats = None
    for k in range(args.steps):
        S.set_rhs(prs, om)
        dt, rc, stats = timed(S, "fp64", args)
        times.append(dt)
        log(f"step {k}: {dt:.3f}s rc={rc}")
    x_batched = S.get_solution()
    ms_step = 1e3 * float(np.mean(times))
    value = ms_step / 1e3 / nsys
    batched_ok = all(q["req"] == 0 for q in stats)
    log(f"fp64 batched: {ms_step:.1f} ms/step, {value:.4f} s/system, ok={batched_ok}, "
        f"iters {[q['iter'] for q in stats]}")

    # ---- mixed precision, batched ---------------------------------------------------------
    mixed = None
    x_mixed = None
    if not args.no_mixed:
        S.set_rhs(prs, om)
        dt, rc, st = timed(S, "mixed", args)
        xm = S.get_solution()
        x_mixed = xm
        mixed = {"s_per_system": dt / nsys, "ok": all(q["req"] == 0 for q in st),
                 "true_rel_res_max": max(q["true_rel_res"] for q in st),
                 "inner_iters": [q["iter"] for q in st], "passes": [q["outer"] for q in st],
                 "fp64_finish": [q["flags"] & 1 for q in st],
                 "x_rel_diff_vs_fp64_max": float(max(np.linalg.norm(xm[k] - x_batched[k]) /
                                                     np.linalg.norm(x_batched[k])
                                                     for k in range(nsys)))}
        log(f"mixed batched: {mixed}")

    # ---- comparison: plain FP64 GPU solver, one system at a time ----------------------------
    n64 = nsys if args.fp64_systems < 0 else min(nsys, args.fp64_systems)
    fp64 = []
    x_single = {}
    for k in range(n64):
        S.set_rhs(prs[k:k + 1], om[k:k + 1])
        dt, rc, st = timed(S, "fp64", args)
        x64 = S.get_solution()[0]
        x_single[k] = x64
        d = np.linalg.norm(x64 - x_batched[k]) / np.linalg.norm(x64)
        fp64.append({"f": float(cases[k][0]), "pol": "X" if cases[k][1] else "Y", "s": dt,
                     "iter": st[0]["iter"], "batched_iter": stats[k]["iter"], "req": st[0]["req"],
                     "true_rel_res": st[0]["true_rel_res"], "batched_vs_single_rel": d})
        log(f"fp64 system {k}: {fp64[-1]}")
    fp64_per_sys = float(np.mean([r["s"] for r in fp64])) if fp64 else None

    # ---- SpMV roofline ----------------------------------------------------------------------
    S.set_rhs(prs, om)
    spmv_ms, spmv_bytes = S.bench_spmv("fp64", args.spmv_reps)
    spmv32_ms, spmv32_bytes = S.bench_spmv("mixed", args.spmv_reps)
    S.set_rhs(prs[:1], om[:1])
    spmv64_ms, spmv64_bytes = S.bench_spmv("fp64", args.spmv_reps)
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9
    ach, ach32, ach64 = gbs(spmv_bytes, spmv_ms), gbs(spmv32_bytes, spmv32_ms), gbs(spmv64_bytes, spmv64_ms)
    log(f"spmv fp64 x{nsys}: {spmv_ms:.3f} ms {ach:.0f} GB/s; fp32-vector x{nsys}: "
        f"{spmv32_ms:.3f} ms {ach32:.0f} GB/s; fp64 x1: {spmv64_ms:.3f} ms {ach64:.0f} GB/s")
    S.set_rhs(prs, om)
    upd_ms, upd_bytes = S.bench_update("fp64", args.spmv_reps)
    achu = gbs(upd_bytes, upd_ms)
    log(f"update pass fp64 x{nsys}: {upd_ms:.3f} ms {achu:.0f} GB/s")
    xp_ms, xp_bytes = S.bench_flush(args.spmv_reps)
    xp_period = S.get_xp_period()
    log(f"x/p pass fp64 x{nsys}: {xp_ms:.3f} ms {gbs(xp_bytes, xp_ms):.0f} GB/s")
    S.close()

    # ---- impedance agreement at a 5x5 surface grid (200 m spacing) ---------------------------
    zcmp = impedance_agreement(M, m, cases, x_batched, x_single, x_mixed)
    log(f"impedance agreement: {zcmp}")

    # ---- adjoint gradient (config 4) against central differences ----------------------------
    adj = None
    if args.adjoint_cells > 0:
        from mtcalc_with_dfp_cocr_amd import adjoint as AD

        t0 = time.perf_counter()
        import io
        from contextlib import redirect_stdout

        buf = io.StringIO()
        with redirect_stdout(buf):
            AD.main(["--fd-cells", str(args.adjoint_cells)])
        adj = json.loads(buf.getvalue().strip().splitlines()[-1])
        adj = {k: adj[k] for k in ("workload", "J", "eps", "h", "rel_err_norm", "rel_err",
                                   "rel_err_max", "directions",
                                   "seconds", "timings")}
        log(f"adjoint (config 4): {adj} ({time.perf_counter() - t0:.1f}s)")

    # ---- CPU baseline: the oracle to convergence on one system, beside the GPU's answer ----
    cpu = cpu_all = oracle_cmp = None
    if not args.no_cpu:
        nproc, aff, granted = host_cores()
        k = args.cpu_system
        ref_threads = max(1, granted - 1)  # the reference's cap (ControlOMP.cpp:46)
        xo, so = cpu_solve(M, m, ma, cases[k], ref_threads)
        label = f"{cases[k][0]:.3g} Hz {'X' if cases[k][1] else 'Y'}"
        cpu = {"value": so["seconds"], "unit": "s per system (time-to-1e-8)",
               "cores": ref_threads, "kind": "port",
               "sample": (f"oracle FP64 COCR_FP restatement run to the stop test on system {k} "
                          f"({label}, nb={p.nb}) on {ref_threads} threads (the reference's "
                          f"nproc-1 cap; host nproc {nproc}, affinity {aff}, granted {granted}): "
                          f"{so['iter']} iterations, DFP set-up included")}
        log(f"cpu baseline: {cpu}")
        if granted > ref_threads:
            xo2, so2 = cpu_solve(M, m, ma, cases[k], granted)
            cpu_all = {"value": so2["seconds"], "cores": granted, "iter": so2["iter"],
                       "nproc": nproc, "affinity": aff}
            log(f"cpu baseline, all granted cores: {cpu_all}")
        # parity at full size: the GPU's batched and one-at-a-time answers against the oracle's
        import scipy.sparse as sps

        rows = np.repeat(np.arange(p.nb), np.diff(p.ig))
        full = lambda d, g: (sps.csr_matrix((g, (rows, p.jg)), shape=(p.nb, p.nb))
                             + sps.csr_matrix((g, (p.jg, rows)), shape=(p.nb, p.nb))
                             + sps.diags(d)).tocsr()
        Kf, Sf = full(ks[0], ks[1]), full(ks[2], ks[3])

        def true_rel(v):
            b = M.to_complex(prs[k])
            xc = M.to_complex(v)
            return float(np.linalg.norm(b - (Kf @ xc + 1j * om[k] * (Sf @ xc)))
                         / np.linalg.norm(b))

        oracle_cmp = {"system": k, "f": float(cases[k][0]), "pol": "X" if cases[k][1] else "Y",
                      "oracle_iter": so["iter"], "oracle_mrs_rel": so["residual_rel"],
                      "oracle_true_rel": true_rel(xo),
                      "gpu_batched_iter": stats[k]["iter"],
                      "gpu_batched_true_rel": true_rel(x_batched[k]),
                      "x_batched_vs_oracle": float(np.linalg.norm(x_batched[k] - xo)
                                                   / np.linalg.norm(xo))}
        if k in x_single:
            oracle_cmp.update({"gpu_single_iter": fp64[k]["iter"],
                               "gpu_single_true_rel": true_rel(x_single[k]),
                               "x_single_vs_oracle": float(np.linalg.norm(x_single[k] - xo)
                                                           / np.linalg.norm(xo))})
        # Z at the 5x5 grid with the oracle's solution in place of the GPU's for this system
        # (the other polarization of the frequency from the batched GPU solve)
        xz = x_batched.copy()
        xz[k] = xo
        zo = impedance_agreement(M, m, cases, x_batched, {q: xz[q] for q in range(nsys)}, None)
        fi = [q for q, f in enumerate(zo["freqs"]) if abs(f - cases[k][0]) < 1e-12 * f][0]
        oracle_cmp["Z_batched_vs_oracle_max_rel"] = zo["batched_vs_single_fp64"][fi]
        log(f"oracle vs GPU at full size: {oracle_cmp}")

    spb = 1 << (nsys - 1).bit_length()
    traffic = None if args.no_traffic else measure_traffic(args, spb)
    log(f"spmv traffic: {traffic}")
    def kernel_line(kname, b, ms, tr):
        a = gbs(b, ms)
        t = tr["FETCH_SIZE"]["bytes"] + tr["WRITE_SIZE"]["bytes"] if tr else None
        return {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": a / HBM_PEAK_GBS, "traffic": t, "traffic_counters": tr,
                "kernel": kname, "algorithmic_bytes": b, "avg_ms": ms}

    spmv_line = kernel_line(f"{SPMV_KERNELS[tune['batched_fp64']['spmv_layout']]}"
                            f"<double,{spb},0> (batched FP64 SpMV)", spmv_bytes, spmv_ms,
                            traffic["spmv"] if traffic else None)
    # in the grid order the update pass forms w from the fused preconditioner's node help
    upd_name = f"k_xr_fused<double,{spb},true{',true' if args.order_block == -1 else ''}>"
    upd_line = kernel_line(f"{upd_name} (batched FP64 update pass)",
                           upd_bytes, upd_ms, traffic["update"] if traffic else None)
    xp_line = kernel_line(f"k_flush_xp<{spb},{xp_period}> (batched FP64 x/p pass, every "
                          f"{xp_period}th iteration)",
                          xp_bytes, xp_ms, traffic["flush"] if traffic else None)
    # the dominant kernel: both run once per iteration, the longer one
    roof = upd_line if upd_ms > spmv_ms else spmv_line
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (complex128)",
        "data": "synthetic (seeded random conductivity model)",
        "config": {"workload": f"batched FP64 COCR + folded preconditioner, {args.nx}x{args.ny}x"
                               f"{args.nz} cells (nb={p.nb}), {args.nfreq} freqs 1e-3..1e2 Hz x 2 "
                               f"pol = {nsys} systems, MRS stop at {args.eps:g}",
                   "systems": nsys, "nb": int(p.nb), "parallelism": "single GPU",
                   "order_block": args.order_block,
                   "launch": tune},
        "roofline": roof,
        "spmv_fp64_batched": spmv_line,
        "update_fp64_batched": upd_line,
        "xp_fp64_batched": xp_line,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "oracle_vs_gpu_full_size": oracle_cmp,
        "gradient_setup_s": t_grad,
        "fp64_gpu_per_system_s": fp64_per_sys,
        "speedup_vs_fp64_gpu": (fp64_per_sys / value) if fp64_per_sys else None,
        "batched_ok": batched_ok,
        "batched_iters": [q["iter"] for q in stats],
        "mixed_batched": mixed,
        "spmv_fp32vec_batched": {"achieved_gbs": ach32, "avg_ms": spmv32_ms,
                                 "algorithmic_bytes": spmv32_bytes, "frac": ach32 / HBM_PEAK_GBS},
        "spmv_fp64_single": {"achieved_gbs": ach64, "avg_ms": spmv64_ms,
                             "algorithmic_bytes": spmv64_bytes, "frac": ach64 / HBM_PEAK_GBS},
        "fp64_systems": fp64,
        "impedance_agreement": zcmp,
        "adjoint_config4": adj,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
