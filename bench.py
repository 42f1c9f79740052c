#!/usr/bin/env python3
"""Benchmark: time-to-1e-6 FP64 residual of the refined hipPCG solve on the
3D lognormal pressure Poisson 128^3, plus the SpMV / fused-Chebyshev kernel
bandwidth against the MI355X HBM roofline, the plain-FP64 GPU solve beside it,
and the same-host CPU PCG(DIC) baseline.

One step = one complete solve from psi = 0 to the FP64 L1 residual 1e-6
(refined: FP32 Chebyshev-Jacobi PCG inner iterations, constant-mode deflation,
FP64 refinement on the GPU).  Inputs are resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--workload poisson128]
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from metal4foam_amd.profsum import slot_matrix_bytes  # noqa: E402

METRIC = ("time-to-1e-6 FP64 residual & SpMV HBM GB/s (%roofline), "
          "refined vs plain FP64")
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (generator, solver, description)
    "poisson128": ("config1", "pcg",
                   "3D pressure Poisson 128^3 (2,097,152 cells, 6,242,304 faces), "
                   "lognormal face coefficients, cell 0 pinned; hipPCG refined"),
    "convdiff96": ("config2", "bicg",
                   "3D upwind convection-diffusion 96^3 (884,736 cells), asymmetric; "
                   "hipPBiCGStab refined"),
    "poisson2d100": ("config0", "pcg",
                     "2D pressure Poisson 100x100x1 (10,000 cells), unit faces; "
                     "hipPCG refined"),
    "batch64": ("config3", "batch",
                "64 independent systems, cube sides U{20..80} (8k-512k cells), 32 lognormal "
                "Poisson (hipPCG) + 32 convection-diffusion (hipPBiCGStab), 5 RHS each: "
                "320 refined solves per step"),
    "adjoint128": ("config1", "adjoint",
                   "adjoint of L = 0.5|x - x_t|^2 through the 128^3 lognormal Poisson: "
                   "forward hipPCG + transpose solve + face/diag/source gradients"),
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def numpy_residual(m, x):
    """sum|b - A x| / normFactor with plain numpy (independent of the GPU)."""
    lower = m.upper if m.lower is None else m.lower
    ax = m.diag * x
    ax += np.bincount(m.lower_addr, weights=m.upper * x[m.upper_addr], minlength=m.n)
    ax += np.bincount(m.upper_addr, weights=lower * x[m.lower_addr], minlength=m.n)
    u = m.diag + np.bincount(m.lower_addr, weights=m.upper, minlength=m.n) + np.bincount(
        m.upper_addr, weights=lower, minlength=m.n)
    return np.abs(m.source - ax).sum(), u


def spmv_bytes(n, nf, symmetric, extra_vectors=0):
    """Algorithmic bytes of one FP32 row-gather LDU SpMV: diag, x, y, the two
    start arrays, and per face the coefficient(s) and the two address arrays
    the row gather reads (every array counted once)."""
    face = 16 if symmetric else 20
    return 20 * n + 8 + face * nf + 4 * extra_vectors * n


def cheb_step_bytes(n, nf, symmetric):
    """Interior fused Chebyshev step: SpMV of z_in (diag, z_in, starts, faces)
    plus reads r, rD, w and writes w, z_out (y itself is never stored)."""
    face = 16 if symmetric else 20
    return 36 * n + 8 + face * nf


def to_slot_bytes(ldu_bytes, n, nf, symmetric, patterned_rows):
    """The same FP32 kernel on the row-slot layout: the two row-start arrays
    and the per-face arrays are replaced by the row-slot matrix (coefficient
    planes, slot bytes, offsets of explicit tiles, tile descriptors)."""
    face = 16 if symmetric else 20
    return (ldu_bytes - (8 * n + 8 + face * nf)
            + slot_matrix_bytes(n, not symmetric, 4, patterned_rows))


def cpu_baseline(m, solver, sample_iters):
    """Oracle PCG(DIC) / PBiCGStab(DILU) in FP64 on one host core.

    sample_iters == 0: the full solve to 1e-6, timed, and its solution
    returned so the GPU results can state their error against it.  Otherwise
    a bounded number of iterations is timed and scaled to the iteration count
    of the full solve (committed fixture written by the oracle); no solution.
    Returns (baseline dict, x or None)."""
    import oracle

    fpath = os.path.join(ROOT, "tests", "golden", "cpu_baseline_iterations.json")
    full = json.load(open(fpath)) if os.path.exists(fpath) else {}
    key = m.name
    max_iter = sample_iters if sample_iters > 0 else 20000
    t0 = time.perf_counter()
    if solver == "pcg":
        x, perf = oracle.pcg(m, precon=oracle.PRE_DIC, tol=1e-6, maxIter=max_iter)
    else:
        x, perf = oracle.pbicgstab(m, precon=oracle.PRE_DIC, tol=1e-6, maxIter=max_iter)
    dt = time.perf_counter() - t0
    name = "PCG(DIC)" if solver == "pcg" else "PBiCGStab(DILU)"
    done = perf.nIterations
    if perf.converged:
        est, sample = dt, (f"{name} FP64, 1 core, the full solve to 1e-6 timed "
                           f"({done} iterations, final residual {perf.finalResidual:.3e})")
    else:
        iters_full = full.get(key, {}).get("iterations") or done
        est = dt / max(done, 1) * iters_full
        sample = (f"{name} FP64, 1 core, {done} iterations timed ({dt:.2f} s) and scaled to "
                  f"the {iters_full} iterations the full solve to 1e-6 takes")
        x = None
    return {
        "value": round(1e3 * est, 3),
        "unit": "ms",
        "cores": 1,
        "kind": "port",
        "sample": sample,
        "cpu": platform.processor() or platform.machine(),
        "nproc": os.cpu_count(),
    }, x


def rel_diff(a, b):
    """max |a - b| / max |b|"""
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="poisson128", choices=sorted(WORKLOADS))
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--host-steps", type=int, default=2,
                    help="solves timed with host arrays (PCIe copies included)")
    ap.add_argument("--cpu-iters", type=int, default=0,
                    help="0: time the full CPU solve and compare solutions; N: time N "
                         "iterations and scale")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--spmv-tune", type=int, default=1,
                    help="also time the plain SpMV on the launch shape hipfoam_autotune picks")
    ap.add_argument("--cold-reps", type=int, default=10,
                    help="SpMV launches timed after a MALL-flushing copy")
    ap.add_argument("--adjoint-precision", choices=["fp32", "fp64"], default="fp32",
                    help="adjoint128: transpose solve refined (FP32 inner) or plain FP64")
    ap.add_argument("--fd-faces", type=int, default=0,
                    help="adjoint128: finite-difference check on this many faces and cells")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU: the LDU solve is not "
                          "sharded across GPUs (see DESIGN.md for how it would shard); "
                          "run with --gpus 1"}))
        sys.exit(2)

    from metal4foam_amd import hipfoam as H
    from metal4foam_amd import meshgen as mg

    gen, solver, desc = WORKLOADS[args.workload]
    if solver == "batch":
        return run_batch(args, H, mg, desc)
    if solver == "adjoint":
        return run_adjoint(args, H, mg, desc)
    m = getattr(mg, gen)()
    sym = m.lower is None
    kind = H.PCG if solver == "pcg" else H.PBICGSTAB
    log(f"workload {m.name}: n={m.n} faces={m.nfaces}")

    ctx = H.Context(0)
    log(ctx.description)
    ctx.set_topology(m)
    D = H.DeviceArray
    d_diag = D.from_host(ctx, m.diag)
    d_upper = D.from_host(ctx, m.upper)
    d_lower = None if sym else D.from_host(ctx, m.lower)
    d_src = D.from_host(ctx, m.source)
    d_zero = D.from_host(ctx, np.zeros(m.n))
    d_psi = D.from_host(ctx, np.zeros(m.n))
    ctl = H.default_controls(maxRestarts=8, maxIter=20000)

    def step(fp64=False):
        d_psi.copy_from(d_zero)
        return ctx.solve_device(kind, d_diag, d_upper, d_lower, d_psi, d_src, ctl, fp64=fp64)

    for _ in range(args.warmup):
        perf = step()
    ctx.synchronize()
    log("warmup done")

    # ---- timed: refined solve -----------------------------------------------------
    ctx.synchronize()
    t0 = time.perf_counter()
    perfs = [step() for _ in range(args.steps)]
    ctx.synchronize()
    t1 = time.perf_counter()
    ms = 1e3 * (t1 - t0) / max(args.steps, 1)
    perf = perfs[-1]
    psi = d_psi.download()
    rsum, _ = numpy_residual(m, psi)
    true_res = rsum / perf.normFactor
    log(f"refined: {ms:.3f} ms/solve, iters={perf.nIterations} restarts={perf.restarts} "
        f"res={perf.finalResidual:.3e} (numpy {true_res:.3e}) converged={perf.converged}")
    if not (perf.converged and true_res < 1e-6):
        log("ERROR: refined solve did not reach 1e-6")

    # ---- PCIe-inclusive: host arrays in and out, as the OpenFOAM plugin passes them
    host_ms = None
    if args.host_steps > 0:
        hpsi = np.zeros(m.n)
        lower_h = None if sym else m.lower
        ctx.solve(kind, m.diag, m.upper, lower_h, hpsi, m.source, ctl)
        t0 = time.perf_counter()
        for _ in range(args.host_steps):
            hpsi[:] = 0.0
            ctx.solve(kind, m.diag, m.upper, lower_h, hpsi, m.source, ctl)
        host_ms = 1e3 * (time.perf_counter() - t0) / args.host_steps
        log(f"refined with host arrays (PCIe included): {host_ms:.3f} ms/solve")

    # ---- plain FP64 comparator ----------------------------------------------------
    fp64_ms = None
    fp64_perf = None
    psi64 = None
    if args.fp64_steps > 0:
        step(fp64=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            fp64_perf = step(fp64=True)
        ctx.synchronize()
        fp64_ms = 1e3 * (time.perf_counter() - t0) / args.fp64_steps
        psi64 = d_psi.download()
        log(f"plain fp64: {fp64_ms:.3f} ms/solve, iters={fp64_perf.nIterations} "
            f"res={fp64_perf.finalResidual:.3e}")

    # ---- kernel roofline: per-launch HIP events on the library stream -------------
    dom_class = 0 if solver == "pcg" else 2
    ctx.set_kernel_timing(True, dom_class)
    step()
    nl, secs = ctx.kernel_timing()
    ctx.set_kernel_timing(False, 0)
    layout = ctx.active_schedule()
    slots = layout == 2
    patterned, _ = ctx.slot_layout()
    if solver == "pcg":
        ref_kbytes = cheb_step_bytes(m.n, m.nfaces, sym)
        kname = "k_cheb_step<float, false> (interior Chebyshev step fused with its SpMV)"
    else:
        # mean of k_bicg_py_spmv (reads r, rD, p, AyA, rA0; writes p, yA, AyA)
        # and k_bicg_sz_spmv (reads r, rD, AyA; writes sA, zA, tA)
        ref_kbytes = spmv_bytes(m.n, m.nfaces, sym, extra_vectors=5)
        kname = "k_bicg_py_spmv / k_bicg_sz_spmv (BiCGStab SpMVs fused with their vector updates)"
    kbytes = to_slot_bytes(ref_kbytes, m.n, m.nfaces, sym, patterned) if slots else ref_kbytes
    kavg = secs / max(nl, 1)
    kgbs = kbytes / kavg / 1e9 if nl else 0.0
    ref_kgbs = ref_kbytes / kavg / 1e9 if nl else 0.0
    log(f"dominant kernel {kname} ({'row-slot' if slots else 'LDU face'} layout): {nl} "
        f"launches, avg {1e6 * kavg:.2f} us, {kgbs:.0f} GB/s ({100 * kgbs / HBM_PEAK_GBS:.1f}% "
        f"of HBM peak); LDU-format equivalent {ref_kgbs:.0f} GB/s")

    # plain FP32 SpMV of the inner operator (the metric's SpMV GB/s)
    x = D.from_host(ctx, np.random.default_rng(0).standard_normal(m.n).astype(np.float32))
    y = D(ctx, (m.n,), np.float32)
    ctx.set_kernel_timing(True, 5)
    for _ in range(args.spmv_reps):
        ctx.amul_device(x, y, H.FP32)
    ns, ssecs = ctx.kernel_timing()
    ctx.set_kernel_timing(False, 0)
    ref_sbytes = spmv_bytes(m.n, m.nfaces, sym)
    sbytes = to_slot_bytes(ref_sbytes, m.n, m.nfaces, sym, patterned) if slots else ref_sbytes
    savg = ssecs / max(ns, 1)
    sgbs = sbytes / savg / 1e9 if ns else 0.0
    ref_sgbs = ref_sbytes / savg / 1e9 if ns else 0.0
    log(f"plain FP32 SpMV: avg {1e6 * savg:.2f} us, {sgbs:.0f} GB/s "
        f"(LDU-format equivalent {ref_sgbs:.0f} GB/s)")

    # the same SpMV with the 256 MB MALL (Infinity Cache) flushed before every
    # launch by a 2 GiB device-to-device copy
    cold = None
    if args.cold_reps > 0:
        big = 256 * 1024 * 1024
        fa, fb = D(ctx, (big,), np.float64), D(ctx, (big,), np.float64)
        ctx.set_kernel_timing(True, 5)
        for _ in range(args.cold_reps):
            fb.copy_from(fa)
            ctx.amul_device(x, y, H.FP32)
        nc, csecs = ctx.kernel_timing()
        ctx.set_kernel_timing(False, 0)
        fa.free()
        fb.free()
        cavg = csecs / max(nc, 1)
        cold = {"avg_launch_us": round(1e6 * cavg, 3),
                "GB_s": round(sbytes / cavg / 1e9, 1),
                "frac": round(sbytes / cavg / 1e9 / HBM_PEAK_GBS, 4),
                "reference_format_GB_s": round(ref_sbytes / cavg / 1e9, 1)}
        log(f"plain FP32 SpMV after a MALL flush: avg {1e6 * cavg:.2f} us, "
            f"{cold['GB_s']:.0f} GB/s")

    # the same SpMV on the launch shape the library's autotuner picks for it
    # (hipfoam_autotune times every SpMV schedule, row form, workgroup size,
    # XCD mapping and grid limit); the solves above ran on the automatic shape
    tuned = None
    if args.spmv_tune and args.spmv_reps > 0:
        keep = ctx.get_schedule()
        best = ctx.autotune()
        ctx.set_kernel_timing(True, 5)
        for _ in range(args.spmv_reps):
            ctx.amul_device(x, y, H.FP32)
        nt, tsecs = ctx.kernel_timing()
        ctx.set_kernel_timing(False, 0)
        ctx.set_schedule(spmvSchedule=keep.spmvSchedule, blockSize=keep.blockSize,
                         xcdRemap=keep.xcdRemap, graphChunk=keep.graphChunk,
                         gridLimit=keep.gridLimit, rowForm=keep.rowForm)
        tbytes = sbytes if (slots and best.spmvSchedule == 2) else ref_sbytes
        tavg = tsecs / max(nt, 1)
        tuned = {"schedule": {"spmvSchedule": best.spmvSchedule, "blockSize": best.blockSize,
                              "xcdRemap": best.xcdRemap, "gridLimit": best.gridLimit,
                              "rowForm": best.rowForm},
                 "bytes_per_launch": tbytes,
                 "avg_launch_us": round(1e6 * tavg, 3),
                 "GB_s": round(tbytes / tavg / 1e9, 1),
                 "frac": round(tbytes / tavg / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"plain FP32 SpMV on the autotuned shape {tuned['schedule']}: avg "
            f"{1e6 * tavg:.2f} us, {tuned['GB_s']:.0f} GB/s "
            f"({100 * tuned['frac']:.1f}% of HBM peak)")

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        # counters measured on the layout this run uses only
        same = tj.get("layout", "LDU face") == ("row-slot" if slots else "LDU face")
        if same and slots:
            same = tj.get("patterned_rows") == patterned
        if same:
            traffic = tj.get("bytes_per_launch")

    cpu = None
    errors = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu, x_cpu = cpu_baseline(m, solver, args.cpu_iters)
        log(f"cpu baseline: {cpu['value']} ms ({cpu['sample']})")
        if x_cpu is not None:
            # both solutions meet the same 1e-6 normalised residual; their
            # difference is what that tolerance leaves, not an error of either
            errors = {"refined_vs_cpu_fp64_max_rel": rel_diff(psi, x_cpu)}
            if psi64 is not None:
                errors["plain_fp64_vs_cpu_fp64_max_rel"] = rel_diff(psi64, x_cpu)
            log("solution vs CPU FP64 solve: " +
                ", ".join(f"{k} {v:.3e}" for k, v in errors.items()))

    out = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "cells": m.n,
            "faces": m.nfaces,
            "tolerance": 1e-6,
            "inner": "FP32 Chebyshev(8)-Jacobi PCG" if solver == "pcg" else
                     "FP32 Jacobi BiCGStab",
            "refinement": "FP64 on GPU",
            "schedule": ctx.get_schedule().as_dict(),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(kgbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(kgbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": kname,
            "layout": "row-slot" if slots else "LDU face",
            "patterned_rows": patterned if slots else None,
            "bytes_per_launch": kbytes,
            "reference_format_bytes": ref_kbytes,
            "reference_format_GB_s": round(ref_kgbs, 1),
            "avg_launch_us": round(1e6 * kavg, 3),
            "launches": nl,
        },
        "spmv": {
            "kernel": "k_spmv<float> (plain FP32 SpMV of the inner operator)",
            "layout": "row-slot" if slots else "LDU face",
            "bytes_per_launch": sbytes,
            "avg_launch_us": round(1e6 * savg, 3),
            "GB_s": round(sgbs, 1),
            "frac": round(sgbs / HBM_PEAK_GBS, 4),
            "reference_format_bytes": ref_sbytes,
            "reference_format_GB_s": round(ref_sgbs, 1),
            "mall_flushed": cold,
            "autotuned": tuned,
        },
        "solve": {
            "iterations": perf.nIterations,
            "restarts": perf.restarts,
            "final_residual": perf.finalResidual,
            "numpy_residual": true_res,
            "converged": bool(perf.converged and true_res < 1e-6),
            "deflated": perf.deflated,
            "polished": perf.polished,
            "host_arrays_ms": None if host_ms is None else round(host_ms, 3),
            "vs_cpu_fp64": errors,
        },
        "fp64_plain": None if fp64_perf is None else {
            "ms": round(fp64_ms, 4),
            "iterations": fp64_perf.nIterations,
            "final_residual": fp64_perf.finalResidual,
            "converged": bool(fp64_perf.converged),
            "refined_speedup": round(fp64_ms / ms, 3),
        },
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(cpu["value"] / ms, 1)
    print(json.dumps(out), flush=True)


def _common(args, value, unit, ms, desc, higher):
    return {
        "metric": METRIC,
        "value": value,
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": higher,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": desc},
    }


def run_batch(args, H, mg, desc):
    """Batched throughput: 320 refined solves (64 systems x 5 RHS) per step, after
    the batch autotuner has picked each batch's schedule and RHS group."""
    systems = mg.config3()
    nrhs = systems[0].source.shape[1]
    sym = [m for m in systems if m.lower is None]
    asym = [m for m in systems if m.lower is not None]
    log(f"batch: {len(sym)} symmetric + {len(asym)} asymmetric systems, {nrhs} RHS, "
        f"{sum(m.n for m in systems):,} cells")
    # one context per batch: each keeps the schedule its autotuner picks
    ctxs = [H.Context(0), H.Context(0)]
    ctx = ctxs[0]
    log(ctx.description)
    batches = [H.Batch(ctxs[0], H.PCG, sym, nrhs), H.Batch(ctxs[1], H.PBICGSTAB, asym, nrhs)]
    for bt in batches:
        for k, m in enumerate(bt.systems):
            bt.set_rhs(k, m.source, layout="cell_major")
    # batch autotuner (outside the timed region): SpMV schedule, row form,
    # XCD mapping and RHS group from timings of the batched SpMV
    tuned = []
    for name, bt in zip(("pcg", "bicg"), batches):
        sc, grp = bt.autotune()
        tuned.append({"batch": name, "spmvSchedule": sc.spmvSchedule, "rowForm": sc.rowForm,
                      "xcdRemap": sc.xcdRemap, "rhsGroup": grp})
        log(f"batch autotune ({name}): schedule {sc.spmvSchedule}, rowForm {sc.rowForm}, "
            f"xcdRemap {sc.xcdRemap}, RHS group {grp}")
    ctl = H.default_controls(maxRestarts=8, maxIter=20000)

    def step():
        out = []
        for bt in batches:
            bt.reset_psi()
            out.append(bt.solve(ctl))
        return out

    for _ in range(args.warmup):
        step()
    for c in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        perfs = step()
    for c in ctxs:
        c.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / max(args.steps, 1)
    nsolves = nrhs * len(systems)
    flat = [p for ps in perfs for p in ps]
    conv = sum(1 for p in flat if p.converged)
    # independent check of every segment's FP64 residual with numpy
    worst = 0.0
    q = 0
    for bt, ps in zip(batches, perfs):
        for k, m in enumerate(bt.systems):
            psi = bt.psi(k)
            for r in range(nrhs):
                mm = mg.LduMatrix(m.n, m.lower_addr, m.upper_addr, m.owner_start, m.losort,
                                  m.losort_start, m.diag, m.upper, m.lower,
                                  m.source[:, r].copy())
                rs, _ = numpy_residual(mm, psi[r])
                worst = max(worst, rs / ps[k * nrhs + r].normFactor)
                q += 1
    iters = [p.nIterations for p in flat]
    log(f"batch: {ms:.2f} ms per step, {nsolves / (ms / 1e3):.1f} solves/s, converged "
        f"{conv}/{nsolves}, inner iterations min/median/max {min(iters)}/"
        f"{int(np.median(iters))}/{max(iters)}, worst numpy residual {worst:.2e}")
    out = _common(args, round(nsolves / (ms / 1e3), 2), "solves/s", ms, desc, True)
    out["config"].update({"systems": len(systems), "rhs": nrhs,
                          "cells": int(sum(m.n for m in systems)), "tolerance": 1e-6})
    out["batch"] = {"solves": nsolves, "converged": conv, "worst_numpy_residual": worst,
                    "iterations_min_median_max": [min(iters), int(np.median(iters)), max(iters)],
                    "autotuned": tuned}
    out["roofline"] = None
    out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def run_adjoint(args, H, mg, desc):
    """Forward + adjoint time on the 128^3 Poisson, with a finite-difference check
    of 32 face and 32 cell gradients (h = 1e-6, tight plain-FP64 GPU solves)."""
    m = mg.config1()
    xt = mg.config4_target(m.n)
    ctx = H.Context(0)
    log(ctx.description)
    ctx.set_topology(m)
    ctl = H.default_controls(maxRestarts=8, maxIter=20000)
    # g = x - x_t has a non-zero sum: on the pinned Poisson the reference's
    # zero-mean deflation stalls on such a right-hand side, so the refined
    # transpose solve projects its inner right-hand side along u = A 1
    # (deflate = 3); --adjoint-precision fp64 solves it in plain FP64 instead
    ctl_adj = H.default_controls(maxRestarts=8, maxIter=20000, deflate=3)
    adj_prec = H.FP64 if args.adjoint_precision == "fp64" else H.FP32

    def fwd_bwd():
        x = np.zeros(m.n)
        pf = ctx.solve(H.PCG, m.diag, m.upper, None, x, m.source, ctl)
        pb = ctx.adjoint(H.PCG, m.diag, m.upper, None, x, x - xt,
                         ctl_adj if adj_prec == H.FP32 else ctl, precision=adj_prec)
        return x, pf, pb

    for _ in range(args.warmup):
        fwd_bwd()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x, pf, (pb, lam, du, _dl, dd, db) = fwd_bwd()
    ms = 1e3 * (time.perf_counter() - t0) / max(args.steps, 1)
    log(f"adjoint: forward+backward {ms:.1f} ms (host arrays), forward it={pf.nIterations} "
        f"res={pf.finalResidual:.2e}, adjoint ({args.adjoint_precision}) it={pb.nIterations} "
        f"res={pb.finalResidual:.2e}")
    # the same gradients from a plain-FP64 transpose solve at the same tolerance
    p64, lam64, du64, _dl64, dd64, db64 = ctx.adjoint(H.PCG, m.diag, m.upper, None, x, x - xt,
                                                       ctl, precision=H.FP64)
    grad_diff = max(float(np.abs(a - b).max() / np.abs(b).max())
                    for a, b in ((du, du64), (dd, dd64), (db, db64)))
    log(f"adjoint gradients vs FP64 transpose solve: max relative difference {grad_diff:.2e}")
    fd = None
    if args.fd_faces > 0:
        rng = np.random.default_rng(1004)
        tight = H.default_controls(tolerance=1e-14, maxIter=20000)
        # gradients of the tightly converged solution
        x = np.zeros(m.n)
        ctx.solve(H.PCG, m.diag, m.upper, None, x, m.source, tight, fp64=True)
        pa, lam, du, _dl, dd, db = ctx.adjoint(H.PCG, m.diag, m.upper, None, x, x - xt,
                                               tight, precision=H.FP64)
        log(f"adjoint (FP64 transpose solve, tol 1e-14): it={pa.nIterations} "
            f"res={pa.finalResidual:.2e}")
        h = 1e-6

        def solve(d, u):
            y = np.zeros(m.n)
            ctx.solve(H.PCG, d, u, None, y, m.source, tight, fp64=True)
            return y

        def dloss(yp, ym):
            # L(+h) - L(-h) = 0.5 sum (yp - ym)(yp + ym - 2 x_t): no cancellation of L
            return 0.5 * np.dot(yp - ym, yp + ym - 2.0 * xt)

        worst = 0.0
        faces = rng.choice(m.nfaces, args.fd_faces, replace=False)
        cells = rng.choice(m.n, args.fd_faces, replace=False)
        scale = max(np.abs(du).max(), np.abs(dd).max())
        for f in faces:
            u1, u2 = m.upper.copy(), m.upper.copy()
            u1[f] += h
            u2[f] -= h
            g = dloss(solve(m.diag, u1), solve(m.diag, u2)) / (2 * h)
            worst = max(worst, abs(g - du[f]) / max(abs(du[f]), 1e-3 * scale))
        for i in cells:
            d1, d2 = m.diag.copy(), m.diag.copy()
            d1[i] += h
            d2[i] -= h
            g = dloss(solve(d1, m.upper), solve(d2, m.upper)) / (2 * h)
            worst = max(worst, abs(g - dd[i]) / max(abs(dd[i]), 1e-3 * scale))
        fd = {"faces": int(args.fd_faces), "cells": int(args.fd_faces), "h": h,
              "max_rel_err": worst}
        log(f"adjoint FD check: max relative error {worst:.2e} over {args.fd_faces} faces and "
            f"{args.fd_faces} cells")
    out = _common(args, round(ms, 3), "ms", ms, desc, False)
    out["config"].update({"cells": m.n, "faces": m.nfaces, "tolerance": 1e-6})
    out["adjoint"] = {"forward_iterations": pf.nIterations, "adjoint_iterations": pb.nIterations,
                      "adjoint_precision": ("fp32 refined, deflate 3 (pinned projection)"
                                            if adj_prec == H.FP32 else "fp64"),
                      "adjoint_residual": pb.finalResidual,
                      "gradients_vs_fp64_adjoint_max_rel_diff": grad_diff,
                      "adjoint_converged": bool(pb.converged),
                      "fd_check": fd}
    out["roofline"] = None
    out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
