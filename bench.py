#!/usr/bin/env python3
"""Benchmark: time-to-1e-6 FP64 residual of the refined hipPCG solve on the
3D lognormal pressure Poisson 128^3, plus the SpMV / fused-Chebyshev kernel
bandwidth against the MI355X HBM roofline, the plain-FP64 GPU solve beside it,
and the same-host CPU PCG(DIC) baseline.

One step = one complete solve from psi = 0 to the FP64 L1 residual 1e-6
(refined: FP32 Chebyshev-Jacobi PCG inner iterations, constant-mode deflation,
FP64 refinement on the GPU).  Inputs are resident in HBM before timing starts.
The line also carries small meshes (10k-cell Poisson and convection-diffusion,
24^3) with the inner loop as one persistent launch against the launched loop,
and the 2D Poisson at the reference benchmark's cavity sizes (400 to 640,000
cells) beside the reference's published per-call times and the same-host CPU
PCG(DIC).

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
    "poisson256": ("config1", "pcg",
                   "3D pressure Poisson 256^3 (16,777,216 cells), lognormal face "
                   "coefficients, cell 0 pinned; hipPCG refined (past the MALL)"),
    "adjoint128": ("config1", "adjoint",
                   "adjoint of L = 0.5|x - x_t|^2 through the 128^3 lognormal Poisson: "
                   "forward hipPCG + transpose solve + face/diag/source gradients"),
}


WORKLOAD_ARGS = {"poisson256": (256,)}


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


def _omp_threads():
    """Threads for the all-cores CPU line: OMP_NUM_THREADS when set (16 on
    the MI355X boxes, their CPU share), else every core."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    return int(v) if v.isdigit() and int(v) > 0 else (os.cpu_count() or 1)


def _relmax(a, b):
    """relMax of two fields: max |a - b| / max |b| (the measure of the
    reference's compare_fields.py:38-53, with b the reference field)."""
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


CPU_RUNS = 3  # each CPU time is the median of this many runs


_NATIVE = []


def _native_oracle(oracle):
    """The CPU solvers built once on this host with -O3 -march=native (None when
    that build fails; the prebuilt -march=x86-64-v2 build then runs)."""
    import tempfile

    if not _NATIVE:
        oracle.lib()
        _NATIVE.append(oracle.build_native(tempfile.mkdtemp(prefix="hipfoam_cpu_")))
    return _NATIVE[0]


def cpu_baseline(m, solver, sample_iters):
    """CPU solves on this host, in FP64, built here with -O3 -march=native:

    * the baseline proper: OpenFOAM PCG(DIC) / PBiCGStab(DILU), 1 core (DIC and
      DILU sweeps are sequential; OpenFOAM runs serially here, as the
      reference's benchmark does);
    * PCG(diagonal) on all host cores (OpenMP, row-gather SpMV), PCG only;
    * the CPU-vs-CPU floor: how far the diagonal-preconditioned CPU solve
      lands from the DIC (DILU) one at the same 1e-6 tolerance.

    Each time is the median of CPU_RUNS runs.  sample_iters == 0: every
    solve runs to 1e-6 and is timed whole, and the
    DIC (DILU) solution is returned so the GPU results can state their error
    against it.  Otherwise a bounded number of iterations of the baseline is
    timed and scaled to the iteration count of the full solve (committed
    fixture written by the oracle); no solution, no floor.
    Returns (baseline dict, x or None, floor dict or None)."""
    import oracle

    native = _native_oracle(oracle)
    prev = oracle.swap_library(native) if native is not None else None
    build = ("gcc -O3 -march=native -fopenmp, built on this host" if native is not None else
             "the prebuilt oracle (-O3 -march=x86-64-v2): the native build failed")
    try:
        fpath = os.path.join(ROOT, "tests", "golden", "cpu_baseline_iterations.json")
        full = json.load(open(fpath)) if os.path.exists(fpath) else {}
        max_iter = sample_iters if sample_iters > 0 else 20000
        runs = []
        for _ in range(CPU_RUNS):
            t0 = time.perf_counter()
            if solver == "pcg":
                x, perf = oracle.pcg(m, precon=oracle.PRE_DIC, tol=1e-6, maxIter=max_iter)
            else:
                x, perf = oracle.pbicgstab(m, precon=oracle.PRE_DIC, tol=1e-6,
                                           maxIter=max_iter)
            runs.append(time.perf_counter() - t0)
        dt = float(np.median(runs))
        name = "PCG(DIC)" if solver == "pcg" else "PBiCGStab(DILU)"
        done = perf.nIterations
        if perf.converged:
            est, sample = dt, (f"{name} FP64, 1 core, the full solve to 1e-6 timed, median "
                               f"of {CPU_RUNS} runs ({done} iterations, final residual "
                               f"{perf.finalResidual:.3e})")
        else:
            iters_full = full.get(m.name, {}).get("iterations") or done
            est = dt / max(done, 1) * iters_full
            sample = (f"{name} FP64, 1 core, {done} iterations timed ({dt:.2f} s) and scaled "
                      f"to the {iters_full} iterations the full solve to 1e-6 takes")
            x = None
        out = {
            "value": round(1e3 * est, 3),
            "unit": "ms",
            "cores": 1,
            "kind": "port",
            "sample": sample,
            "build": build,
            "cpu": _cpu_model(),
            "nproc": os.cpu_count(),
        }
        floor = None
        if x is not None:
            if solver == "pcg":
                th = _omp_threads()
                druns = []
                for _ in range(CPU_RUNS):
                    t0 = time.perf_counter()
                    xd, pd = oracle.pcg_diag_par(m, tol=1e-6, maxIter=20000, threads=th)
                    druns.append(time.perf_counter() - t0)
                out["all_cores"] = {
                    "value": round(1e3 * float(np.median(druns)), 3), "unit": "ms",
                    "cores": th if native is not None else 1,
                    "solver": "PCG(diagonal) FP64, OpenMP row-gather SpMV",
                    "runs": CPU_RUNS,
                    "iterations": pd.nIterations, "converged": bool(pd.converged)}
                other = "PCG(diagonal)"
            else:
                xd, pd = oracle.pbicgstab(m, precon=oracle.PRE_DIAGONAL, tol=1e-6,
                                          maxIter=20000)
                other = "PBiCGStab(diagonal)"
            if pd.converged:
                floor = {"cpu_vs_cpu_floor_relmax": _relmax(xd, x),
                         "compared": f"{other} vs {name}, both FP64 CPU to 1e-6"}
        return out, x, floor
    finally:
        if prev is not None:
            oracle.swap_library(prev)


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def rel_diff(a, b):
    """max |a - b| / max |b|"""
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _traffic(tag, kernel_layout, patterned):
    """Calibrated HBM bytes per launch from profiles/traffic_<tag>.json, when
    those counters were taken on the layout this run uses."""
    tfile = os.path.join(ROOT, "profiles", f"traffic_{tag}.json")
    if not os.path.exists(tfile):
        return None, None
    tj = json.load(open(tfile))
    if tj.get("layout", "LDU face") != kernel_layout:
        return None, None
    if kernel_layout == "row-slot" and tj.get("patterned_rows") != patterned:
        return None, None
    return tj.get("bytes_per_launch"), os.path.relpath(tfile, ROOT)


def _entry(kernel, layout, nbytes, ref_bytes, launches, secs, traffic=None, tfile=None):
    avg = secs / max(launches, 1)
    gbs = nbytes / avg / 1e9 if launches else 0.0
    return {
        "kernel": kernel,
        "layout": layout,
        "bytes_per_launch": int(nbytes),
        "reference_format_bytes": int(ref_bytes),
        "launches": int(launches),
        "avg_launch_us": round(1e6 * avg, 3),
        "GB_s": round(gbs, 1),
        "frac": round(gbs / HBM_PEAK_GBS, 4),
        "reference_format_GB_s": round(ref_bytes / avg / 1e9, 1) if launches else 0.0,
        "traffic": traffic,
        "traffic_frac_of_algorithmic": (round(traffic / nbytes, 3) if traffic else None),
        "traffic_source": tfile,
    }


def measure_kernels(H, ctx, m, solver, spmv_reps, tag, solve_fn):
    """Dominant solve kernel and plain FP32 SpMV of one mesh on the context's
    current schedule, timed with per-launch HIP events on the library stream.

    PCG: the interior Chebyshev step k_cheb_step<float, false> (class 0);
    BiCGStab: the two SpMVs fused with their vector updates (class 2).
    Bytes are the kernel's own algorithmic bytes on the layout it runs (the
    row-slot layout, or the LDU face arrays, priced on the reference's own
    format, when the schedule is 0 or the mesh falls back to it)."""
    sym = m.lower is None
    D = H.DeviceArray
    ctx.set_kernel_timing(True, 0 if solver == "pcg" else 2)
    perf = solve_fn()
    nl, secs = ctx.kernel_timing()
    ctx.set_kernel_timing(False, 0)
    slots = ctx.active_schedule() == 2
    patterned, _ = ctx.slot_layout()
    layout = "row-slot" if slots else "LDU face"
    if solver == "pcg":
        ref_kbytes = cheb_step_bytes(m.n, m.nfaces, sym)
        kname = "k_cheb_step<float, false> (interior Chebyshev step fused with its SpMV)"
    else:
        # mean of k_bicg_py_spmv (reads r, rD, p, AyA, rA0; writes p, yA, AyA)
        # and k_bicg_sz_spmv (reads r, rD, AyA; writes sA, zA, tA)
        ref_kbytes = spmv_bytes(m.n, m.nfaces, sym, extra_vectors=5)
        kname = ("k_bicg_py_spmv / k_bicg_sz_spmv (BiCGStab SpMVs fused with their "
                 "vector updates)")
    kbytes = to_slot_bytes(ref_kbytes, m.n, m.nfaces, sym, patterned) if slots else ref_kbytes
    ktraf, ktfile = _traffic(tag, layout, patterned)
    solve_k = _entry(kname, layout, kbytes, ref_kbytes, nl, secs, ktraf, ktfile)
    solve_k["patterned_rows"] = patterned if slots else None
    spmv = None
    if spmv_reps > 0:
        x = D.from_host(ctx, np.random.default_rng(0).standard_normal(m.n).astype(np.float32))
        y = D(ctx, (m.n,), np.float32)
        ctx.set_kernel_timing(True, 5)
        for _ in range(spmv_reps):
            ctx.amul_device(x, y, H.FP32)
        ns, ssecs = ctx.kernel_timing()
        ctx.set_kernel_timing(False, 0)
        ref_sbytes = spmv_bytes(m.n, m.nfaces, sym)
        sbytes = to_slot_bytes(ref_sbytes, m.n, m.nfaces, sym, patterned) if slots else \
            ref_sbytes
        straf, stfile = _traffic(tag + "_spmv", layout, patterned)
        spmv = _entry("k_spmv<float> (plain FP32 SpMV of the inner operator)", layout, sbytes,
                      ref_sbytes, ns, ssecs, straf, stfile)
        x.free()
        y.free()
    return perf, solve_k, spmv


def extra_mesh(H, mg, args, gen_name, gen_args, solver, tag, what, schedule=None):
    """Kernel rooflines (and the time to 1e-6) of another mesh, on a context
    of its own: the past-MALL mesh, the forced LDU face schedule, the
    convection-diffusion BiCGStab kernels."""
    m = gen_name(mg) if callable(gen_name) else getattr(mg, gen_name)(*gen_args)
    ctx = H.Context(0)
    ctx.set_topology(m)
    if schedule is not None:
        ctx.set_schedule(spmvSchedule=schedule)
    sym = m.lower is None
    kind = H.PCG if solver == "pcg" else H.PBICGSTAB
    D = H.DeviceArray
    d = [D.from_host(ctx, a) for a in (m.diag, m.upper)]
    lo = None if sym else D.from_host(ctx, m.lower)
    src = D.from_host(ctx, m.source)
    psi = D.from_host(ctx, np.zeros(m.n))
    zero = D.from_host(ctx, np.zeros(m.n))
    ctl = H.default_controls(maxRestarts=8, maxIter=20000)

    def solve():
        psi.copy_from(zero)
        return ctx.solve_device(kind, d[0], d[1], lo, psi, src, ctl)

    solve()
    ctx.synchronize()
    t0 = time.perf_counter()
    p = solve()
    ctx.synchronize()
    ms = 1e3 * (time.perf_counter() - t0)
    _, solve_k, spmv = measure_kernels(H, ctx, m, solver, args.spmv_reps, tag, solve)
    rsum, _ = numpy_residual(m, psi.download())
    pat, expl, frows = ctx.slot_tiles()
    out = {"what": what, "mesh": m.name, "cells": m.n, "faces": m.nfaces,
           "schedule": ctx.get_schedule().as_dict(),
           "active_spmv_schedule": ctx.active_schedule(),
           "rows_patterned_explicit_facetile": [pat, expl, frows],
           "solve_ms": round(ms, 3), "iterations": p.nIterations,
           "converged": bool(p.converged and rsum / p.normFactor < 1e-6),
           "solve_kernel": solve_k, "spmv": spmv}
    log(f"{tag}: {m.name} solve {ms:.2f} ms ({p.nIterations} it.); {solve_k['kernel'][:24]} "
        f"{solve_k['avg_launch_us']} us {100 * solve_k['frac']:.1f}% ({solve_k['layout']}); "
        f"SpMV {spmv['avg_launch_us'] if spmv else '-'} us "
        f"{100 * spmv['frac'] if spmv else 0:.1f}%")
    for a in (*d, lo, src, psi, zero):
        if a is not None:
            a.free()
    ctx.close()
    return out


def small_meshes(H, mg, reps=20):
    """Refined solves of small meshes, where a launched kernel costs about as
    much as the work it does: the whole inner loop as one persistent launch
    on one XCD (the automatic schedule) against the launched loop (persistent
    0), median of `reps` device-resident solves each."""
    cases = [
        ("poisson2d100", mg.config0(), H.PCG,
         "2D pressure Poisson 100x100x1 (10,000 cells), hipPCG refined"),
        ("convdiff2d100", mg.convection_diffusion(100, 100, 1, seed=1002), H.PBICGSTAB,
         "2D upwind convection-diffusion 100x100x1 (10,000 cells), hipPBiCGStab refined"),
        ("poisson24", mg.poisson(24, 24, 24, seed=1001, lognormal=True), H.PCG,
         "3D lognormal Poisson 24^3 (13,824 cells), hipPCG refined"),
    ]
    out = {}
    for tag, m, kind, what in cases:
        ctx = H.Context(0)
        ctx.set_topology(m)
        D = H.DeviceArray
        arrs = [D.from_host(ctx, a) for a in (m.diag, m.upper)]
        lo = None if m.lower is None else D.from_host(ctx, m.lower)
        src = D.from_host(ctx, m.source)
        psi = D.from_host(ctx, np.zeros(m.n))
        zero = D.from_host(ctx, np.zeros(m.n))
        ctl = H.default_controls(maxRestarts=8, maxIter=20000)
        res = {}
        for name, p in (("persistent", -1), ("launched", 0)):
            ctx.set_schedule(persistent=p)
            ts = []
            for k in range(reps + 2):
                psi.copy_from(zero)
                ctx.synchronize()
                t0 = time.perf_counter()
                perf = ctx.solve_device(kind, arrs[0], arrs[1], lo, psi, src, ctl)
                ctx.synchronize()
                ts.append(time.perf_counter() - t0)
            rsum, _ = numpy_residual(m, psi.download())
            res[name] = (1e3 * float(np.median(ts[2:])), perf,
                         bool(perf.converged and rsum / perf.normFactor < 1e-6))
        launches, abandoned = ctx.persistent_stats()
        tp, pp, cp = res["persistent"]
        tl, pl, cl = res["launched"]
        out[tag] = {"what": what, "cells": m.n, "solve_ms": round(tp, 4),
                    "launched_loop_ms": round(tl, 4), "speedup": round(tl / tp, 2),
                    "iterations": pp.nIterations, "restarts": pp.restarts,
                    "converged": cp and cl, "same_iterations": pp.nIterations == pl.nIterations,
                    "persistent_launches": launches, "abandoned": abandoned}
        log(f"small mesh {tag}: persistent {tp:.3f} ms, launched loop {tl:.3f} ms "
            f"({pp.nIterations} it., {launches} persistent launches, {abandoned} abandoned)")
        for a in (*arrs, lo, src, psi, zero):
            if a is not None:
                a.free()
        ctx.close()
    return out


# The reference's published pressure-solve table (BENCHMARK.md:59-69):
# icoFoam 2D lid-driven cavity, fvMatrix::solve.p total / calls over 20 time
# steps (40 pressure solves, warm-started from the previous field), OpenFOAM
# v2506 on an Apple M3 Pro.  (side, cells, CPU PCG+DIC ms, metalPCG ms).
REFERENCE_CAVITY_P = (
    (20, 400, 0.066, 1.109),
    (50, 2500, 2.082, 5.838),
    (70, 4900, 6.434, 8.331),
    (100, 10000, 20.52, 11.63),
    (200, 40000, 173.8, 34.40),
    (400, 160000, 1500.8, 249.7),
    (800, 640000, 10508.3, 1787.8),
)


# The reference's momentum table (BENCHMARK.md:112-119): icoFoam cavity 200x200
# (40,000 cells), fvMatrix::solve.U total / calls, Ux and Uy per call, Apple
# M3 Pro: CPU PBiCGStab with the diagonal preconditioner, and metalPBiCGStab.
REFERENCE_CAVITY_U = {"cells": 40000, "cpu_pbicgstab_diagonal_ms": 11.16,
                      "metalpbicgstab_ms": 13.75}


def momentum_matrix(mg):
    """Stand-in for the cavity's 200x200 momentum matrix: upwind
    convection-diffusion (umax 2) with a transient diagonal term (shift 1) that
    makes PBiCGStab(diagonal) converge in 15 iterations, as the benchmark's
    momentum matrices converge in about 15; two N(0,1) right-hand sides, the
    Ux and Uy components of one call."""
    return mg.convection_diffusion(200, 200, 1, seed=1005, nrhs=2, shift=1.0, umax=2.0,
                                   name="momentum2d_200x200")


def reference_momentum(H, mg, reps=20):
    """One momentum call of the reference benchmark as a plugin makes it: two
    refined hipPBiCGStab solves (Ux, Uy) of the same matrix on host arrays,
    each with hipfoam_set_topology's content check, from psi = 0 to 1e-6;
    median of `reps` calls."""
    m = momentum_matrix(mg)
    comps = [np.ascontiguousarray(m.source[:, k]) for k in range(2)]
    ctx = H.Context(0)
    ctl = H.default_controls(maxRestarts=8, maxIter=20000)
    ts = []
    for _ in range(reps + 2):
        psis = [np.zeros(m.n), np.zeros(m.n)]
        t0 = time.perf_counter()
        perfs = []
        for k in range(2):
            ctx.set_topology(m)
            perfs.append(ctx.solve(H.PBICGSTAB, m.diag, m.upper, m.lower, psis[k], comps[k],
                                   ctl))
        ts.append(time.perf_counter() - t0)
    conv = True
    for k in range(2):
        mk = m
        mk.source = comps[k]
        rsum, _ = numpy_residual(mk, psis[k])
        conv = conv and bool(perfs[k].converged and rsum / perfs[k].normFactor < 1e-6)
    ms = 1e3 * float(np.median(ts[2:]))
    ref = REFERENCE_CAVITY_U
    out = {"what": "2D upwind convection-diffusion 200x200x1 with a transient diagonal term, "
                   "Ux and Uy per call, refined hipPBiCGStab on host arrays (the benchmark's "
                   "cavity momentum call)",
           "cells": m.n, "call_ms": round(ms, 4),
           "iterations": [p.nIterations for p in perfs],
           "restarts": [p.restarts for p in perfs], "converged": conv,
           "reference_cpu_pbicgstab_diagonal_ms": ref["cpu_pbicgstab_diagonal_ms"],
           "reference_metalpbicgstab_ms": ref["metalpbicgstab_ms"],
           "vs_reference_metalpbicgstab": round(ref["metalpbicgstab_ms"] / ms, 1)}
    log(f"reference momentum call: {ms:.3f} ms (iterations {out['iterations']}); reference "
        f"metalPBiCGStab {ref['metalpbicgstab_ms']} ms")
    ctx.close()
    return out


def cpu_reference_momentum(mg):
    """The CPU side of the momentum call: OpenFOAM PBiCGStab with the diagonal
    preconditioner (the reference's CPU solver for it), FP64, 1 core, Ux and Uy
    (median of CPU_RUNS calls)."""
    import oracle

    native = _native_oracle(oracle)
    prev = oracle.swap_library(native) if native is not None else None
    try:
        m = momentum_matrix(mg)
        comps = [np.ascontiguousarray(m.source[:, k]) for k in range(2)]
        runs, its = [], []
        for _ in range(CPU_RUNS):
            t0 = time.perf_counter()
            its = []
            for k in range(2):
                x, perf = oracle.pbicgstab(m, precon=oracle.PRE_DIAGONAL, tol=1e-6,
                                           maxIter=20000, source=comps[k])
                its.append(perf.nIterations)
            runs.append(time.perf_counter() - t0)
        return {"cpu_pbicgstab_diagonal_ms": round(1e3 * float(np.median(runs)), 4),
                "cpu_iterations": its, "cpu_runs": len(runs)}
    finally:
        if prev is not None:
            oracle.swap_library(prev)


def reference_size_meshes(mg, sides):
    """2D pressure Poisson side x side x 1 (unit faces, cell 0 pinned, zero-mean
    U[-1,1] right-hand side) at the reference benchmark's mesh sizes."""
    return [mg.poisson(n, n, 1, seed=1000, lognormal=False, name=f"poisson2d_{n}x{n}")
            for n in sides]


def reference_sizes(H, mg, sides, reps=20):
    """Refined hipPCG at the reference benchmark's cavity sizes, each solve from
    psi = 0 to the FP64 L1 residual 1e-6 (a cold start: at least the work of
    the benchmark's warm-started calls).  Two times per mesh, medians of
    `reps` solves (5 from 160,000 cells): device arrays (the solve alone) and
    one plugin call on host arrays (hipfoam_set_topology's content check,
    staging, copies both ways and the solve), set beside the reference's
    published per-call times."""
    ref = {r[0]: r for r in REFERENCE_CAVITY_P}
    out, solutions = {}, {}
    ctx = H.Context(0)
    for m, side in zip(reference_size_meshes(mg, sides), sides):
        ctx.set_topology(m)
        D = H.DeviceArray
        d, u = D.from_host(ctx, m.diag), D.from_host(ctx, m.upper)
        b = D.from_host(ctx, m.source)
        psi = D.from_host(ctx, np.zeros(m.n))
        zero = D.from_host(ctx, np.zeros(m.n))
        ctl = H.default_controls(maxRestarts=8, maxIter=20000)
        k = reps if m.n < 160000 else 5
        ts = []
        for _ in range(k + 2):
            psi.copy_from(zero)
            ctx.synchronize()
            t0 = time.perf_counter()
            perf = ctx.solve_device(H.PCG, d, u, None, psi, b, ctl)
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        x_dev = psi.download()
        rsum, _ = numpy_residual(m, x_dev)
        th = []
        hp = np.zeros(m.n)
        for _ in range(k + 2):
            hp[:] = 0.0
            t0 = time.perf_counter()
            ctx.set_topology(m)
            ph = ctx.solve(H.PCG, m.diag, m.upper, None, hp, m.source, ctl)
            th.append(time.perf_counter() - t0)
        dev = 1e3 * float(np.median(ts[2:]))
        host = 1e3 * float(np.median(th[2:]))
        e = {"cells": m.n, "solve_ms": round(dev, 4), "host_arrays_ms": round(host, 4),
             "iterations": perf.nIterations, "restarts": perf.restarts,
             "converged": bool(perf.converged and ph.converged and rsum / perf.normFactor < 1e-6)}
        if side in ref:
            _, _, cpu_ms, gpu_ms = ref[side]
            e.update({"reference_cpu_dic_ms": cpu_ms, "reference_metalpcg_ms": gpu_ms,
                      "vs_reference_metalpcg": round(gpu_ms / host, 1)})
        out[f"{side}x{side}"] = e
        solutions[f"{side}x{side}"] = x_dev
        log(f"reference size {side}x{side}: {dev:.3f} ms device arrays, {host:.3f} ms host "
            f"arrays ({perf.nIterations} it.); reference metalPCG "
            f"{ref[side][3] if side in ref else '-'} ms")
        for a in (d, u, b, psi, zero):
            a.free()
    ctx.close()
    return out, solutions


def cpu_reference_sizes(mg, sides, solutions=None):
    """The CPU baseline at the reference benchmark's sizes: OpenFOAM PCG(DIC),
    FP64, 1 core (median of CPU_RUNS runs; one run from 640,000 cells).  With
    the GPU solutions (name -> psi), each entry also states the GPU
    solution's relMax against the CPU one, beside the CPU-vs-CPU floor:
    PCG(diagonal) against PCG(DIC), both to 1e-6, as the reference states its
    accuracy (BENCHMARK.md:23-40)."""
    import oracle

    native = _native_oracle(oracle)
    prev = oracle.swap_library(native) if native is not None else None
    try:
        out = {}
        for m, side in zip(reference_size_meshes(mg, sides), sides):
            runs = []
            for _ in range(CPU_RUNS if m.n < 640000 else 1):
                t0 = time.perf_counter()
                x, perf = oracle.pcg(m, precon=oracle.PRE_DIC, tol=1e-6, maxIter=20000)
                runs.append(time.perf_counter() - t0)
            e = {"cpu_dic_ms": round(1e3 * float(np.median(runs)), 4),
                 "cpu_iterations": perf.nIterations, "cpu_converged": bool(perf.converged),
                 "cpu_runs": len(runs)}
            tag = f"{side}x{side}"
            if solutions is not None and tag in solutions:
                e["gpu_vs_cpu_relmax"] = _relmax(solutions[tag], x)
                if m.n <= 160000:
                    xd, pd = oracle.pcg(m, precon=oracle.PRE_DIAGONAL, tol=1e-6, maxIter=20000)
                    if pd.converged:
                        e["cpu_vs_cpu_floor_relmax"] = _relmax(xd, x)
            out[tag] = e
        return out
    finally:
        if prev is not None:
            oracle.swap_library(prev)


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
                    help="0: time the full CPU solves and compare solutions; N: time N "
                         "iterations of the baseline and scale")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--spmv-tune", type=int, default=1,
                    help="also time the plain SpMV on the launch shape hipfoam_autotune picks")
    ap.add_argument("--spmv-schedule", type=int, default=2, choices=[0, 1, 2],
                    help="SpMV schedule of the solve (2: row slots where the mesh allows "
                         "them, the default; 0: LDU face arrays)")
    ap.add_argument("--cold-reps", type=int, default=10,
                    help="SpMV launches timed after a MALL-flushing copy")
    ap.add_argument("--extras", type=int, default=1,
                    help="poisson128: also measure the past-MALL 256^3 mesh, the LDU face "
                         "schedule, the 96^3 BiCGStab kernels, renumbered meshes, the small "
                         "meshes and the reference's cavity sizes (--ref-sizes)")
    ap.add_argument("--small-meshes", type=int, default=1,
                    help="poisson128 / poisson2d100 with --extras: the small-mesh section "
                         "(persistent against launched loops)")
    ap.add_argument("--ref-sizes", default="20,50,70,100,200,400,800",
                    help="poisson128 / poisson2d100 with --extras: 2D Poisson sides of the "
                         "reference benchmark's cavity table to solve (empty: none)")
    ap.add_argument("--ref-momentum", type=int, default=1,
                    help="poisson128 / poisson2d100 with --extras: the reference benchmark's "
                         "momentum call (two BiCGStab components on 40,000 cells)")
    ap.add_argument("--adjoint-precision", choices=["fp32", "fp64"], default="fp32",
                    help="adjoint128: transpose solve refined (FP32 inner) or plain FP64")
    ap.add_argument("--fd-faces", type=int, default=0,
                    help="adjoint128: finite-difference check on this many faces and cells")
    ap.add_argument("--batch-rhs", choices=["as-specified", "zero-mean"],
                    default="as-specified",
                    help="batch64: Poisson right-hand sides as drawn (N(0,1), solved with "
                         "the pinned projection, deflate 3) or projected to zero mean "
                         "(solved with the reference's deflation)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU: the LDU solve is not "
                          "sharded across GPUs; run with --gpus 1"}))
        sys.exit(2)

    from metal4foam_amd import hipfoam as H
    from metal4foam_amd import meshgen as mg

    gen, solver, desc = WORKLOADS[args.workload]
    if solver == "batch":
        return run_batch(args, H, mg, desc)
    if solver == "adjoint":
        return run_adjoint(args, H, mg, desc)
    gen_args = WORKLOAD_ARGS.get(args.workload, ())
    m = getattr(mg, gen)(*gen_args)
    sym = m.lower is None
    kind = H.PCG if solver == "pcg" else H.PBICGSTAB
    log(f"workload {m.name}: n={m.n} faces={m.nfaces}")

    ctx = H.Context(0)
    log(ctx.description)
    ctx.set_topology(m)
    if args.spmv_schedule != 2:
        ctx.set_schedule(spmvSchedule=args.spmv_schedule)
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

    # ---- PCIe-inclusive: host arrays in and out, as the OpenFOAM plugin passes
    # them, with the plugin's per-solve hipfoam_set_topology (a content check)
    host_ms = None
    if args.host_steps > 0:
        hpsi = np.zeros(m.n)
        lower_h = None if sym else m.lower
        ctx.solve(kind, m.diag, m.upper, lower_h, hpsi, m.source, ctl)
        t0 = time.perf_counter()
        for _ in range(args.host_steps):
            hpsi[:] = 0.0
            ctx.set_topology(m)
            ctx.solve(kind, m.diag, m.upper, lower_h, hpsi, m.source, ctl)
        host_ms = 1e3 * (time.perf_counter() - t0) / args.host_steps
        log(f"refined with host arrays (PCIe and topology check included): "
            f"{host_ms:.3f} ms/solve")

    # ---- cost of hipfoam_set_topology on an unchanged mesh, by check mode
    topo = {}
    for name, mode in (("content", H.TOPO_CONTENT), ("pointers", H.TOPO_POINTERS)):
        ctx.set_topology_check(mode)
        ctx.set_topology(m)
        t0 = time.perf_counter()
        for _ in range(5):
            ctx.set_topology(m)
        topo[name] = round(1e3 * (time.perf_counter() - t0) / 5, 4)
    ctx.set_topology_check(H.TOPO_CONTENT)
    log(f"hipfoam_set_topology on an unchanged mesh: content hash {topo['content']} ms, "
        f"pointers {topo['pointers']} ms")

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
    tag = args.workload + ("" if args.spmv_schedule == 2 else f"_schedule{args.spmv_schedule}")
    _, solve_k, spmv = measure_kernels(H, ctx, m, solver, args.spmv_reps, tag, step)
    slots = solve_k["layout"] == "row-slot"
    log(f"dominant kernel {solve_k['kernel']} ({solve_k['layout']} layout): "
        f"{solve_k['launches']} launches, avg {solve_k['avg_launch_us']} us, "
        f"{solve_k['GB_s']:.0f} GB/s ({100 * solve_k['frac']:.1f}% of HBM peak); LDU-format "
        f"equivalent {solve_k['reference_format_GB_s']:.0f} GB/s")
    if spmv:
        log(f"plain FP32 SpMV: avg {spmv['avg_launch_us']} us, {spmv['GB_s']:.0f} GB/s "
            f"(LDU-format equivalent {spmv['reference_format_GB_s']:.0f} GB/s)")
    sbytes = spmv["bytes_per_launch"] if spmv else 0
    ref_sbytes = spmv["reference_format_bytes"] if spmv else 0

    # the same SpMV with the 256 MB MALL (Infinity Cache) flushed before every
    # launch by a 2 GiB device-to-device copy
    x = D.from_host(ctx, np.random.default_rng(0).standard_normal(m.n).astype(np.float32))
    y = D(ctx, (m.n,), np.float32)
    cold = None
    if args.cold_reps > 0 and spmv:
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
    x.free()
    y.free()

    # ---- small meshes: the persistent inner loop against the launched one
    small = None
    if args.extras and args.small_meshes and args.workload in ("poisson128", "poisson2d100"):
        small = small_meshes(H, mg)

    # ---- the reference benchmark's cavity sizes (2D Poisson, 400 to 640k cells)
    sides = [int(v) for v in args.ref_sizes.split(",") if v.strip()]
    sizes = None
    momentum = None
    if args.extras and args.workload in ("poisson128", "poisson2d100") and sides:
        sizes, size_solutions = reference_sizes(H, mg, sides)
    if args.extras and args.workload in ("poisson128", "poisson2d100") and args.ref_momentum:
        momentum = reference_momentum(H, mg)

    # ---- more rooflines: a mesh past the 256 MB MALL, the LDU face schedule
    # priced on the LDU arrays, the BiCGStab kernels of the 96^3 workload
    extras = None
    if args.extras and args.workload == "poisson128" and args.spmv_schedule == 2:
        extras = {
            "past_mall": extra_mesh(H, mg, args, "config1", (256,), "pcg", "poisson256",
                                    "256^3 lognormal Poisson (working set ~0.7 GB, past the "
                                    "256 MB MALL): served from HBM"),
            "ldu_face_schedule": extra_mesh(H, mg, args, "config1", (), "pcg",
                                            "poisson128_schedule0",
                                            "128^3 with SpMV schedule 0 forced: the LDU face "
                                            "arrays (ownerStart/losortStart, upper, lowerL, "
                                            "uppAddr, losortLow), priced on their own bytes",
                                            schedule=0),
            "convdiff96_bicgstab": extra_mesh(H, mg, args, "config2", (), "bicg", "convdiff96",
                                              "96^3 convection-diffusion, refined "
                                              "hipPBiCGStab: the BiCGStab SpMV kernels"),
            "random_order128": extra_mesh(
                H, mg, args, lambda g: g.random_renumbering(g.config1(), 11), (), "pcg",
                "poisson128_random", "the 128^3 Poisson with its cells in a random order: "
                "every tile holds a row of more than 3 owner faces, so the whole mesh runs "
                "the LDU face loops (and its gathers have no locality)"),
            "rcm_order128": extra_mesh(
                H, mg, args, lambda g: g.rcm_renumbering(g.random_renumbering(g.config1(), 11)),
                (), "pcg", "poisson128_rcm", "the random order renumbered by reverse "
                "Cuthill-McKee (as OpenFOAM's renumberMesh can): rows of at most 3 faces a "
                "side again, row slots in explicit tiles"),
        }

    cpu = None
    errors = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu, x_cpu, floor = cpu_baseline(m, solver, args.cpu_iters)
        if sizes is not None:
            cpu["reference_sizes"] = cpu_reference_sizes(mg, sides, size_solutions)
            for tag, c in cpu["reference_sizes"].items():
                sizes[tag]["cpu_dic_ms"] = c["cpu_dic_ms"]
                for k in ("gpu_vs_cpu_relmax", "cpu_vs_cpu_floor_relmax"):
                    if k in c:
                        sizes[tag][k] = c[k]
                sizes[tag]["speedup_vs_cpu"] = round(
                    c["cpu_dic_ms"] / sizes[tag]["host_arrays_ms"], 2)
            log("reference sizes, CPU PCG(DIC) 1 core: " +
                ", ".join(f"{t} {c['cpu_dic_ms']} ms" for t, c in cpu["reference_sizes"].items()))
        if momentum is not None:
            cm = cpu_reference_momentum(mg)
            cpu["reference_momentum"] = cm
            momentum["cpu_pbicgstab_diagonal_ms"] = cm["cpu_pbicgstab_diagonal_ms"]
            momentum["speedup_vs_cpu"] = round(cm["cpu_pbicgstab_diagonal_ms"] /
                                               momentum["call_ms"], 2)
            log(f"reference momentum, CPU PBiCGStab(diagonal) 1 core: "
                f"{cm['cpu_pbicgstab_diagonal_ms']} ms (iterations {cm['cpu_iterations']})")
        log(f"cpu baseline: {cpu['value']} ms ({cpu['sample']}; {cpu['build']})")
        if cpu.get("all_cores"):
            log(f"cpu all cores: {cpu['all_cores']}")
        if x_cpu is not None:
            # both solutions meet the same 1e-6 normalised residual; their
            # difference is what that tolerance leaves, not an error of either.
            # The floor is the same measure between two CPU solvers.
            errors = {"refined_vs_cpu_fp64_relmax": _relmax(psi, x_cpu)}
            if psi64 is not None:
                errors["plain_fp64_vs_cpu_fp64_relmax"] = _relmax(psi64, x_cpu)
            if floor:
                errors.update(floor)
            log("solution vs CPU FP64 solve: " +
                ", ".join(f"{k} {v:.3e}" if isinstance(v, float) else f"{k} {v}"
                          for k, v in errors.items()))

    roof = dict(solve_k)
    roof.update({"bound": "hbm", "achieved": solve_k["GB_s"], "peak": HBM_PEAK_GBS,
                 "unit": "GB/s"})
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
        "roofline": roof,
        "spmv": None if spmv is None else dict(spmv, mall_flushed=cold, autotuned=tuned),
        "rooflines_more": extras,
        "small_meshes": small,
        "reference_sizes": sizes,
        "reference_momentum": momentum,
        "solve": {
            "iterations": perf.nIterations,
            "restarts": perf.restarts,
            "final_residual": perf.finalResidual,
            "numpy_residual": true_res,
            "converged": bool(perf.converged and true_res < 1e-6),
            "deflated": perf.deflated,
            "polished": perf.polished,
            "host_arrays_ms": None if host_ms is None else round(host_ms, 3),
            "topology_check_ms": topo,
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
        if cpu.get("all_cores"):
            out["speedup_vs_cpu_all_cores"] = round(cpu["all_cores"]["value"] / ms, 1)
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
    as_specified = args.batch_rhs == "as-specified"
    systems = mg.config3(zero_mean_poisson=not as_specified)
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
    # as specified, the Poisson right-hand sides are N(0,1) draws with a
    # non-zero sum: the pinned projection (deflate 3) solves them; with the
    # reference's zero-mean projection some stall until the polish
    ctl = H.default_controls(maxRestarts=8, maxIter=20000, deflate=3 if as_specified else 2)

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
                          "cells": int(sum(m.n for m in systems)), "tolerance": 1e-6,
                          "poisson_rhs": ("N(0,1) as drawn (non-zero sum), deflate 3"
                                          if as_specified else
                                          "N(0,1) projected to zero mean, deflate auto")})
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
