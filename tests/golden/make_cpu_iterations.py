"""Writes cpu_baseline_iterations.json: iteration counts of the oracle's
FP64 PCG(DIC) / PBiCGStab(DILU) full solves to 1e-6 on the benchmark
workloads (used to scale the bounded CPU-baseline sample in bench.py).

    python tests/golden/make_cpu_iterations.py
"""

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from metal4foam_amd import meshgen as mg  # noqa: E402

out = {}
for gen, solver in (("config0", "pcg"), ("config2", "bicg"), ("config1", "pcg")):
    m = getattr(mg, gen)()
    t0 = time.perf_counter()
    if solver == "pcg":
        _, perf = oracle.pcg(m, precon=oracle.PRE_DIC, tol=1e-6)
    else:
        _, perf = oracle.pbicgstab(m, precon=oracle.PRE_DIC, tol=1e-6)
    dt = time.perf_counter() - t0
    out[m.name] = {"solver": "PCG(DIC)" if solver == "pcg" else "PBiCGStab(DILU)",
                   "iterations": perf.nIterations, "final_residual": perf.finalResidual,
                   "converged": bool(perf.converged), "seconds_here": round(dt, 2)}
    print(m.name, out[m.name], flush=True)
json.dump(out, open(os.path.join(HERE, "cpu_baseline_iterations.json"), "w"), indent=1)
