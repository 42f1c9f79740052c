"""bench.py's one-line JSON contract, on a small workload in a child process:
the required keys and types, a time-like value with higher_is_better false,
a roofline whose fraction is achieved / peak and below 1 and whose achieved
rate is its bytes per launch over its average launch time, a CPU baseline
timed on one core with its sample stated, the small-mesh and reference-size
sections, and --gpus above 1 refused with a message."""

import json
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]

pytestmark = pytest.mark.gpu


def _bench(*args, timeout=300):
    print(f"[bench-contract] bench.py {' '.join(args)}", flush=True)
    r = subprocess.run([sys.executable, "bench.py", *args], cwd=ROOT, capture_output=True,
                       text=True, timeout=timeout)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    return r, (json.loads(lines[-1]) if lines else None)


def test_bench_line_contract():
    r, out = _bench("--workload", "poisson2d100", "--steps", "2", "--warmup", "1",
                    "--fp64-steps", "1", "--host-steps", "1", "--spmv-reps", "5",
                    "--spmv-tune", "0", "--cold-reps", "0", "--extras", "0")
    assert r.returncode == 0, r.stderr[-3000:]
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "cpu_baseline"):
        assert key in out, key
    assert out["n_gpus"] == 1 and out["steps"] == 2 and out["warmup"] == 1
    assert out["unit"] == "ms" and out["higher_is_better"] is False
    assert out["value"] > 0 and out["value"] == pytest.approx(out["ms_per_step"])
    assert out["vs_baseline"] is None and out["data"] == "synthetic"
    assert "workload" in out["config"] and "model" not in out["config"]
    rf = out["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert 0 < rf["frac"] < 1
    # the line rounds frac to 4 decimals and achieved to 0.1 GB/s
    assert rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], abs=6e-5)
    assert rf["achieved"] == pytest.approx(
        rf["bytes_per_launch"] / (rf["avg_launch_us"] * 1e-6) / 1e9, rel=2e-3)
    cpu = out["cpu_baseline"]
    assert cpu["cores"] == 1 and cpu["kind"] in ("port", "reference")
    assert cpu["value"] > 0 and cpu["unit"] == "ms" and cpu["sample"]
    assert out["solve"]["converged"] and out["solve"]["numpy_residual"] < 1e-6


def test_bench_small_meshes():
    """The small-mesh section: every case converged on the persistent inner
    loop (never abandoned) in the same iterations as the launched loop."""
    r, out = _bench("--workload", "poisson2d100", "--steps", "1", "--warmup", "1",
                    "--fp64-steps", "0", "--host-steps", "0", "--spmv-reps", "2",
                    "--spmv-tune", "0", "--cold-reps", "0", "--no-cpu", "--extras", "1",
                    "--ref-sizes", "", "--ref-momentum", "0")
    assert r.returncode == 0, r.stderr[-3000:]
    assert out["reference_sizes"] is None and out["reference_momentum"] is None
    small = out["small_meshes"]
    assert set(small) == {"poisson2d100", "convdiff2d100", "poisson24"}
    for tag, e in small.items():
        assert e["converged"] and e["same_iterations"], tag
        assert e["persistent_launches"] >= 1 and e["abandoned"] == 0, tag
        assert e["solve_ms"] > 0 and e["launched_loop_ms"] > 0, tag


def test_bench_reference_sizes():
    """The reference benchmark's cavity sizes and its momentum call: each
    solve converged, the published M3 Pro per-call times carried unchanged,
    the same-host CPU times taken from the CPU leg, and the ratios their
    quotients."""
    published = {"20x20": (0.066, 1.109), "50x50": (2.082, 5.838),
                 "100x100": (20.52, 11.63)}
    r, out = _bench("--workload", "poisson2d100", "--steps", "1", "--warmup", "0",
                    "--fp64-steps", "0", "--host-steps", "0", "--spmv-reps", "2",
                    "--spmv-tune", "0", "--cold-reps", "0", "--extras", "1",
                    "--ref-sizes", "20,50,100")
    assert r.returncode == 0, r.stderr[-3000:]
    sizes = out["reference_sizes"]
    assert set(sizes) == set(published)
    cpu = out["cpu_baseline"]["reference_sizes"]
    for tag, (cpu_ref, gpu_ref) in published.items():
        e = sizes[tag]
        side = int(tag.split("x")[0])
        assert e["cells"] == side * side and e["converged"], tag
        assert e["solve_ms"] > 0 and e["host_arrays_ms"] > 0, tag
        assert (e["reference_cpu_dic_ms"], e["reference_metalpcg_ms"]) == (cpu_ref, gpu_ref)
        assert e["vs_reference_metalpcg"] == pytest.approx(gpu_ref / e["host_arrays_ms"],
                                                           abs=0.051)
        assert cpu[tag]["cpu_converged"] and e["cpu_dic_ms"] == cpu[tag]["cpu_dic_ms"] > 0
        # the reference's accuracy measure beside its CPU-vs-CPU floor
        assert 0 < e["gpu_vs_cpu_relmax"] < 1e-4 and 0 <= e["cpu_vs_cpu_floor_relmax"] < 1e-4
        assert e["speedup_vs_cpu"] == pytest.approx(e["cpu_dic_ms"] / e["host_arrays_ms"],
                                                    abs=0.0051)
    # the momentum call: Ux and Uy on 40,000 cells, beside the published
    # 11.16 ms (CPU PBiCGStab, diagonal) and 13.75 ms (metalPBiCGStab)
    mo = out["reference_momentum"]
    assert mo["cells"] == 40000 and mo["converged"] and len(mo["iterations"]) == 2
    assert (mo["reference_cpu_pbicgstab_diagonal_ms"], mo["reference_metalpbicgstab_ms"]) == (
        11.16, 13.75)
    assert mo["vs_reference_metalpbicgstab"] == pytest.approx(13.75 / mo["call_ms"], abs=0.051)
    cm = out["cpu_baseline"]["reference_momentum"]
    assert all(0 < i < 40 for i in cm["cpu_iterations"])
    assert mo["cpu_pbicgstab_diagonal_ms"] == cm["cpu_pbicgstab_diagonal_ms"] > 0
    assert mo["speedup_vs_cpu"] == pytest.approx(mo["cpu_pbicgstab_diagonal_ms"] / mo["call_ms"],
                                                 abs=0.0051)


def test_bench_refuses_several_gpus():
    r, out = _bench("--gpus", "2", timeout=120)
    assert r.returncode != 0
    assert out is not None and "one GPU" in out["error"]
