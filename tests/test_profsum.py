"""The measurement tooling on the CPU: the algorithmic bytes that bench.py
prices its rooflines on agree with the ones the rocprof summaries under
profiles/ use, they reproduce the per-launch figures of the 128^3 workload
(86,245,376 B for the interior Chebyshev step on the row-slot layout,
175,374,344 B in the LDU format, 52,690,944 B for the plain SpMV), and the
per-wavefront counter summary divides each counter by the wavefront count."""

import csv
import importlib.util
import pathlib
import subprocess
import sys

import pytest

from metal4foam_amd import profsum

ROOT = pathlib.Path(__file__).resolve().parents[1]

N128, F128 = 128 ** 3, 3 * 128 * 128 * 127


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_faces_of_the_128_cube():
    assert F128 == 6_242_304


@pytest.mark.parametrize("sym", [True, False])
@pytest.mark.parametrize("patterned", [N128, 0, N128 // 2])
def test_bench_and_profiles_price_kernels_alike(bench, sym, patterned):
    asym = not sym
    step = bench.cheb_step_bytes(N128, F128, sym)
    assert step == profsum._ldu_bytes("k_cheb_step<float, false>", N128, F128, asym)
    spmv = bench.spmv_bytes(N128, F128, sym)
    assert spmv == profsum._ldu_bytes("k_spmv<float>", N128, F128, asym)
    for name, ldu in (("k_cheb_step<float, false>", step), ("k_spmv<float>", spmv)):
        assert bench.to_slot_bytes(ldu, N128, F128, sym, patterned) == profsum._bytes(
            name, N128, F128, asym, slots=True, patterned_rows=patterned)
    # the BiCGStab SpMVs bench averages: (py + sz) / 2
    py = profsum._ldu_bytes("k_bicg_py_spmv<float>", N128, F128, asym)
    sz = profsum._ldu_bytes("k_bicg_sz_spmv<float>", N128, F128, asym)
    assert bench.spmv_bytes(N128, F128, sym, extra_vectors=5) == (py + sz) // 2


def test_published_128_figures(bench):
    step = bench.cheb_step_bytes(N128, F128, True)
    assert step == 175_374_344
    assert bench.to_slot_bytes(step, N128, F128, True, N128) == 86_245_376
    spmv = bench.spmv_bytes(N128, F128, True)
    assert bench.to_slot_bytes(spmv, N128, F128, True, N128) == 52_690_944
    # 41.125 B per row: 12 B of coefficients, the slot byte, 1/8 B of tile
    # descriptor, 4 B diag, r, rD, z read, w read and written, z written
    assert 86_245_376 / N128 == pytest.approx(41.125)


def test_wave_summary_divides_by_waves(tmp_path):
    path = tmp_path / "run_counter_collection.csv"
    fields = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    rows = []
    for launch in range(4):
        scale = 1.0 if launch else 0.01  # a stopped launch, left out
        for c, v in (("SQ_WAVES", 100), ("SQ_INSTS_VMEM_RD", 6200), ("SQ_INSTS_VMEM_WR", 1000),
                     ("SQ_INSTS_VALU", 29000), ("SQ_INSTS_SALU", 31700),
                     ("SQ_INSTS_SMEM", 3800), ("SQ_WAVE_CYCLES", 751000),
                     ("SQ_WAIT_ANY", 542100), ("TCP_TCC_READ_REQ", 1000),
                     ("TCP_TCC_READ_REQ_LATENCY", 927000)):
            rows.append({"Kernel_Name": "void hipfoam::kb_cheb_step<false, false>(args)",
                         "Counter_Name": c, "Counter_Value": v * scale})
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        w.writerows(rows)
    out = subprocess.run([sys.executable, "-m", "metal4foam_amd.profsum_waves", "--pmc",
                          str(path)], cwd=ROOT, capture_output=True, text=True, check=True)
    line = next(l for l in out.stdout.splitlines() if "kb_cheb_step<false, false>" in l)
    cells = [c.strip() for c in line.strip("|").split("|")]
    assert cells[1:] == ["100", "62.0", "10.0", "290", "317", "38.0", "7,510", "5,421", "72 %",
                         "927"]
