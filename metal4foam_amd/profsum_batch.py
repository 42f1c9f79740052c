"""Summarise a rocprofv3 kernel trace of the batched workload (64 systems x 5
right-hand sides, ``bench.py --workload batch64``) into a short Markdown table.

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT/batch -o run -- \\
        python3 bench.py --workload batch64 --steps 1 --warmup 0
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- \\
        python3 bench.py --workload batch64 --steps 1 --warmup 0
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- \\
        python3 bench.py --workload batch64 --steps 1 --warmup 0
    python -m metal4foam_amd.profsum_batch --trace OUT/batch/run_kernel_trace.csv \\
        [--fetch OUT/fetch/run_counter_collection.csv \\
         --write OUT/write/run_counter_collection.csv] \\
        --command "..." > profiles/kb_cheb_step_batch64.md

FETCH_SIZE / WRITE_SIZE (KiB per dispatch) are scaled by the k_copy<double>
calibration of the single-system profile (profiles/traffic_poisson128.json:
FETCH_SIZE x 1.998, WRITE_SIZE x 1.0) and averaged over the full launches,
those within 10 % of the kernel's largest count.

The iteration kernels run one workgroup per (256-row tile, group of RHS) and
skip the right-hand sides whose segment has stopped, so a launch moves fewer
bytes as segments converge.  The bandwidth column therefore uses the launches
of the first iterations, while every segment still runs: the launches within
10 % of the kernel's longest.  Their bytes are priced on the batch's layout:
row slots with every tile patterned (one slot byte per row, three FP32
coefficient planes for the symmetric batch and six for the asymmetric one,
one 32-byte descriptor per tile), the two diagonal planes of a system (exact
and deflated; a segment reads one diag and one 1/diag value per row, which
the other RHS of the system then read from cache) and the per-RHS vectors
each kernel reads and writes.
"""

from __future__ import annotations

import argparse
import collections
import csv
import re

# per-RHS vector bytes per row, and whether the kernel reads the matrix slots
# and the diagonal / 1/diagonal planes (bytes per row)
KERNELS = {
    "kb_cheb_step<false, false>": ("sym", 20, True, 8),   # r, z, w in; w, z out
    "kb_cheb_step<true, false>": ("sym", 12, True, 8),    # r in; w, z out
    "kb_cheb_step<false, true>": ("sym", 16, True, 8),    # r, z, w in; z out
    "kb_pcg_p_spmv": ("sym", 16, True, 4),                # w, p in; p, Ap out
    "kb_pcg_update": ("sym", 24, False, 0),               # delta, r in/out; p, Ap in
    "kb_bicg_py": ("asym", 20, False, 4),                 # r, p, AyA in; p, y out
    "kb_bicg_y_spmv": ("asym", 12, True, 4),              # y, rA0 in; AyA out
    "kb_bicg_sz": ("asym", 16, False, 4),                 # r, AyA in; s, z out
    "kb_bicg_z_spmv": ("asym", 12, True, 4),              # z, s in; t out
    "kb_bicg_update_rho": ("asym", 32, False, 0),         # delta in/out; y, z, s, t, rA0 in; r out
}


def _short(name: str) -> str:
    name = re.sub(r"^void ", "", name)
    name = name.replace("hipfoam::", "")
    name = re.sub(r"\(.*$", "", name) if not name.startswith("kb_cheb_step") else \
        re.sub(r">\(.*$", ">", name)
    return name


def batch_geometry():
    """Rows, tiles and RHS count of the symmetric and asymmetric batches."""
    from . import meshgen as mg

    systems = mg.config3()
    nrhs = systems[0].source.shape[1]
    geo = {}
    for kind, sel in (("sym", [m for m in systems if m.lower is None]),
                      ("asym", [m for m in systems if m.lower is not None])):
        rows = sum(m.n for m in sel)
        tiles = sum((m.n + 255) // 256 for m in sel)
        geo[kind] = (rows, tiles, len(sel))
    return geo, nrhs


def launch_bytes(kernel: str, geo, nrhs: int) -> int:
    kind, vec, matrix, diag = KERNELS[kernel]
    rows, tiles, _ = geo[kind]
    mat = 0
    if matrix:
        planes = 3 if kind == "sym" else 6
        mat = planes * 4 * rows + rows + 32 * tiles
    return mat + diag * rows + nrhs * vec * rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--peak-gbs", type=float, default=8000.0)
    ap.add_argument("--full", type=float, default=0.9,
                    help="launches at least this fraction of the kernel's longest count as full")
    ap.add_argument("--top", type=int, default=14)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--fetch-scale", type=float, default=1.998)
    ap.add_argument("--write-scale", type=float, default=1.0)
    ap.add_argument("--note", action="append", default=[],
                    help="a paragraph printed after the tables")
    args = ap.parse_args()

    geo, nrhs = batch_geometry()
    durs = collections.defaultdict(list)
    with open(args.trace) as f:
        for r in csv.DictReader(f):
            d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            durs[_short(r["Kernel_Name"])].append(d / 1e3)

    print("# Batched refined solves, 64 systems x 5 RHS: kernel durations and bandwidth "
          "of the full launches (MI355X)\n")
    for c in args.command:
        print(f"    {c}")
    print()
    s, a = geo["sym"], geo["asym"]
    print(f"Workload: {s[2]} symmetric systems ({s[0]:,} rows, {s[1]:,} tiles, hipPCG) and "
          f"{a[2]} asymmetric systems ({a[0]:,} rows, {a[1]:,} tiles, hipPBiCGStab), "
          f"{nrhs} RHS each; row slots, every tile patterned. Durations: rocprofv3 kernel "
          f"trace.\n")
    print("| kernel | launches | total ms | avg us | full launches | full avg us | "
          "MB per full launch | GB/s | % of 8 TB/s |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|---:|")
    order = sorted(durs.items(), key=lambda kv: -sum(kv[1]))[:args.top]
    for name, ds in order:
        tot = sum(ds)
        row = f"| `{name}` | {len(ds)} | {tot / 1e3:.2f} | {tot / len(ds):.2f} |"
        if name in KERNELS:
            mx = max(ds)
            full = [d for d in ds if d >= args.full * mx]
            avg = sum(full) / len(full)
            b = launch_bytes(name, geo, nrhs)
            gbs = b / (avg * 1e-6) / 1e9
            row += (f" {len(full)} | {avg:.1f} | {b / 1e6:.1f} | {gbs:.0f} | "
                    f"{100 * gbs / args.peak_gbs:.1f} |")
        else:
            row += "  |  |  |  |  |"
        print(row)
    print()
    if "kb_spmv" in durs:
        print("`kb_spmv` is the batch autotuner's timing kernel (40 candidates, before the "
              "timed step).\n")
    if args.fetch and args.write:
        cnt = {}
        for path, name, scale in ((args.fetch, "FETCH_SIZE", args.fetch_scale),
                                  (args.write, "WRITE_SIZE", args.write_scale)):
            vals = collections.defaultdict(list)
            with open(path) as f:
                for r in csv.DictReader(f):
                    if r.get("Counter_Name", name) == name:
                        vals[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
            cnt[name] = {k: scale * 1024.0 * sum(x for x in v if x >= args.full * max(v)) /
                         len([x for x in v if x >= args.full * max(v)])
                         for k, v in vals.items() if v and max(v) > 0}
        print("HBM traffic of the full launches (FETCH_SIZE x %.3f + WRITE_SIZE x %.3f):\n"
              % (args.fetch_scale, args.write_scale))
        print("| kernel | fetch MB | write MB | traffic MB | algorithmic MB | traffic / algorithmic |")
        print("|---|---:|---:|---:|---:|---:|")
        for k in KERNELS:
            if k in cnt["FETCH_SIZE"] and k in cnt["WRITE_SIZE"]:
                fb, wb = cnt["FETCH_SIZE"][k], cnt["WRITE_SIZE"][k]
                alg = launch_bytes(k, geo, nrhs)
                print(f"| `{k}` | {fb / 1e6:.1f} | {wb / 1e6:.1f} | {(fb + wb) / 1e6:.1f} | "
                      f"{alg / 1e6:.1f} | {(fb + wb) / alg:.2f} |")
        print()
    print("Bytes per full launch: matrix slots (when the kernel applies A) + diagonal "
          "planes + RHS x per-RHS vector bytes, per row: " + "; ".join(
              f"`{k}` {v[1]} B/RHS{' + slots' if v[2] else ''}{f' + {v[3]} B diag' if v[3] else ''}"
              for k, v in KERNELS.items()) + ".")
    for n in args.note:
        print()
        print(n)


if __name__ == "__main__":
    main()
