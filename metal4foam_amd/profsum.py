"""Summarise rocprofv3 CSV output of the LDU solve into a short Markdown table.

Inputs are what these commands write (run from the repository root):

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o run -- python3 bench.py ...
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- python3 bench.py ...
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- python3 bench.py ...

    python -m metal4foam_amd.profsum --trace OUT/trace/run_kernel_trace.csv \\
        --fetch OUT/fetch/run_counter_collection.csv \\
        --write OUT/write/run_counter_collection.csv --cells N --faces F \\
        [--asym] --title "..." --command "..." > profiles/<kernel>_<workload>.md

Launches shorter than --noop-us are counted apart: once the device-side
convergence test has stopped the loop, the launches still queued return at
entry, and would otherwise dilute the averages.

FETCH_SIZE / WRITE_SIZE (KiB per dispatch) are scaled by what the same
counter reads for k_copy<double>, a copy of known size (8 B read + 8 B
written per element).  Counter values below 5 % of a kernel's largest
dispatch are the stopped launches and are left out of its mean.

--layout slots prices the SpMV kernels on the row-slot layout instead of the
row-start arrays and per-face arrays: six coefficient planes (three for a
symmetric matrix), one slot byte per row, three packed int16 offset words per
row outside patterned tiles (--patterned-rows, as hipfoam_slot_layout
reports) and one 32-byte descriptor per 256-row tile.  --traffic-json writes the calibrated
HBM bytes per dispatch of one kernel (--traffic-kernel) as JSON.
"""

from __future__ import annotations

import argparse
import collections
import csv
import re


def slot_matrix_bytes(n: int, asym: bool, coef: int, patterned_rows: int) -> int:
    """Matrix bytes one pass over the row-slot layout reads: the coefficient
    planes (six, or three for a symmetric matrix whose neighbour-side
    coefficients are gathered from their owner rows), one slot byte per row,
    three packed int16 offset words per row of an explicit tile and one
    32-byte descriptor per 256-row tile."""
    planes = 6 if asym else 3
    tiles = (n + 255) // 256
    return planes * coef * n + n + 12 * (n - patterned_rows) + 32 * tiles


def _bytes(name: str, n: int, nf: int, asym: bool, slots: bool = False,
           patterned_rows: int = 0) -> int | None:
    """Algorithmic bytes per launch of the kernels on the solve path."""
    b = _ldu_bytes(name, n, nf, asym)
    if b is None or not slots or "k_pcg_update" in name or "k_bicg_update" in name:
        return b
    if "<float" in name:
        face, coef = (20 if asym else 16), 4
    elif "<double" in name:
        face, coef = (28 if asym else 24), 8
    else:
        return b
    # row starts and per-face arrays out, the row-slot matrix in
    return b - (8 * n + 8 + face * nf) + slot_matrix_bytes(n, asym, coef, patterned_rows)


def _ldu_bytes(name: str, n: int, nf: int, asym: bool) -> int | None:
    """LDU-format algorithmic bytes per launch, matched on the kernel name's
    prefix (template arguments after the first ones are finalisers)."""
    face32 = 20 if asym else 16
    face64 = 28 if asym else 24
    table = (
        ("k_cheb_step<float, false>", 36 * n + 8 + face32 * nf),
        ("k_cheb_step<float, true>", 32 * n + 8 + face32 * nf),
        # previous iteration's update applied (every launch but an inner
        # solve's first): reads r, Ap, delta, p, rD; writes r, delta, w, z
        ("k_cheb_first<float, false,", 48 * n + 8 + face32 * nf),
        ("k_pcg_p_spmv<float", 28 * n + 8 + face32 * nf),
        ("k_pcg_update<float", 24 * n),
        ("k_spmv<float", 20 * n + 8 + face32 * nf),
        ("k_cheb_step<double, false>", 64 * n + 8 + face64 * nf),
        ("k_cheb_first<double, false,", 88 * n + 8 + face64 * nf),
        ("k_pcg_p_spmv<double", 48 * n + 8 + face64 * nf),
        ("k_bicg_py_spmv<float", 44 * n + 8 + face32 * nf),
        ("k_bicg_sz_spmv<float", 36 * n + 8 + face32 * nf),
        ("k_bicg_update_rho<float", 32 * n),
    )
    for prefix, b in table:
        if prefix in name:
            return b
    return None


def _short(name: str) -> str:
    name = re.sub(r"\(.*", "", name)
    return name.replace("void ", "").replace("hipfoam::", "")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--cells", type=int, required=True)
    ap.add_argument("--faces", type=int, required=True)
    ap.add_argument("--asym", action="store_true")
    ap.add_argument("--noop-us", type=float, default=6.0)
    ap.add_argument("--top", type=int, default=12)
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--peak-gbs", type=float, default=8000.0)
    ap.add_argument("--layout", choices=["faces", "slots"], default="faces")
    ap.add_argument("--patterned-rows", type=int, default=0)
    ap.add_argument("--traffic-json")
    ap.add_argument("--traffic-kernel", default="k_cheb_step<float, false>")
    ap.add_argument("--traffic", action="append", default=[],
                    help="PATH=KERNEL[|KERNEL...]: write the calibrated traffic of these "
                         "kernels (their mean, as the benchmark averages them) to PATH")
    ap.add_argument("--workload", default="")
    a = ap.parse_args()
    slots = a.layout == "slots"

    durs = collections.defaultdict(list)
    for r in csv.DictReader(open(a.trace)):
        durs[r["Kernel_Name"]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)

    def counters(path):
        out = collections.defaultdict(list)
        if path:
            for r in csv.DictReader(open(path)):
                out[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        for k, v in out.items():
            top = max(v)
            out[k] = [x for x in v if x >= 0.05 * top] if top > 0 else v
        return out

    fetch, write = counters(a.fetch), counters(a.write)

    def mean(v):
        return sum(v) / len(v) if v else float("nan")

    copy = [k for k in fetch if "k_copy<double>" in k]
    fscale = wscale = None
    if copy and a.fetch and a.write:
        # k_copy<double> over the cells: 8 B read and 8 B written per cell
        known = 8.0 * a.cells
        fscale = known / (1024.0 * mean(fetch[copy[0]]))
        wscale = known / (1024.0 * mean(write[copy[0]]))

    print(f"# {a.title}\n")
    for c in a.command:
        print(f"    {c}")
    print()
    lay = (f"row-slot layout, {a.patterned_rows:,} rows in patterned tiles" if slots
           else "LDU face layout")
    print(f"Workload: {a.cells:,} cells, {a.faces:,} faces, "
          f"{'asymmetric' if a.asym else 'symmetric'}, {lay}. Durations: rocprofv3 kernel trace; "
          f"for row kernels, launches of at least {a.noop_us} us (shorter launches returned at "
          "entry after the device-side stop and are listed apart).\n")
    if fscale:
        print(f"HBM counter calibration on k_copy<double> ({8 * a.cells:,} B read, "
              f"{8 * a.cells:,} B written): FETCH_SIZE x {fscale:.3f}, WRITE_SIZE x {wscale:.3f}.\n")
    print("| kernel | launches (stopped) | avg us | algorithmic MB | GB/s | % of 8 TB/s "
          "| traffic MB (fetch+write) | traffic / algorithmic |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|")
    rows = sorted(durs.items(), key=lambda kv: -sum(kv[1]))[: a.top]
    for name, v in rows:
        b = _bytes(name, a.cells, a.faces, a.asym, slots, a.patterned_rows)
        # the cut applies to row kernels; one-workgroup kernels are short anyway
        real = [x for x in v if x >= a.noop_us] if b is not None else list(v)
        stopped = len(v) - len(real)
        avg = mean(real) if real else mean(v)
        gbs = b / (avg * 1e3) if (b and real) else None
        tr = None
        if fscale and name in fetch and name in write:
            tr = (mean(fetch[name]) * fscale + mean(write[name]) * wscale) * 1024.0
        print(f"| `{_short(name)}` | {len(real)} ({stopped}) | {avg:.2f} | "
              f"{'' if b is None else f'{b / 1e6:.1f}'} | {'' if gbs is None else f'{gbs:.0f}'} | "
              f"{'' if gbs is None else f'{100 * gbs / a.peak_gbs:.1f}'} | "
              f"{'' if tr is None else f'{tr / 1e6:.1f}'} | "
              f"{'' if (tr is None or b is None) else f'{tr / b:.2f}'} |")


    jobs = list(a.traffic)
    if a.traffic_json:
        jobs.append(f"{a.traffic_json}={a.traffic_kernel}")
    if jobs and fscale:
        import json

        for job in jobs:
            path, kernels = job.split("=", 1)
            wanted = kernels.split("|")
            names = []
            for w in wanted:
                hit = [k for k in fetch if w in k and k in write]
                if hit:
                    names.append(hit[0])
            if len(names) != len(wanted):
                continue
            f_kib = mean([mean(fetch[k]) for k in names])
            w_kib = mean([mean(write[k]) for k in names])
            alg = [_bytes(k, a.cells, a.faces, a.asym, slots, a.patterned_rows) for k in names]
            out = {
                "kernel": " / ".join(wanted),
                "workload": a.workload,
                "layout": "row-slot" if slots else "LDU face",
                "patterned_rows": a.patterned_rows if slots else None,
                "fetch_kib_per_dispatch": round(f_kib, 1),
                "write_kib_per_dispatch": round(w_kib, 1),
                "calibration": {"kernel": "k_copy<double>", "bytes_read": 8 * a.cells,
                                "bytes_written": 8 * a.cells, "fetch_scale": round(fscale, 4),
                                "write_scale": round(wscale, 4)},
                "bytes_per_launch": int(round((f_kib * fscale + w_kib * wscale) * 1024.0)),
                "algorithmic_bytes_per_launch": (int(round(mean(alg))) if None not in alg
                                                 else None),
                "commands": a.command,
            }
            with open(path, "w") as fh:
                json.dump(out, fh, indent=1)
                fh.write("\n")


if __name__ == "__main__":
    main()
