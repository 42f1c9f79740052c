"""Per-wavefront instruction and wait counts of iteration kernels from
rocprofv3 --pmc runs (SQ_*, TCP_* counters): how many vector-memory, VALU,
scalar-ALU and scalar-memory instructions one wavefront issues, how long it
lives, how much of that it waits, and the mean L2 read latency its vector
loads see.  Each counter is averaged over the kernel's launches in the top
quartile of its value (the launches that do the full work: every segment
still running, no launch stopped at entry).  By default the batched
iteration kernels; --kernel and --title select others.

    python -m metal4foam_amd.profsum_waves --pmc A.csv --pmc B.csv \\
        [--kernel NAME ...] [--title TEXT] [--command CMD ...] [--note TEXT ...]
"""

import argparse
import collections
import csv

KERNELS = [
    "kb_cheb_step<true, false>", "kb_cheb_step<false, false>", "kb_cheb_step<false, true>",
    "kb_pcg_p_spmv", "kb_pcg_update", "kb_bicg_py", "kb_bicg_y_spmv", "kb_bicg_sz",
    "kb_bicg_z_spmv", "kb_bicg_update_rho",
]


def _short(name):
    return name.replace("void ", "").replace("hipfoam::", "").split("(")[0]


def load(paths):
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for p in paths:
        with open(p) as f:
            for r in csv.DictReader(f):
                vals[_short(r["Kernel_Name"])][r["Counter_Name"]].append(
                    float(r["Counter_Value"]))
    out = {}
    for k, cs in vals.items():
        out[k] = {}
        for c, v in cs.items():
            v = sorted(v)
            top = v[len(v) * 3 // 4:] or v
            out[k][c] = sum(top) / len(top)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pmc", action="append", required=True,
                    help="run_counter_collection.csv of one --pmc pass (repeatable)")
    ap.add_argument("--kernel", action="append", default=[],
                    help="kernel name as the trace shows it, without 'hipfoam::' and the "
                         "argument list (repeatable; default: the batched kernels)")
    ap.add_argument("--title", default="Batched iteration kernels, 64 systems x 5 RHS: "
                    "per-wavefront instructions, wave life and L2 read latency (MI355X)")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--note", action="append", default=[])
    a = ap.parse_args()
    res = load(a.pmc)
    print(f"# {a.title}\n")
    for c in a.command:
        print(f"    {c}")
    print()
    print("Per wavefront (SQ_INSTS_* / SQ_WAVES, SQ_WAVE_CYCLES / SQ_WAVES, SQ_WAIT_ANY / "
          "SQ_WAVES); L2 read latency = TCP_TCC_READ_REQ_LATENCY / TCP_TCC_READ_REQ, in "
          "cycles.\n")
    print("| kernel | waves | VMEM reads | VMEM writes | VALU | SALU | SMEM | wave life "
          "(cycles) | waiting (cycles) | waiting share | L2 read latency |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    for k in a.kernel or KERNELS:
        r = res.get(k)
        if not r or not r.get("SQ_WAVES"):
            continue
        w = r["SQ_WAVES"]

        def per(c):
            return r.get(c, 0.0) / w
        life, wait = per("SQ_WAVE_CYCLES"), per("SQ_WAIT_ANY")
        req = r.get("TCP_TCC_READ_REQ", 0.0)
        lat = r.get("TCP_TCC_READ_REQ_LATENCY", 0.0) / req if req else float("nan")
        share = wait / life if life else float("nan")
        print(f"| `{k}` | {w:,.0f} | {per('SQ_INSTS_VMEM_RD'):.1f} | "
              f"{per('SQ_INSTS_VMEM_WR'):.1f} | {per('SQ_INSTS_VALU'):.0f} | "
              f"{per('SQ_INSTS_SALU'):.0f} | {per('SQ_INSTS_SMEM'):.1f} | {life:,.0f} | "
              f"{wait:,.0f} | {100 * share:.0f} % | {lat:.0f} |")
    for n in a.note:
        print()
        print(n)


if __name__ == "__main__":
    main()
