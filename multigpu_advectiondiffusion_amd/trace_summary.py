"""Summaries of rocprofv3 kernel traces for the profiles/ notes.

    python -m multigpu_advectiondiffusion_amd.trace_summary TRACE_CSV
        [--kernel REGEX --last N]
    python -m multigpu_advectiondiffusion_amd.trace_summary --pmc COUNTER_CSV...
        [--kernel REGEX] [--points N] [--calib-kib K --calib-n C]

TRACE_CSV is the `*_kernel_trace.csv` that `rocprofv3 --kernel-trace
--output-format csv` writes (one row per dispatch with start and end
timestamps in ns).  Prints a markdown table of every kernel's dispatches
(calls, total, average, min, max) and, with --kernel/--last, the same figures
over only the last N dispatches of the kernels matching REGEX: for a
benchmark that runs W warm-up steps and then K timed steps of L launches,
--last K*L selects the timed launches, whose average is the one to set
beside the benchmark's own event timing.

With --pmc the files are `*_counter_collection.csv` files of `rocprofv3
--pmc` runs (one row per dispatch and counter).  Per kernel it prints the
launch shape and each counter's mean per dispatch; with --points N (grid
points one dispatch covers) also each counter per 64 points, i.e. per wave
for a kernel that gives every lane one point of each plane it visits.
FETCH_SIZE and WRITE_SIZE (KiB) are scaled by what they read for the last C
`__amd_rocclr_copyBuffer` dispatches, plain copies of K KiB each (--calib-kib):
factor = K / mean counter of those copies; the scaled figure is in bytes.
--skip-idle leaves out, per kernel and counter, dispatches that read under 5 %
of that kernel's largest value of the counter: stage launches that found the
end time already reached and returned at once.  --grid N keeps only the
dispatches of N threads (one problem size of a run that launches several).
"""
from __future__ import annotations

import argparse
import csv
import re
import sys
from collections import OrderedDict


def read_dispatches(path: str):
    """[(kernel name, start ns, end ns)] in dispatch order."""
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            name = r.get("Kernel_Name") or r.get("KernelName") or r.get("Name")
            t0 = r.get("Start_Timestamp") or r.get("BeginNs") or r.get("Start")
            t1 = r.get("End_Timestamp") or r.get("EndNs") or r.get("End")
            did = r.get("Dispatch_Id") or r.get("Index") or len(rows)
            if name is None or t0 is None or t1 is None:
                raise ValueError(f"{path}: no kernel name / timestamp columns")
            rows.append((int(did), name, int(t0), int(t1)))
    rows.sort(key=lambda x: (x[2], x[0]))
    return [(n, a, b) for _, n, a, b in rows]


def stats(durs):
    n = len(durs)
    tot = sum(durs)
    return {"calls": n, "total_ns": tot, "avg_ns": tot / n if n else 0.0, "min_ns": min(durs) if n else 0,
            "max_ns": max(durs) if n else 0}


def table(groups) -> str:
    total = sum(sum(d) for d in groups.values()) or 1
    out = ["| kernel | calls | total ns | average ns | min ns | max ns | % |", "|---|---|---|---|---|---|---|"]
    for name, durs in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        s = stats(durs)
        short = name if len(name) <= 90 else name[:87] + "..."
        out.append(f"| `{short}` | {s['calls']} | {s['total_ns']} | {s['avg_ns']:.0f} | {s['min_ns']} | "
                   f"{s['max_ns']} | {100.0 * s['total_ns'] / total:.2f} |")
    return "\n".join(out)


COPY_KERNEL = "__amd_rocclr_copyBuffer"


def read_counters(paths):
    """{kernel name: {"shape": {...}, "dispatches": {id: {counter: value}}}}
    over every counter-collection CSV in `paths` (counters of one dispatch
    summed over any per-XCD / per-SE rows rocprofv3 splits them into)."""
    out = OrderedDict()
    for path in paths:
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                name = r["Kernel_Name"]
                k = out.setdefault(name, {"shape": {}, "dispatches": OrderedDict()})
                k["shape"] = {key: r.get(key) for key in ("Grid_Size", "Workgroup_Size", "VGPR_Count",
                                                          "Accum_VGPR_Count", "LDS_Block_Size", "Scratch_Size")}
                did = (path, int(r["Dispatch_Id"]))
                d = k["dispatches"].setdefault(did, {"_grid": int(r["Grid_Size"])})
                c = r["Counter_Name"]
                d[c] = d.get(c, 0.0) + float(r["Counter_Value"])
    return out


def copy_factors(kernels, calib_kib: float, calib_n: int):
    """{counter: factor} for FETCH_SIZE / WRITE_SIZE from the last calib_n
    copy dispatches of each counter (the calibration copies run last)."""
    fac = {}
    copies = kernels.get(COPY_KERNEL)
    if not copies or calib_kib <= 0:
        return fac
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        vals = [d[ctr] for d in copies["dispatches"].values() if ctr in d][-calib_n:]
        if vals and sum(vals) > 0:
            fac[ctr] = calib_kib / (sum(vals) / len(vals))
    return fac


def pmc_report(paths, kernel_rx=None, points=0, calib_kib=0.0, calib_n=3, skip_idle=False, grid=0) -> str:
    kernels = read_counters(paths)
    fac = copy_factors(kernels, calib_kib, calib_n)
    lines = []
    for ctr, f in fac.items():
        lines.append(f"copy calibration {ctr}: last {calib_n} copies of {calib_kib:.0f} KiB -> factor {f:.5f}")
    rx = re.compile(kernel_rx) if kernel_rx else None
    for name, k in kernels.items():
        if name == COPY_KERNEL or (rx and not rx.search(name)):
            continue
        disp = [d for d in k["dispatches"].values() if not grid or d["_grid"] == grid]
        if not disp:
            continue
        sh = dict(k["shape"], Grid_Size=disp[-1]["_grid"])
        lines.append(f"\n{name}")
        lines.append(f"  dispatches {len(disp)}; grid {sh['Grid_Size']} threads, workgroup {sh['Workgroup_Size']}, "
                     f"VGPR {sh['VGPR_Count']} (+{sh['Accum_VGPR_Count']} AGPR), LDS {sh['LDS_Block_Size']} B, "
                     f"scratch {sh['Scratch_Size']} B")
        ctrs = sorted({c for d in disp for c in d if c != "_grid"})
        for c in ctrs:
            vals = [d[c] for d in disp if c in d]
            if skip_idle:
                top = max(vals)
                vals = [v for v in vals if v >= 0.05 * top]
            mean = sum(vals) / len(vals)
            txt = f"  {c}: mean {mean:.6g} per dispatch (min {min(vals):.6g}, max {max(vals):.6g}, n {len(vals)})"
            if c in fac:
                txt += f" -> {mean * fac[c] * 1024:.6g} bytes"
            elif points and c.startswith("SQ_INSTS"):
                txt += f" = {mean / (points / 64.0):.1f} per 64 points"
            lines.append(txt)
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("files", nargs="+")
    ap.add_argument("--pmc", action="store_true", help="files are counter-collection CSVs")
    ap.add_argument("--kernel", help="regular expression selecting kernels (of --last, or of the --pmc report)")
    ap.add_argument("--last", type=int, default=0, help="only the last N dispatches of --kernel")
    ap.add_argument("--points", type=int, default=0, help="grid points per dispatch (--pmc)")
    ap.add_argument("--calib-kib", type=float, default=0.0, help="KiB each calibration copy moves (--pmc)")
    ap.add_argument("--calib-n", type=int, default=3, help="number of calibration copies, the last ones (--pmc)")
    ap.add_argument("--skip-idle", action="store_true", help="drop dispatches under 5%% of the kernel's largest value")
    ap.add_argument("--grid", type=int, default=0, help="only dispatches of this grid size in threads (--pmc)")
    args = ap.parse_args(argv)
    if args.pmc:
        print(pmc_report(args.files, args.kernel, args.points, args.calib_kib, args.calib_n, args.skip_idle, args.grid))
        return 0
    if len(args.files) != 1:
        ap.error("a kernel trace summary reads one file")
    disp = read_dispatches(args.files[0])
    groups = OrderedDict()
    for name, a, b in disp:
        groups.setdefault(name, []).append(b - a)
    print(f"All dispatches ({len(disp)}):\n")
    print(table(groups))
    if args.kernel:
        rx = re.compile(args.kernel)
        sel = [(n, a, b) for n, a, b in disp if rx.search(n)]
        if args.last:
            sel = sel[-args.last:]
        g = OrderedDict()
        for name, a, b in sel:
            g.setdefault(name, []).append(b - a)
        s = stats([b - a for _, a, b in sel])
        span = (sel[-1][2] - sel[0][1]) if sel else 0
        print(f"\nLast {len(sel)} dispatches matching `{args.kernel}`:\n")
        print(table(g))
        print(f"\nselected: {s['calls']} dispatches, average {s['avg_ns']:.0f} ns, "
              f"first start to last end {span} ns ({span / max(1, s['calls']):.0f} ns per dispatch)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
