"""Summaries of rocprofv3 kernel traces for the profiles/ notes.

    python -m multigpu_advectiondiffusion_amd.trace_summary TRACE_CSV
        [--kernel REGEX --last N] [--top K]

TRACE_CSV is the `*_kernel_trace.csv` that `rocprofv3 --kernel-trace
--output-format csv` writes (one row per dispatch with start and end
timestamps in ns).  Prints a markdown table of every kernel's dispatches
(calls, total, average, min, max) and, with --kernel/--last, the same figures
over only the last N dispatches of the kernels matching REGEX: for a
benchmark that runs W warm-up steps and then K timed steps of L launches,
--last K*L selects the timed launches, whose average is the one to set
beside the benchmark's own event timing.
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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--kernel", help="regular expression selecting the kernel(s) of --last")
    ap.add_argument("--last", type=int, default=0, help="only the last N dispatches of --kernel")
    args = ap.parse_args(argv)
    disp = read_dispatches(args.trace)
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
