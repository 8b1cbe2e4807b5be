"""Benchmark: periodic 3-D inviscid Burgers, WENO5-Z (p=2, eps=1e-40) with
global Lax-Friedrichs splitting and SSP-RK3, 256^3 FP64 on one MI355X.

A step is one full SSP-RK3 step (three fused stage launches) over the whole
grid; inputs are resident in HBM before the timed region starts.  Prints one
JSON line (the contract the harness reads).

    python bench.py [--steps K] [--warmup W] [--n 256] [--scheme Z2|JS]
                    [--no-cpu] [--quick] [--gpus 1]

It also reports the FP64 VALU issue roofline of the same launches (the
ceiling that binds; instruction counts from the committed counter profile).
After the timed region (and the CPU baseline) it runs the other measurements
the metric names, plus the reference's own runs, as extra fields:
  - batched: 64 uneven problems, one persistent launch per stage, against
    one-by-one launches (200 steps);
  - adjoint: cost ratio and finite-difference agreement (128^3, 400 steps,
    16 directions);
  - fp32_storage_512: FP32-storage error against FP64 (512^3 to t = 0.5);
  - reference_cases: the reference's own viscous Burgers runs with the fused
    reference step, against the kernel times it published;
  - burgers_js/z2_256_to_t0.5: the 256^3 runs to t = 0.5 (past shocks);
  - z_slabs: one 256x256x32 slab (one GPU's share of an 8-way 256^3
    decomposition) and 8 slabs on one GPU, against the unsplit grid;
  - weno7_256: the WENO7 prototype's scheme at 256^3.
--quick skips them.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpts/s and % HBM roofline, WENO5-Z 256³ FP64; adjoint/forward cost; batched vs 1-by-1"
HBM_PEAK_GBS = 8000.0
# Compulsory HBM traffic of one SSP-RK3 step with the fused stages, FP64:
# stage 1 reads u^n, writes u1 (16 B/pt); stages 2 and 3 read their input and
# u^n and write (24 B/pt each) -> 64 B/pt/step, 64/3 B/pt per stage launch.
BYTES_PER_PT_STEP = 64.0
LAUNCHES_PER_STEP = 3


def read_traffic(workload: str):
    """HBM bytes per stage launch from the committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", f"{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


# FP64 VALU issue peak: one wave64 FP64 instruction per CU per cycle (64 FMA
# lanes per CU, the 78.6 TFLOP/s vector rate), 256 CUs at 2.4 GHz.
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 2.4e9


def read_valu(workload: str):
    """VALU wave-instructions per stage launch from the committed rocprofv3
    --pmc summary (SQ_INSTS_VALU), or None."""
    path = os.path.join(ROOT, "profiles", f"{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"SQ_INSTS_VALU:\s*([0-9.eE+]+) per launch", open(path).read())
    return float(m.group(1)) if m else None


def run_to_half(scheme: str, n: int = 256, seed: int = 3) -> dict:
    """The 256^3 periodic run from the seeded Fourier field to t = 0.5 (past
    shock formation), one hipGraph replay per step; steps after t = 0.5 are
    no-ops on the device, so the loop asks for a few more than needed."""
    from multigpu_advectiondiffusion_amd import hip, solvers

    s = solvers.PeriodicBurgers(n, scheme=scheme, cfl=0.4)
    s.set_initial(solvers.ic_fourier(n, seed=seed), t0=0.0, tend=0.5)
    s.steps(1)  # capture
    s.set_initial(solvers.ic_fourier(n, seed=seed), t0=0.0, tend=0.5)
    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    e0.record(s.stream)
    st = s.run(batch=64)
    e1.record(s.stream)
    e1.synchronize()
    wall = e0.elapsed_ms(e1) / 1e3
    return {"steps": st.steps, "t": st.t, "wall_s": round(wall, 4), "gpts": round(n ** 3 * st.steps / wall / 1e9, 3)}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n: int, scheme: str, seed: int):
    """The CPU oracle (OpenMP restatement of the reference routines), built
    -O3 -march=native for this host, timed on a bounded sample of the same
    workload: whole SSP-RK3 steps of the n^3 grid, as many as fit in about
    10 s (at most 20).  Threads are placed one per core, close together."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import numpy as np

    import oracle as O
    from multigpu_advectiondiffusion_amd import solvers

    kw = {"Z2": O.Z2_GLOBAL, "JS": O.JS_GLOBAL}[scheme]
    # the native build computes the same bits as the tested one
    a = solvers.ic_fourier(32, seed=seed)
    b = a.copy()
    O.burgers_step(a, 2.0 / 32, 0.4, 0.0, 10.0, **kw)
    O.burgers_step(b, 2.0 / 32, 0.4, 0.0, 10.0, native=True, **kw)
    assert np.array_equal(a, b), "native oracle build differs from the tested one"
    u = solvers.ic_fourier(n, seed=seed)
    steps, t = 0, 0.0
    t0 = time.perf_counter()
    while steps < 20 and (t < 10.0 or steps == 0):
        O.burgers_step(u, 2.0 / n, 0.4, 0.0, 10.0, native=True, **kw)
        steps += 1
        t = time.perf_counter() - t0
    assert np.isfinite(u).all()
    return {
        "value": round(n ** 3 * steps / t / 1e9, 5),
        "unit": "Gpts/s",
        "cores": O.num_threads(native=True),
        "kind": "port",
        "sample": f"{steps} SSP-RK3 steps of the {n}^3 {scheme} periodic grid in {t:.2f} s",
        "build": "gcc -O3 -march=native -fopenmp -ffp-contract=off (oracle/Makefile native)",
        "cpu_model": _cpu_model(),
        "host_cpus": os.cpu_count(),
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the stage launches speed up over the first ~20 steps after the GPU has
    # been idle (420-540 us, then ~385 us per launch at 256^3): the default
    # warm-up runs into the steady state before the timed steps
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--scheme", default="Z2", choices=["Z2", "JS"])
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="skip the batched / adjoint / reference / fp32 / slab studies")
    ap.add_argument("--copy-calib", action="store_true",
                    help="after timing, run 3 device copies of one field (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU z-slab decomposition is a design, not a built path.", file=sys.stderr)
        return 2

    from multigpu_advectiondiffusion_amd import hip, solvers

    n = args.n
    s = solvers.PeriodicBurgers(n, scheme=args.scheme, cfl=0.4)
    u0 = solvers.ic_fourier(n, seed=args.seed)
    s.set_initial(u0, t0=0.0, tend=float("inf"))
    graph = not args.no_graph
    s.steps(args.warmup, graph=graph)
    s.stream.synchronize()

    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    e0.record(s.stream)
    s.steps(args.steps, graph=graph)
    e1.record(s.stream)
    e1.synchronize()
    hip.synchronize()
    ms = e0.elapsed_ms(e1)
    st = s.state()
    assert st.steps == args.warmup + args.steps

    npts = n ** 3
    sec = ms / 1e3
    gpts = npts * args.steps / sec / 1e9
    launch_s = sec / (args.steps * LAUNCHES_PER_STEP)
    bytes_per_launch = BYTES_PER_PT_STEP / LAUNCHES_PER_STEP * npts
    achieved = bytes_per_launch / launch_s / 1e9
    workload = f"burgers3d_weno5{args.scheme.lower()}_{n}"
    out = {
        "metric": METRIC,
        "value": round(gpts, 4),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: 0.5 + 8 seeded Fourier modes (SplitMix64 seed %d)" % args.seed,
        "config": {
            "workload": workload,
            "grid": [n, n, n],
            "scheme": {"Z2": "WENO5-Z p=2 eps=1e-40", "JS": "WENO5-JS eps=1e-6"}[args.scheme],
            "flux_split": "global Lax-Friedrichs alpha=max|u| per stage",
            "time": "SSP-RK3, dt=0.4*dx/max|u^n| computed on device",
            "boundary": "periodic on [-1,1)^3",
            "hipgraph": graph,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(workload),
            "kernel": "k_burgers_stage (one launch per SSP-RK3 stage)",
            "avg_launch_us": round(launch_s * 1e6, 2),
        },
        "cpu_baseline": None,
    }
    valu = read_valu(workload) if (n == 256 and args.scheme == "Z2") else None
    if valu:
        # the ceiling that binds: FP64 VALU issue (the counts are per launch of
        # this workload, taken from the committed counter profile)
        rate = valu / launch_s
        out["valu_roofline"] = {"bound": "fp64 valu issue", "instructions_per_launch": valu,
                                "achieved": round(rate / 1e9, 1), "peak": round(VALU_PEAK_WAVE_INSTR_PER_S / 1e9, 1),
                                "unit": "G wave-instr/s", "frac": round(rate / VALU_PEAK_WAVE_INSTR_PER_S, 4)}
    if args.copy_calib:
        # plain device-to-device copies of one field (n^3 * 8 bytes each way),
        # the known-size reference for scaling FETCH_SIZE / WRITE_SIZE
        for _ in range(3):
            s.w1.copy_from(s.u, s.stream)
        s.stream.synchronize()
    print(f"timed {args.steps} steps: {ms / args.steps:.3f} ms/step, {gpts:.3f} Gpts/s", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.scheme, args.seed)
        print("cpu baseline done", file=sys.stderr, flush=True)
    if not args.quick:
        from multigpu_advectiondiffusion_amd import studies

        b = studies.batched(steps=200)
        out["batched"] = {k: b[k] for k in ("speedup", "speedup_vs_eager", "batched_gpts", "one_by_one_gpts",
                                             "one_by_one_eager_gpts", "problems", "points", "steps",
                                             "results_identical")}
        print("batched study done", file=sys.stderr, flush=True)
        a = studies.adjoint(n=128)
        # per-direction |FD - AD| / |AD| over the 16 directions at three end
        # times (before shocks, and two after): central differences at
        # h = 1e-6, Richardson's (4 D(h/2) - D(h)) / 3 from h = 2e-6, and the
        # max / median over directions for every h of the sweep
        pick = ("max_rel_err", "median_rel_err", "max_err_over_grad_norm")
        fd = {}
        for T in ("T0.02", "T0.2", "T0.5"):
            fd[T] = {"steps": a[T]["steps"],
                     "central_h1e-6": {k: a[T]["central"]["h1e-06"][k] for k in pick},
                     "richardson_h2e-6": {k: a[T]["richardson"]["h2e-06"][k] for k in pick},
                     "max_rel_err_by_h": {h: v["max_rel_err"] for h, v in a[T]["central"].items()},
                     "median_rel_err_by_h": {h: v["median_rel_err"] for h, v in a[T]["central"].items()}}
        out["adjoint"] = {"cost_ratio": a["cost_ratio"], "steps": a["cost_steps"], "grid": 128,
                          "checkpoint_every": a["sqrtN"]["checkpoint_every"],
                          "cost_ratio_store_all": a["store_all"]["cost_ratio"], "fd": fd}
        print("adjoint study done", file=sys.stderr, flush=True)
        r = studies.reference_cases()
        out["reference_cases"] = [{k: c[k] for k in ("grid", "steps", "fused_gpts", "published_gpts",
                                                     "speedup_vs_published", "dropin_gpts", "fused_vs_dropin")}
                                  for c in r["cases"]]
        print("reference cases done", file=sys.stderr, flush=True)
        f = studies.fp32(n=512)
        out["fp32_storage_512"] = {k: f[k] for k in ("l1", "linf", "steps_f64", "gpts_f64", "gpts_f32")}
        print("fp32 study done", file=sys.stderr, flush=True)
        for sch in ("JS", "Z2"):
            out[f"burgers_{sch.lower()}_256_to_t0.5"] = run_to_half(sch)
        print("runs to t = 0.5 done", file=sys.stderr, flush=True)
        sl = studies.slab_cases()
        out["z_slabs"] = {k: sl[k] for k in ("one_slab_256x256x32_gpts", "one_slab_256x256x32_tuned_gpts",
                                             "one_slab_tuned_tile_kchunk", "eight_slabs_256_gpts", "unsplit_256_gpts",
                                             "one_slab_vs_unsplit", "one_slab_tuned_vs_unsplit",
                                             "eight_slabs_vs_unsplit")}
        print("slab study done", file=sys.stderr, flush=True)
        w7 = studies.weno7_cases()
        out["weno7_256"] = w7["grid_256"]
        print("weno7 study done", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
