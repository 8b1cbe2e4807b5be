"""Measurement studies of the non-headline configurations.

    python -m multigpu_advectiondiffusion_amd.studies batched   [--steps 200]
    python -m multigpu_advectiondiffusion_amd.studies fp32      [--n 512]

batched: 64 independent periodic WENO5-Z (p=2, eps=1e-40) FP64 problems with
    edge lengths drawn seeded-uniformly from {32,48,64,96,128} (SplitMix64 seed
    4), each with its own Fourier initial field (seed 4+i) and its own dt from
    CFL 0.4, advanced `steps` steps (a) in one persistent launch per RK stage
    under a hipGraph-captured step and (b) problem by problem with the
    single-problem kernel (each size autotuned, also under a hipGraph).
    Reports both throughputs (points * steps / s) and their ratio.
fp32: the same Burgers setup with WENO5-Z at n^3, run to t = 0.5 once in FP64
    and once with FP32 storage (FP64 split fluxes, smoothness indicators and
    weights); reports the L1 (mean |diff|) and L-infinity error of the FP32
    result against the FP64 one.
Each prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import hip, solvers


class SplitMix64:
    """Same generator as the library's host initial-condition code."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53


def batch_sizes(nprob: int = 64, seed: int = 4, choices=(32, 48, 64, 96, 128)):
    r = SplitMix64(seed)
    return [choices[int(r.uniform() * len(choices))] for _ in range(nprob)]


def batched(steps: int = 200, nprob: int = 64, scheme: str = "Z2") -> dict:
    sizes = batch_sizes(nprob)
    u0s = [solvers.ic_fourier(n, seed=4 + i) for i, n in enumerate(sizes)]
    b = solvers.BurgersBatch(sizes, scheme=scheme)
    # autotune the single-problem launch of each distinct size (one-by-one baseline)
    tuned = {}
    for n in sorted(set(sizes)):
        s = solvers.PeriodicBurgers(n, scheme=scheme, stream=b.stream)
        s.set_initial(u0s[sizes.index(n)])
        tuned[n] = solvers.autotune(s)
        del s
    e0, e1 = hip.Event(), hip.Event()

    def timed(fn):
        b.set_initial(u0s, tend=float("inf"), kchunk=0)
        fn(2)  # warm-up (also builds the graph once)
        b.set_initial(u0s, tend=float("inf"), kchunk=b.batch.kchunk)
        hip.synchronize()
        e0.record(b.stream)
        fn(steps)
        e1.record(b.stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / 1e3

    t_batch = timed(lambda k: b.steps(k, graph=True))
    sol_batch = b.solutions()
    t_serial = timed(lambda k: b.steps_serial(k, graph=True))
    sol_serial = b.solutions()
    identical = all(np.array_equal(x, y) for x, y in zip(sol_batch, sol_serial))
    pts = b.npts
    return {
        "study": "batched",
        "problems": nprob,
        "sizes": {str(n): sizes.count(n) for n in sorted(set(sizes))},
        "points": pts,
        "steps": steps,
        "batched_gpts": pts * steps / t_batch / 1e9,
        "one_by_one_gpts": pts * steps / t_serial / 1e9,
        "speedup": t_serial / t_batch,
        "batched_s": t_batch,
        "one_by_one_s": t_serial,
        "kchunk_batched": b.batch.kchunk,
        "work_items": b.batch.nitems,
        "tuned_single": {str(k): list(v) for k, v in tuned.items()},
        "results_identical": identical,
    }


def fp32(n: int = 512, tend: float = 0.5, seed: int = 3) -> dict:
    u0 = solvers.ic_fourier(n, seed=seed)
    out = {"study": "fp32_storage", "grid": n, "tend": tend, "scheme": "WENO5-Z p=2 eps=1e-40"}
    res = {}
    for prec in ("f64", "f32"):
        s = solvers.PeriodicBurgers(n, scheme="Z2", precision=prec)
        s.set_initial(u0, tend=tend)
        e0, e1 = hip.Event(), hip.Event()
        e0.record(s.stream)
        st = s.run(batch=64)
        e1.record(s.stream)
        e1.synchronize()
        res[prec] = (s.solution().astype(np.float64), st, e0.elapsed_ms(e1) / 1e3)
        del s
    d = res["f32"][0] - res["f64"][0]
    out.update({
        "steps_f64": res["f64"][1].steps, "steps_f32": res["f32"][1].steps,
        "l1": float(np.mean(np.abs(d))), "linf": float(np.max(np.abs(d))),
        "max_abs_f64": float(np.max(np.abs(res["f64"][0]))),
        "gpts_f64": n ** 3 * res["f64"][1].steps / res["f64"][2] / 1e9,
        "gpts_f32": n ** 3 * res["f32"][1].steps / res["f32"][2] / 1e9,
    })
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["batched", "fp32"])
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--n", type=int, default=512)
    args = ap.parse_args(argv)
    if args.study == "batched":
        r = batched(args.steps)
    else:
        r = fp32(args.n)
    print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
