"""Measurement studies of the non-headline configurations.

    python -m multigpu_advectiondiffusion_amd.studies batched   [--steps 200]
    python -m multigpu_advectiondiffusion_amd.studies batched_trace [--steps 20]
    python -m multigpu_advectiondiffusion_amd.studies fp32      [--n 512]
    python -m multigpu_advectiondiffusion_amd.studies adjoint   [--n 128]
    python -m multigpu_advectiondiffusion_amd.studies reference
    python -m multigpu_advectiondiffusion_amd.studies heat
    python -m multigpu_advectiondiffusion_amd.studies slab
    python -m multigpu_advectiondiffusion_amd.studies weno7

batched: 64 independent periodic WENO5-Z (p=2, eps=1e-40) FP64 problems with
    edge lengths drawn seeded-uniformly from {32,48,64,96,128} (SplitMix64 seed
    4), each with its own Fourier initial field (seed 4+i) and its own dt from
    CFL 0.4, advanced `steps` steps (a) in one persistent launch per RK stage
    under a hipGraph-captured step, (b) problem by problem with the
    single-problem kernel, each size autotuned and each step a replayed
    hipGraph, and (c) problem by problem with eager launches of the default
    launch shape.  Reports the throughputs (points * steps / s) and the
    ratios of (a) to (b) and to (c).
fp32: the same Burgers setup with WENO5-Z at n^3, run to t = 0.5 once in FP64
    and once with FP32 storage (FP64 split fluxes, smoothness indicators and
    weights); reports the L1 (mean |diff|) and L-infinity error of the FP32
    result against the FP64 one.
adjoint: J = 1/2 |u(T) - u_target|^2 for WENO5-Z (p=2) Burgers at n^3 FP64,
    u0 the seed-5 Fourier field and u_target the forward solution from u0 +
    1e-2 * uniform(-1,1) noise (numpy seed 5).  Cost: wall time of the
    gradient (forward with sqrt(N) checkpoints + recompute + reverse) over
    the plain forward run, on the run of 400 steps.  Accuracy: the gradient
    along 16 seeded unit directions against central finite differences
    (formed point-wise) at T = 0.02 (before the first shock, which forms
    near t = 0.03-0.04 for this field), 0.2 and 0.5, for h in
    {1e-5, 3e-6, 2e-6, 1e-6, 5e-7, 3e-7, 1e-7}, and
    against the Richardson extrapolation (4 D(h/2) - D(h))/3 at h = 2e-6 and
    1e-6; per direction the relative error |FD - AD| / |AD|.
reference: the reference's own viscous Burgers runs (512^3 and 1000x1000x200
    to tEnd 0.1, 1601x986x35 to 0.4) with the fused reference-step kernel, against
    the kernel seconds the reference published for them, plus the drop-in
    sequence on the same GPU for a few steps.
slab: the headline WENO5-Z (p=2) FP64 run on z slabs with 3 ghost planes:
    one GPU's share of an 8-way 256^3 decomposition (one 256x256x32 slab
    whose exchange wraps onto itself) and all 8 slabs of 256^3 on one GPU,
    against the unsplit 256^3 run; each at least 30 hipGraph-replayed steps
    and 0.2 s after 0.3 s of warm-up steps.  The exchange copies stand in for
    the per-stage neighbour send/recv.
weno7: the WENO7 Burgers prototype (LFWENO7FDM3d.m: Gaussian initial field,
    CFL 0.4, low-storage RK45) at 256^3 FP64, at least 10 steps and 0.2 s
    after 0.3 s of warm-up steps, hipGraph-replayed; the prototype's own 100^3
    run to tEnd 0.4 with its step count and wall time.
The GPU's clock rises over the first ~50 ms of sustained work (stage launches
shorten by up to 30 %), so every timed region above follows warm-up work of
at least 0.3 s.
Each prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import hip, solvers

WARM_S = 0.3  # warm-up wall time before a timed region
MIN_TIMED_S = 0.2  # shortest timed region of the slab and WENO7 studies


def warm(step, seconds: float = WARM_S) -> int:
    """Runs step() (then a device synchronisation) until `seconds` of wall
    time have passed: the GPU clock ramps up over the first ~50 ms of
    sustained work.  Returns the number of calls."""
    import time

    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        step()
        hip.synchronize()
        n += 1
    return n


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
    e0, e1 = hip.Event(), hip.Event()

    def timed(fn):
        b.set_initial(u0s, tend=float("inf"), kchunk=0)
        fn(2)  # builds the graph once
        warm(lambda: fn(4))
        b.set_initial(u0s, tend=float("inf"), kchunk=b.batch.kchunk)
        hip.synchronize()
        e0.record(b.stream)
        fn(steps)
        e1.record(b.stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / 1e3

    # one by one as a plain caller would run them: default launch shape, eager launches
    for n in sorted(set(sizes)):
        solvers.set_launch(scheme, n, 0, 0)
    t_eager = timed(lambda k: b.steps_serial(k, graph=False))
    sol_eager = b.solutions()
    # one by one at its best: each size autotuned, each step a replayed hipGraph
    tuned = {}
    for n in sorted(set(sizes)):
        s = solvers.PeriodicBurgers(n, scheme=scheme, stream=b.stream)
        s.set_initial(u0s[sizes.index(n)])
        tuned[n] = solvers.autotune(s)
        del s
    t_serial = timed(lambda k: b.steps_serial(k, graph=True))
    sol_serial = b.solutions()
    t_batch = timed(lambda k: b.steps(k, graph=True))
    sol_batch = b.solutions()
    identical = all(np.array_equal(x, y) and np.array_equal(x, z) for x, y, z in zip(sol_batch, sol_serial, sol_eager))
    pts = b.npts
    return {
        "study": "batched",
        "problems": nprob,
        "sizes": {str(n): sizes.count(n) for n in sorted(set(sizes))},
        "points": pts,
        "steps": steps,
        "batched_gpts": pts * steps / t_batch / 1e9,
        "one_by_one_gpts": pts * steps / t_serial / 1e9,
        "one_by_one_eager_gpts": pts * steps / t_eager / 1e9,
        "speedup": t_serial / t_batch,
        "speedup_vs_eager": t_eager / t_batch,
        "batched_s": t_batch,
        "one_by_one_s": t_serial,
        "one_by_one_eager_s": t_eager,
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


def adjoint(n: int = 128, steps: int = 400, ndir: int = 16, h: float = 1e-6,
            hs=(1e-5, 3e-6, 1e-6, 3e-7, 1e-7)) -> dict:
    u0 = solvers.ic_fourier(n, seed=5)
    noise = np.random.default_rng(5).uniform(-1.0, 1.0, size=u0.shape)
    out = {"study": "adjoint", "grid": n, "scheme": "WENO5-Z p=2 eps=1e-40", "fd_h": h, "directions": ndir}

    def forward(u, tend, stream=None):
        s = solvers.PeriodicBurgers(n, scheme="Z2", stream=stream)
        s.set_initial(u, tend=tend)
        st = s.run(batch=64)
        return s.solution(), st

    # cost on the `steps`-step run: find its end time, then time both paths
    s = solvers.PeriodicBurgers(n, scheme="Z2")
    s.set_initial(u0, tend=float("inf"))
    s.steps(steps)
    T = s.state().t
    target, _ = forward(u0 + 1e-2 * noise, T)
    e0, e1 = hip.Event(), hip.Event()

    def timed(fn):
        hip.synchronize()
        e0.record(s.stream)
        r = fn()
        e1.record(s.stream)
        e1.synchronize()
        return r, e0.elapsed_ms(e1) / 1e3

    def fwd():
        s.set_initial(u0, tend=T)
        return s.run(batch=steps + 8)

    fwd()  # warm-up
    _, t_fwd = timed(fwd)
    out.update({"cost_T": T, "forward_s": t_fwd})
    # sqrt(N) checkpoints (the default), then every stage input kept (no recompute)
    for name, every in (("sqrtN", None), ("store_all", steps + 8)):
        adj = solvers.BurgersAdjoint(n, scheme="Z2", max_steps=steps + 8, every=every, stream=s.stream)
        adj(u0, target, T)  # warm-up
        (J, g, nst), t_adj = timed(lambda: adj(u0, target, T))
        out[name] = {"checkpoint_every": adj.every, "adjoint_s": t_adj, "cost_ratio": t_adj / t_fwd,
                     "workspace_gb": adj.workspace.nbytes / 1e9, "steps": nst}
        del adj
    out["cost_steps"] = out["sqrtN"]["steps"]
    out["cost_ratio"] = out["sqrtN"]["cost_ratio"]
    adj = solvers.BurgersAdjoint(n, scheme="Z2", max_steps=steps + 8, stream=s.stream)
    rng = np.random.default_rng(2026)
    # t = 0.02 is before the first shock forms (t ~ 0.03-0.04 for this field:
    # steepest initial slope ~25), 0.2 and 0.5 after
    for tend in (0.02, 0.2, 0.5):
        tg, _ = forward(u0 + 1e-2 * noise, tend)
        J, g, nst = adj(u0, tg, tend)
        gnorm = float(np.linalg.norm(g))
        res = {"steps": nst, "J": J, "grad_norm": gnorm}
        dirs = []
        for _ in range(ndir):
            v = rng.standard_normal(u0.shape)
            v /= np.linalg.norm(v)
            dirs.append(v)
        ad = np.array([float(np.sum(g * v)) for v in dirs])
        D = {}
        for hh in sorted(set(hs) | {2 * h, h, h / 2}, reverse=True):
            fd = []
            for v in dirs:
                up, _ = forward(u0 + hh * v, tend)
                um, _ = forward(u0 - hh * v, tend)
                # J(u+) - J(u-) = 1/2 sum (u+ - u-)(u+ + u- - 2 target), formed point-wise
                fd.append(float(np.sum(0.5 * (up - um) * (up + um - 2 * tg)) / (2 * hh)))
            D[hh] = np.array(fd)

        def errs(est):
            rel = np.abs(est - ad) / np.abs(ad)
            return {"max_rel_err": float(np.max(rel)), "median_rel_err": float(np.median(rel)),
                    "max_err_over_grad_norm": float(np.max(np.abs(est - ad)) / gnorm)}

        res["directional_derivative_median_abs"] = float(np.median(np.abs(ad)))
        res["central"] = {f"h{hh:g}": errs(D[hh]) for hh in sorted(D, reverse=True)}
        # Richardson: (4 D(h/2) - D(h)) / 3 cancels the h^2 term of the central difference
        rich = {hh: (4 * D[hh / 2] - D[hh]) / 3 for hh in (2 * h, h)}
        res["richardson"] = {f"h{hh:g}": errs(est) for hh, est in rich.items()}
        res[f"h{h:g}"] = res["central"][f"h{h:g}"]
        # every direction: the adjoint's value and the relative error of each estimate
        res["per_direction"] = {
            "adjoint": [float(x) for x in ad],
            **{f"rel_err_central_h{hh:g}": [float(x) for x in np.abs(D[hh] - ad) / np.abs(ad)]
               for hh in sorted(D, reverse=True)},
            **{f"rel_err_richardson_h{hh:g}": [float(x) for x in np.abs(est - ad) / np.abs(ad)]
               for hh, est in rich.items()},
        }
        out[f"T{tend}"] = res
    return out


# The reference's own Burgers runs (SingleGPU/Burgers3d_WENO5/Run.m:1-51):
# viscous (C = 1e-5) local-LF WENO5-Z p=1 eps=1e-6, frozen 3-cell ghosts,
# Gaussian initial condition (Call_Init3d ic 3), CFL 0.3, L = W = H = 2,
# dt = min(CFL dx, 0.9/(2C sum 1/dx^2)), the last step clipped to tEnd
# (main.cpp:128-136); per run the published kernel seconds and RK3 step
# count.  Run.m keeps tEnd = 0.1, which gives the 86 and 167 steps printed for
# the first two grids; the 1067 steps printed for 1601x986x35 are tEnd = 0.4.
REFERENCE_CASES = [((512, 512, 512), 39.37, 86, 0.1), ((1000, 1000, 200), 247.54, 167, 0.1),
                   ((1601, 986, 35), 563.49, 1067, 0.4)]


def reference_cases(cases=REFERENCE_CASES, dropin_steps: int = 3) -> dict:
    from . import fileio
    from . import reference_api as R
    from .hip import DeviceArray

    out = {"study": "reference_cases", "scheme": "WENO5-Z p=1 eps=1e-6, local LF, 13-point viscous C=1e-5",
           "cases": []}
    e0, e1 = hip.Event(), hip.Event()
    for (nx, ny, nz), published_s, published_steps, tend in cases:
        L, C, cfl = 2.0, 1e-5, 0.3
        dx, dy, dz = L / (nx - 1), L / (ny - 1), L / (nz - 1)
        kc = (C / (12 * dx * dx), C / (12 * dy * dy), C / (12 * dz * dz))
        R.CopyToConstantMemory(*kc)
        dt = min(cfl * dx / 1.0, 1.0 / (2 * C * (1 / dx / dx + 1 / dy / dy + 1 / dz / dz)) * 0.9)
        t, dts = 0.0, []
        while t < tend:
            d = tend - t if t + dt > tend else dt
            dts.append(d)
            t += d
        u0 = fileio.init3d_burgers(3, (nx, ny, nz), dx, dy, dz)
        s = solvers.ReferenceBurgers((nx, ny, nz), dx, dy, dz, scheme="Z1", viscous=True)
        s.set_initial(u0)
        s.steps(1, dt)  # graph capture
        warm(lambda: s.steps(2, dt))
        s.set_initial(u0)
        hip.synchronize()
        e0.record(s.stream)
        s.steps(len(dts) - 1, dt)
        s.steps(1, dts[-1])
        e1.record(s.stream)
        e1.synchronize()
        t_fused = e0.elapsed_ms(e1) / 1e3
        fused = s.solution()
        del s
        npts = nx * ny * nz
        case = {"grid": [nx, ny, nz], "steps": len(dts), "fused_s": t_fused,
                "fused_gpts": npts * len(dts) / t_fused / 1e9,
                "published_kernel_s": published_s, "published_steps": published_steps,
                "published_gpts": npts * published_steps / published_s / 1e9,
                "speedup_vs_published": (npts * len(dts) / t_fused) / (npts * published_steps / published_s),
                "finite": bool(np.isfinite(fused).all())}
        if dropin_steps:
            # the drop-in sequence (5 launches and a copy per stage) on the same GPU, a few steps
            R.set_adv_scheme()
            du, duo, dLu = DeviceArray.from_host(u0), DeviceArray(u0.shape), DeviceArray(u0.shape)
            dLu.zero()
            warm(lambda: R.burgers_reference_step(du, duo, dLu, nx, ny, nz, dx, dy, dz, dt))
            hip.synchronize()
            e0.record(None)
            for _ in range(dropin_steps):
                R.burgers_reference_step(du, duo, dLu, nx, ny, nz, dx, dy, dz, dt)
            e1.record(None)
            e1.synchronize()
            t_drop = e0.elapsed_ms(e1) / 1e3
            case.update({"dropin_steps": dropin_steps, "dropin_gpts": npts * dropin_steps / t_drop / 1e9,
                         "fused_vs_dropin": (npts * len(dts) / t_fused) / (npts * dropin_steps / t_drop)})
            del du, duo, dLu
        out["cases"].append(case)
    return out


def heat_cases(dropin_steps: int = 5) -> dict:
    """Heat: the 7-point forward-Euler step (64^3, 100 steps, the parity
    configuration; 512^3, HBM-sized) and the reference's own 13-point SSP-RK3
    run (SingleGPU/Diffusion3d_baselineCode/Run.m:1-25: 201x199x200, C = 1,
    L = 10, 605 steps of dt = 0.9/(2C sum 1/dx^2), Call_Init3D ic 3 at
    t0 = 0.1) fused, against its published kernel seconds, with the
    corrected CalcError norms at the final time, and the same step at 512^3.
    A plain device copy of a 512^3 field gives the bandwidth a streaming
    kernel can reach."""
    from . import fileio
    from . import reference_api as R
    from .hip import DeviceArray

    out = {"study": "heat"}
    e0, e1 = hip.Event(), hip.Event()

    def timed(stream, fn):
        hip.synchronize()
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / 1e3

    # what a plain device-to-device copy of one 512^3 field moves (read + write)
    a, b = DeviceArray((512, 512, 512)), DeviceArray((512, 512, 512))
    warm(lambda: b.copy_from(a))
    t = timed(None, lambda: [b.copy_from(a) for _ in range(10)])
    out["device_copy_512"] = {"GBs": 2 * 8.0 * 512 ** 3 * 10 / t / 1e9}
    del a, b
    for n, steps in ((64, 100), (512, 20)):
        h = solvers.Heat7(n)
        h.set_initial(solvers.ic_gauss_noise(n, seed=1))
        h.steps(steps)  # graph capture
        warm(lambda: h.steps(steps))
        t = timed(h.stream, lambda: h.steps(steps))
        gbs = 16.0 * n ** 3 * steps / t / 1e9
        out[f"heat7_fe_{n}"] = {"steps": steps, "gpts": n ** 3 * steps / t / 1e9, "algorithmic_GBs": gbs,
                                "hbm_frac": gbs / 8000.0}
        del h
    nx, ny, nz, C, L, iters, published_s = 201, 199, 200, 1.0, 10.0, 605, 5.81
    dx, dy, dz = L / (nx - 1), L / (ny - 1), L / (nz - 1)
    dt = 1 / (2 * C * (1 / dx / dx + 1 / dy / dy + 1 / dz / dz)) * 0.9
    kc = (C / (12 * dx * dx), C / (12 * dy * dy), C / (12 * dz * dz))
    R.CopyToConstantMemory(*kc)
    u0 = fileio.init3d_heat(3, (nx, ny, nz), dx, dy, dz)
    h = solvers.ReferenceHeat((nx, ny, nz))
    h.set_initial(u0)
    h.steps(1, dt)  # graph capture
    warm(lambda: h.steps(20, dt))
    h.set_initial(u0)
    t = timed(h.stream, lambda: h.steps(iters, dt))
    u = h.solution()
    npts = nx * ny * nz
    l1, l2, linf = fileio.heat_error_norms(u, 0.1 + dt * iters, dx, dy, dz)
    case = {"grid": [nx, ny, nz], "steps": iters, "fused_s": t, "fused_gpts": npts * iters / t / 1e9,
            "algorithmic_GBs": 64.0 * npts * iters / t / 1e9, "published_kernel_s": published_s,
            "published_gpts": npts * iters / published_s / 1e9, "speedup_vs_published": published_s / t,
            "final_time": 0.1 + dt * iters, "error_l1": l1, "error_l2": l2, "error_linf": linf}
    du, duo, dLu = DeviceArray.from_host(u0), DeviceArray(u0.shape), DeviceArray(u0.shape)
    dLu.zero()
    warm(lambda: R.heat_reference_step(du, duo, dLu, nx, ny, nz, dt))
    td = timed(None, lambda: [R.heat_reference_step(du, duo, dLu, nx, ny, nz, dt) for _ in range(dropin_steps)])
    case.update({"dropin_gpts": npts * dropin_steps / td / 1e9,
                 "fused_vs_dropin": (npts * iters / t) / (npts * dropin_steps / td)})
    out["reference_heat13_rk3"] = case
    del h, du, duo, dLu
    # the same fused step on a 512^3 grid (bandwidth once launches are long;
    # the timed call includes its two initial field copies, ~2 %)
    n, steps = 512, 20
    h = solvers.ReferenceHeat((n, n, n))
    h.set_initial(solvers.ic_gauss_noise(n, seed=1))
    h.steps(1, 1e-6)  # graph capture
    warm(lambda: h.steps(2, 1e-6))
    t = timed(h.stream, lambda: h.steps(steps, 1e-6))
    out["heat13_rk3_512"] = {"steps": steps, "gpts": n ** 3 * steps / t / 1e9,
                             "algorithmic_GBs": 64.0 * n ** 3 * steps / t / 1e9,
                             "hbm_frac": 64.0 * n ** 3 * steps / t / 1e9 / 8000.0}
    return out


def slab_cases(steps: int = 30) -> dict:
    e0, e1 = hip.Event(), hip.Event()

    def rate(solver, npts):
        solver.set_initial(solvers.ic_fourier((solver.nx, solver.ny, solver.nz), seed=3))
        calls = warm(lambda: solver.steps(5))
        # at least `steps` steps and MIN_TIMED_S (the warm-up ran 5 * calls steps in WARM_S)
        k = max(steps, int(5 * calls * MIN_TIMED_S / WARM_S) + 1)
        hip.synchronize()
        e0.record(solver.stream)
        solver.steps(k)
        e1.record(solver.stream)
        e1.synchronize()
        return npts * k / (e0.elapsed_ms(e1) / 1e3) / 1e9

    out = {"study": "slab", "scheme": "WENO5-Z p=2 eps=1e-40", "steps": steps}
    one = solvers.SlabBurgers((256, 256, 32), 1)
    out["one_slab_256x256x32_gpts"] = rate(one, 256 * 256 * 32)
    # the slab launch reads the launch table entry of the unsplit grid of its
    # shape: tune that entry, then time the slab again
    p = solvers.PeriodicBurgers((256, 256, 32), stream=one.stream)
    p.set_initial(solvers.ic_fourier((256, 256, 32), seed=3))
    out["one_slab_tuned_tile_kchunk"] = list(solvers.autotune(p))
    del p
    out["one_slab_256x256x32_tuned_gpts"] = rate(one, 256 * 256 * 32)
    del one
    eight = solvers.SlabBurgers((256, 256, 256), 8)
    out["eight_slabs_256_gpts"] = rate(eight, 256 ** 3)
    del eight
    whole = solvers.PeriodicBurgers(256)
    out["unsplit_256_gpts"] = rate(whole, 256 ** 3)
    out["one_slab_vs_unsplit"] = out["one_slab_256x256x32_gpts"] / out["unsplit_256_gpts"]
    out["one_slab_tuned_vs_unsplit"] = out["one_slab_256x256x32_tuned_gpts"] / out["unsplit_256_gpts"]
    out["eight_slabs_vs_unsplit"] = out["eight_slabs_256_gpts"] / out["unsplit_256_gpts"]
    return out


def weno7_cases(steps: int = 10) -> dict:
    e0, e1 = hip.Event(), hip.Event()
    out = {"study": "weno7", "scheme": "WENO7 (Balsara-Shu forms, eps 1e-6), local LF, LSRK(5,4)"}
    s = solvers.Weno7Burgers(256)
    s.set_initial(s.gaussian(), tend=float("inf"))
    calls = warm(lambda: s.steps(2))
    k = max(steps, int(2 * calls * MIN_TIMED_S / WARM_S) + 1)
    hip.synchronize()
    e0.record(s.stream)
    s.steps(k)
    e1.record(s.stream)
    e1.synchronize()
    t = e0.elapsed_ms(e1) / 1e3
    out["grid_256"] = {"steps": k, "gpts": 256 ** 3 * k / t / 1e9, "ms_per_step": 1e3 * t / k,
                       "stage_launches_per_step": 5}
    del s
    s = solvers.Weno7Burgers(100, cfl=0.4)
    s.set_initial(s.gaussian(), tend=0.4)
    hip.synchronize()
    e0.record(s.stream)
    s.steps(400)  # more than the run needs; steps after tEnd are no-ops on the device
    e1.record(s.stream)
    e1.synchronize()
    tt, dt, n = s.state()
    out["prototype_100_to_0.4"] = {"steps": n, "t": tt, "wall_s": e0.elapsed_ms(e1) / 1e3,
                                   "max_u": float(np.max(np.abs(s.solution())))}
    return out


def adjoint_trace(n: int = 128, steps: int = 100, copy_calib: bool = False) -> dict:
    """One store-all adjoint of a `steps`-step n^3 WENO5-Z run: the workload
    of the adjoint kernel trace (forward stages, then per stage k_vjp_xy,
    k_vjp_z and k_vjp_tail).  copy_calib: afterwards, three device copies of
    one n^3 FP64 field (the known-size reference for HBM counters)."""
    u0 = solvers.ic_fourier(n, seed=5)
    s = solvers.PeriodicBurgers(n, scheme="Z2")
    s.set_initial(u0, tend=float("inf"))
    s.steps(steps)
    T = s.state().t
    adj = solvers.BurgersAdjoint(n, scheme="Z2", max_steps=steps + 8, every=steps + 8, stream=s.stream)
    J, _, nst = adj(u0, 0.9 * u0, T)
    if copy_calib:
        for _ in range(3):
            s.w1.copy_from(s.u, s.stream)
        s.stream.synchronize()
    return {"study": "adjoint_trace", "grid": n, "steps": nst, "J": J}


def batched_trace(steps: int = 20, nprob: int = 64, scheme: str = "Z2") -> dict:
    """The batched path alone on the batched study's 64 problems: the
    autotuner's timing runs, then `steps` hipGraph-replayed steps with the
    work list it picked (the workload of the batched kernel trace)."""
    sizes = batch_sizes(nprob)
    u0s = [solvers.ic_fourier(n, seed=4 + i) for i, n in enumerate(sizes)]
    b = solvers.BurgersBatch(sizes, scheme=scheme)
    b.set_initial(u0s, tend=float("inf"), kchunk=0)
    b.steps(steps, graph=True)
    b.stream.synchronize()
    return {"study": "batched_trace", "problems": nprob, "points": b.npts, "steps": steps,
            "kchunk_batched": b.batch.kchunk, "grid": b.batch.grid, "work_items": b.batch.nitems}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("study",
                    choices=["batched", "batched_trace", "fp32", "adjoint", "adjoint_trace", "reference", "heat", "slab",
                             "weno7"])
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--copy-calib", action="store_true",
                    help="adjoint_trace: three device copies of one field afterwards (counter calibration)")
    args = ap.parse_args(argv)
    if args.study == "batched":
        r = batched(args.steps)
    elif args.study == "batched_trace":
        r = batched_trace(args.steps if args.steps != 200 else 20)
    elif args.study == "reference":
        r = reference_cases()
    elif args.study == "heat":
        r = heat_cases()
    elif args.study == "slab":
        r = slab_cases()
    elif args.study == "weno7":
        r = weno7_cases()
    elif args.study == "adjoint":
        r = adjoint(args.n if args.n != 512 else 128)
    elif args.study == "adjoint_trace":
        r = adjoint_trace(args.n if args.n != 512 else 128, args.steps if args.steps != 200 else 100, args.copy_calib)
    else:
        r = fp32(args.n)
    print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
