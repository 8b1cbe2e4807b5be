"""Fused periodic Burgers stages on the GPU vs the CPU oracle.

Both paths evaluate the same IEEE binary64 expression trees (no contraction),
so the comparison is bit-for-bit; the north-star tolerance (1e-12 relative
L-infinity) is asserted as well so a failure reports its size.
"""
import numpy as np
import pytest

import oracle as O
from multigpu_advectiondiffusion_amd import solvers
from multigpu_advectiondiffusion_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu

ORACLE_SCHEME = {"Z2": O.Z2_GLOBAL, "JS": O.JS_GLOBAL, "Z1": dict(weno=O.WENO_Z, p=1, eps=1e-6, lf=O.LF_GLOBAL)}
TOL = 1e-12


def rel_linf(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def oracle_stage(u_in, u_n, stage, dt, d, scheme):
    alpha = O.maxabs(u_in)
    Lu = O.rhs_periodic(u_in, d, **ORACLE_SCHEME[scheme], alpha=alpha)
    out = u_in.copy()
    O.rk_stage_all(out, u_n, Lu, stage, dt)
    return out


@pytest.mark.parametrize("scheme", ["Z2", "JS", "Z1"])
@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 32, 24), (16, 48, 20)])
def test_stages_bitwise(hipdev, scheme, shape):
    nx, ny, nz = shape
    s = solvers.PeriodicBurgers(shape, scheme=scheme, cfl=0.4)
    u0 = solvers.ic_fourier(shape, seed=2)
    s.set_initial(u0, t0=0.0, tend=10.0)
    d = (s.dx, s.dy, s.dz)
    # stage by stage through the C ABI
    s.stage(1, s.u, s.u, s.w1)
    st = s.state()
    dt = 0.4 * s.dx / O.maxabs(u0)
    assert st.dt == dt
    u1 = oracle_stage(u0, u0, 1, dt, d, scheme)
    g1 = s.w1.download()
    assert rel_linf(g1, u1) <= TOL
    assert np.array_equal(g1, u1)
    s.stage(2, s.w1, s.u, s.w2)
    u2 = oracle_stage(u1, u0, 2, dt, d, scheme)
    g2 = s.w2.download()
    assert np.array_equal(g2, u2)
    s.stage(3, s.w2, s.u, s.u)
    u3 = oracle_stage(u2, u0, 3, dt, d, scheme)
    g3 = s.solution()
    assert np.array_equal(g3, u3)
    st = s.state()
    assert st.steps == 1 and st.t == dt
    assert st.amax == O.maxabs(u3)


# The stage kernels shorten the weight divisions only when max|u| <= 2^40 and
# eps is in range (2^-150..1 for Z, 2^-190..1 for JS); the first four cases
# fall outside and take the IEEE sequence, which must match the oracle as
# well.  Within range, the final WENO-Z division is shortened only for
# numerators of magnitude >= 2^-121: "z_tiny" (|u| ~ 1e-20, every numerator
# ~1e-40) takes the IEEE division on every node, "z_zero_half" (u = 0 on half
# the planes: zero numerators there, ordinary ones elsewhere) on some nodes
# of a wave and not others.
@pytest.mark.parametrize("case", ["z_eps_1e-95", "z_eps_1e-60", "js_eps_1e-100", "z_u_times_2^41", "z_tiny",
                                  "z_zero_half"])
def test_stages_ieee_division_path_bitwise(hipdev, case):
    from multigpu_advectiondiffusion_amd import _capi

    shape = (32, 48, 20)
    scheme = "JS" if case.startswith("js") else "Z2"
    w, p, eps = solvers.SCHEMES[scheme]
    u0 = solvers.ic_fourier(shape, seed=9)
    if case == "z_eps_1e-95":
        eps = 1e-95
    elif case == "z_eps_1e-60":
        eps = 1e-60
    elif case == "js_eps_1e-100":
        eps = 1e-100
    elif case == "z_tiny":
        u0 = u0 * 1e-20
    elif case == "z_zero_half":
        u0 = u0.copy()
        u0[: shape[2] // 2] = 0.0
    else:
        u0 = u0 * 2.0 ** 41
    orc = dict(weno=O.WENO_JS if scheme == "JS" else O.WENO_Z, p=p, eps=eps, lf=O.LF_GLOBAL)
    s = solvers.PeriodicBurgers(shape, scheme=scheme)
    s.scheme = _capi.Scheme(w, p, eps)
    s.set_initial(u0, tend=10.0)
    d = (s.dx, s.dy, s.dz)
    u_prev, stages = u0, ((1, s.u, s.w1), (2, s.w1, s.w2), (3, s.w2, s.u))
    for stage, src, dst in stages:
        s.stage(stage, src, s.u, dst)
        if stage == 1:
            dt = s.state().dt
        Lu = O.rhs_periodic(u_prev, d, **orc, alpha=O.maxabs(u_prev))
        ref = u_prev.copy()
        O.rk_stage_all(ref, u0, Lu, stage, dt)
        s.stream.synchronize()
        got = dst.download()
        assert np.isfinite(ref).all()
        assert np.array_equal(got, ref), (stage, np.argwhere(got != ref)[:3])
        u_prev = ref


@pytest.mark.parametrize("scheme", ["Z2", "JS"])
@pytest.mark.parametrize("graph", [False, True])
def test_run_to_half_bitwise(hipdev, scheme, graph):
    """32^3 run to t = 0.5 (past shock formation) against orc_burgers_run."""
    n = 32
    s = solvers.PeriodicBurgers(n, scheme=scheme, cfl=0.4)
    u0 = solvers.ic_fourier(n, seed=2)
    s.set_initial(u0, t0=0.0, tend=0.5)
    st = s.run(batch=8, graph=graph)
    g = s.solution()
    ref = u0.copy()
    steps, t = O.burgers_run(ref, 2.0 / n, 0.4, 0.0, 0.5, **ORACLE_SCHEME[scheme])
    assert st.steps == steps and st.t == t == 0.5
    assert rel_linf(g, ref) <= TOL
    assert np.array_equal(g, ref)


def test_graph_replays_match_eager(hipdev):
    """Several graph replays with host synchronisation in between equal eager steps."""
    n = 32
    u0 = solvers.ic_fourier(n, seed=7)
    a = solvers.PeriodicBurgers(n, scheme="Z2")
    b = solvers.PeriodicBurgers(n, scheme="Z2")
    a.set_initial(u0, tend=10.0)
    b.set_initial(u0, tend=10.0)
    for _ in range(4):
        a.steps(3, graph=True)
        a.stream.synchronize()
        b.steps(3, graph=False)
        assert np.array_equal(a.solution(), b.solution())
    assert a.state().steps == 12


def test_host_threads_share_the_library(hipdev):
    """Four host threads, each with its own stream and grid size, capture and
    replay step graphs at the same time (the graph caches and launch table are
    shared); every result equals the same run done alone.  While one thread
    captures, the others issue only work ordered on their own streams (the
    solvers' copies are stream-ordered): a device-wide synchronisation or a
    null-stream copy would invalidate the capture.  The solvers are created
    before the threads start and freed after they end, since freeing device
    memory may synchronise the device."""
    import threading

    sizes = (32, 48, 32, 64)
    u0s = [solvers.ic_fourier(n, seed=20 + i) for i, n in enumerate(sizes)]

    def run(s, u0, out, i):
        s.set_initial(u0, tend=10.0)
        for _ in range(5):
            s.steps(4, graph=True)
        out[i] = s.solution()

    alone = [None] * len(sizes)
    for i, (n, u0) in enumerate(zip(sizes, u0s)):
        run(solvers.PeriodicBurgers(n, scheme="Z2"), u0, alone, i)
    together = [None] * len(sizes)
    errors = []
    objs = [solvers.PeriodicBurgers(n, scheme="Z2") for n in sizes]
    hipdev.synchronize()

    def guarded(*args):
        try:
            run(*args)
        except Exception as e:  # reported below, with the case
            errors.append(repr(e))

    threads = [threading.Thread(target=guarded, args=(s, u0, together, i))
               for i, (s, u0) in enumerate(zip(objs, u0s))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for a, b in zip(alone, together):
        assert b is not None and np.array_equal(a, b)
    del objs


def test_no_op_after_tend(hipdev):
    n = 32
    s = solvers.PeriodicBurgers(n, scheme="JS")
    s.set_initial(solvers.ic_fourier(n, seed=3), tend=0.05)
    st = s.run(batch=4)
    u = s.solution()
    s.steps(5)
    st2 = s.state()
    assert st2.steps == st.steps and st2.t == st.t == 0.05
    assert np.array_equal(s.solution(), u)


def test_256_one_step_vs_oracle(hipdev):
    """Full-size grid of the WENO5-Z benchmark: one SSP-RK3 step, bitwise."""
    n = 256
    s = solvers.PeriodicBurgers(n, scheme="Z2")
    u0 = solvers.ic_fourier(n, seed=3)
    s.set_initial(u0, tend=10.0)
    s.steps(1, graph=False)
    g = s.solution()
    ref = u0.copy()
    O.burgers_step(ref, 2.0 / n, 0.4, 0.0, 10.0, **O.Z2_GLOBAL)
    assert rel_linf(g, ref) <= TOL
    assert np.array_equal(g, ref)


def test_256_conservation(hipdev):
    """Periodic conservative scheme: sum(u) is invariant up to rounding over many steps."""
    n = 256
    s = solvers.PeriodicBurgers(n, scheme="Z2")
    u0 = solvers.ic_fourier(n, seed=3)
    s.set_initial(u0, tend=10.0)
    s.steps(20)
    g = s.solution()
    assert np.isfinite(g).all()
    m0, m1 = float(np.sum(u0)), float(np.sum(g))
    assert abs(m1 - m0) <= 1e-11 * np.sum(np.abs(u0))
    assert np.max(np.abs(g)) <= np.max(np.abs(u0)) * (1 + 1e-6)


@pytest.mark.parametrize("scheme", ["Z2", "JS"])
def test_fp32_storage_error_vs_fp64(hipdev, scheme):
    """FP32 storage with FP64 weights: L1 / L-infinity distance to the FP64
    run (bit-identical to the oracle) at t = 0.5 on 64^3, reported and bounded."""
    n = 64
    u0 = solvers.ic_fourier(n, seed=3)
    a = solvers.PeriodicBurgers(n, scheme=scheme, precision="f64")
    b = solvers.PeriodicBurgers(n, scheme=scheme, precision="f32")
    a.set_initial(u0, tend=0.5)
    b.set_initial(u0, tend=0.5)
    sa, sb = a.run(batch=16), b.run(batch=16)
    assert sa.t == 0.5 and sb.t == 0.5
    ua, ub = a.solution(), b.solution().astype(np.float64)
    l1 = float(np.mean(np.abs(ub - ua)))
    linf = float(np.max(np.abs(ub - ua)))
    print(f"fp32 vs fp64 {scheme} 64^3 t=0.5: L1 {l1:.3e}  Linf {linf:.3e}  steps {sa.steps}/{sb.steps}")
    assert np.isfinite(ub).all()
    assert l1 < 1e-5 and linf < 1e-2
