"""The fused reference-driver step (mgad_burgers_ref_steps) against the CPU
oracle's restatement of the drop-in sequence (Call_Adv_x/y/z, Call_Diff_,
Call_sspRK; SingleGPU/Burgers3d_WENO5/main.cpp:139-149), bit for bit."""
import numpy as np
import pytest

import oracle as O
from multigpu_advectiondiffusion_amd import fileio, solvers
from multigpu_advectiondiffusion_amd import reference_api as R

pytestmark = pytest.mark.gpu

ORACLE_KW = {"Z1": O.REFERENCE, "Z2": dict(weno=O.WENO_Z, p=2, eps=1e-40, lf=O.LF_LOCAL),
             "JS": dict(weno=O.WENO_JS, p=2, eps=1e-6, lf=O.LF_LOCAL)}


def oracle_steps(u, nsteps, dx, dy, dz, dt, kw, viscous, kc):
    u = u.copy()
    Lu = np.zeros_like(u)
    for _ in range(nsteps):
        uo = u.copy()
        for step in (1, 2, 3):
            O.adv(0, u, Lu, dx, **kw)
            O.adv(1, u, Lu, dy, **kw)
            O.adv(2, u, Lu, dz, **kw)
            if viscous:
                O.diff13_b5(u, Lu, *kc)
            O.ssprk_b5(u, uo, Lu, step, dt)
    return u


@pytest.mark.parametrize("shape", [(32, 24, 20), (7, 7, 7), (45, 33, 17), (70, 9, 12)])
@pytest.mark.parametrize("scheme,viscous", [("Z1", True), ("Z1", False), ("JS", True), ("Z2", False)])
@pytest.mark.parametrize("graph", [False, True])
def test_fused_reference_step_bitwise(hipdev, shape, scheme, viscous, graph):
    nx, ny, nz = shape
    L = 2.0
    dx, dy, dz = L / (nx - 1), L / (ny - 1), L / (nz - 1)
    C = 1e-5
    kc = (C / (12 * dx * dx), C / (12 * dy * dy), C / (12 * dz * dz))
    R.CopyToConstantMemory(*kc)
    u0 = fileio.init3d_burgers(3, shape, dx, dy, dz)
    u0 += 0.05 * np.random.default_rng(7).standard_normal(u0.shape)  # rough data: every WENO branch
    dt = 0.4 * min(dx, dy, dz) / 1.2
    s = solvers.ReferenceBurgers(shape, dx, dy, dz, scheme=scheme, viscous=viscous)
    s.set_initial(u0)
    s.steps(2, dt, graph=graph)
    s.steps(1, dt, graph=graph)
    ref = oracle_steps(u0, 3, dx, dy, dz, dt, ORACLE_KW[scheme], viscous, kc)
    got = s.solution()
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]


@pytest.mark.parametrize("scheme,viscous", [("Z1", True), ("JS", False)])
def test_fused_reference_step_square_tiles_bitwise(hipdev, scheme, viscous):
    """150 x 140 x 70: the 144 x 134 x 64 interior's 5 x 5 tiles of 32 x 32
    (ragged in x and y) fill the CUs in one round of 10 bands, so the step
    runs 32 x 32 workgroups, one per CU."""
    shape = (150, 140, 70)
    nx, ny, nz = shape
    dx, dy, dz = 2.0 / (nx - 1), 2.0 / (ny - 1), 2.0 / (nz - 1)
    kc = tuple(1e-5 / (12 * d * d) for d in (dx, dy, dz))
    R.CopyToConstantMemory(*kc)
    u0 = fileio.init3d_burgers(3, shape, dx, dy, dz)
    u0 += 0.05 * np.random.default_rng(8).standard_normal(u0.shape)
    dt = 0.4 * min(dx, dy, dz) / 1.2
    s = solvers.ReferenceBurgers(shape, dx, dy, dz, scheme=scheme, viscous=viscous)
    s.set_initial(u0)
    s.steps(2, dt, graph=True)
    ref = oracle_steps(u0, 2, dx, dy, dz, dt, ORACLE_KW[scheme], viscous, kc)
    got = s.solution()
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]


@pytest.mark.parametrize("graph", [False, True])
def test_fused_reference_step_ieee_division_path(hipdev, graph):
    """A block of nodes above 2^40 in magnitude: the stages take the IEEE
    weight divisions (the fused kernel's fast-division test is on the stage
    input's max|u|), still bit for bit with the oracle.  The block lies in
    the interior, away from the frozen ghost layers, so every stage updates it."""
    nx, ny, nz = 40, 32, 24
    dx = dy = dz = 2.0 / 39
    kc = (1e-5 / (12 * dx * dx),) * 3
    R.CopyToConstantMemory(*kc)
    u0 = fileio.init3d_burgers(3, (nx, ny, nz), dx, dy, dz)
    blk = u0[8:12, 10:14, 12:20]
    u0[8:12, 10:14, 12:20] = 2.0 ** 41 * (1.0 + 0.1 * np.random.default_rng(3).standard_normal(blk.shape))
    dt = 1e-16
    s = solvers.ReferenceBurgers((nx, ny, nz), dx, dy, dz, scheme="Z1", viscous=True)
    s.set_initial(u0)
    s.steps(2, dt, graph=graph)
    ref = oracle_steps(u0, 2, dx, dy, dz, dt, ORACLE_KW["Z1"], True, kc)
    got = s.solution()
    assert np.isfinite(ref).all()
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]


def test_fused_reference_step_matches_dropins_on_gpu(hipdev):
    """Same u as the GPU drop-in sequence itself (reference_api.burgers_reference_step)."""
    from multigpu_advectiondiffusion_amd.hip import DeviceArray
    nx, ny, nz = 48, 40, 36
    dx = dy = dz = 2.0 / 47
    kc = (1e-5 / (12 * dx * dx),) * 3
    R.CopyToConstantMemory(*kc)
    R.set_adv_scheme()
    u0 = fileio.init3d_burgers(3, (nx, ny, nz), dx, dy, dz)
    dt = 0.4 * dx
    du, duo, dLu = DeviceArray.from_host(u0), DeviceArray(u0.shape), DeviceArray(u0.shape)
    dLu.zero()
    for _ in range(4):
        R.burgers_reference_step(du, duo, dLu, nx, ny, nz, dx, dy, dz, dt)
    s = solvers.ReferenceBurgers((nx, ny, nz), dx, dy, dz)
    s.set_initial(u0)
    s.steps(4, dt)
    assert np.array_equal(s.solution(), du.download())


def test_fused_reference_step_rejects_small_grid(hipdev):
    with pytest.raises(Exception):
        solvers.ReferenceBurgers((6, 8, 8), 0.1, 0.1, 0.1).steps(1, 0.01)


# (67, 9, 35): z chunks of 8 planes and a remainder of 7; (1024, 516, 45): one
# chunk of 41 planes, five passes of the kernel's unrolled plane loop and one
# more plane
@pytest.mark.parametrize("shape", [(24, 20, 18), (5, 6, 7), (67, 9, 35), (1024, 516, 45)])
@pytest.mark.parametrize("graph", [False, True])
def test_fused_reference_heat_bitwise(hipdev, shape, graph):
    """mgad_heat13_rk3_steps against the oracle's Call_Lu3d + Call_RK3d sequence."""
    nx, ny, nz = shape
    L = 10.0
    dx, dy, dz = L / (nx - 1), L / (ny - 1), L / (nz - 1)
    kc = (1.0 / (12 * dx * dx), 1.0 / (12 * dy * dy), 1.0 / (12 * dz * dz))
    R.CopyToConstantMemory(*kc)
    dt = 1.0 / (2 * (1 / dx / dx + 1 / dy / dy + 1 / dz / dz)) * 0.9
    u0 = fileio.init3d_heat(3, shape, dx, dy, dz) + 1e-3 * np.random.default_rng(3).standard_normal((nz, ny, nx))
    h = solvers.ReferenceHeat(shape)
    h.set_initial(u0)
    h.steps(3, dt, graph=graph)
    h.steps(2, dt, graph=graph)
    u, Lu = u0.copy(), np.zeros_like(u0)
    for _ in range(5):
        uo = u.copy()
        for step in (1, 2, 3):
            O.lu3d_13(u, Lu, *kc)
            O.rk3d(u, uo, Lu, step, dt)
    got = h.solution()
    assert np.array_equal(got, u), np.argwhere(got != u)[:5]
