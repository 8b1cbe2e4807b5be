"""The drop-ins of the shared-memory Burgers code, the blocked heat code and
the MPI heat code on the GPU against the oracle, bit for bit, on pitched
arrays whose padding columns must stay untouched, plus each driver's own
sequence of calls over a few steps."""
import numpy as np
import pytest

import oracle as O
from multigpu_advectiondiffusion_amd import _capi
from multigpu_advectiondiffusion_amd import reference_api as R
from multigpu_advectiondiffusion_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu


def field(shape, seed, lo=-1.0, hi=2.0):
    nx, ny, nz = shape
    return np.random.default_rng(seed).uniform(lo, hi, size=(nz, ny, nx))


def padded(a, pitch, seed):
    """a (nz, ny, nx) stored with row pitch `pitch`; padding columns get random values."""
    nz, ny, nx = a.shape
    out = np.random.default_rng(seed).uniform(-5, 5, size=(nz, ny, pitch))
    out[:, :, :nx] = a
    return out


@pytest.fixture(autouse=True)
def reset_scheme():
    R.set_adv_scheme()
    assert _capi.last_error() == 0, "an earlier call left an unread error"
    yield
    R.set_adv_scheme()


def check(dev, host, pad, nx):
    got = dev.download()
    assert np.array_equal(got[:, :, :nx], host), float(np.max(np.abs(got[:, :, :nx] - host)))
    assert np.array_equal(got[:, :, nx:], pad[:, :, nx:])


# ---------------------------------------------------------------------------
# SingleGPU/Burgers3d_WENO5_SharedMem
# ---------------------------------------------------------------------------
# 130 wide: several 63-node row segments, the last one partly outside; 5 x 6
# x 7: every node near a zero-padded end.  "p2" and "big_block" take the
# scheme setter's p = 2 and a block of nodes above 2^40 (IEEE weight
# divisions in some waves, the shortened ones in others).
@pytest.mark.parametrize("shape,pitch", [((40, 36, 24), 45), ((7, 5, 6), 8), ((130, 20, 9), 136), ((16, 70, 40), 16)])
@pytest.mark.parametrize("variant", ["reference", "p2", "big_block"])
def test_sharedmem_sweeps(hipdev, shape, pitch, variant):
    nx, ny, nz = shape
    kw = dict(O.B5S)
    if variant == "p2":
        kw.update(p=2, eps=1e-12)
        R.set_adv_scheme(_capi.WENO_Z, 2, 1e-12, _capi.LF_LOCAL, 0.0)
    u = field(shape, 1)
    if variant == "big_block":
        u[: nz // 2, : ny // 2, : nx // 2] *= 2.0 ** 42
    Lu = field(shape, 2)
    pu, pLu = padded(u, pitch, 3), padded(Lu, pitch, 4)
    du, dLu = DeviceArray.from_host(pu), DeviceArray.from_host(pLu)
    ref = Lu.copy()
    for axis, (call, d) in enumerate(((R.Call_Adv_x_sharedmem, 0.1), (R.Call_Adv_y_sharedmem, 0.07),
                                      (R.Call_Adv_z_sharedmem, 0.13))):
        call(None, None, None, pitch, nx, ny, nz, d, du, dLu)
        O.adv_b5s(axis, u, ref, d, **kw)
        check(dLu, ref, pLu, nx)


def test_sharedmem_driver_steps(hipdev):
    """SingleGPU/Burgers3d_WENO5_SharedMem/main.cpp:150-162 per step: uo <- u,
    then per stage Call_Adv_x/y/z, Call_Diff_ and Call_sspRK (the last two
    are the pitched base wrappers, as in that code)."""
    nx, ny, nz, pitch = 36, 28, 20, 40
    dx, dy, dz, dt = 2 / 35, 2 / 27, 2 / 19, 2e-3
    kx, ky, kz = 1e-3 / (12 * dx * dx), 1e-3 / (12 * dy * dy), 1e-3 / (12 * dz * dz)
    R.CopyToConstantMemory(kx, ky, kz)
    u = field((nx, ny, nz), 5, 0.0, 1.0)
    pu = padded(u, pitch, 6)
    du, duo, dLu = DeviceArray.from_host(pu), DeviceArray.from_host(pu), DeviceArray.from_host(np.zeros_like(pu))
    uo, Lu = u.copy(), np.zeros_like(u)
    for _ in range(3):
        duo.copy_from(du)
        uo[...] = u
        for step in (1, 2, 3):
            R.Call_Adv_x_sharedmem(None, None, None, pitch, nx, ny, nz, dx, du, dLu)
            R.Call_Adv_y_sharedmem(None, None, None, pitch, nx, ny, nz, dy, du, dLu)
            R.Call_Adv_z_sharedmem(None, None, None, pitch, nx, ny, nz, dz, du, dLu)
            R.Call_Diff_pitched(None, None, None, pitch, nx, ny, nz, du, dLu)
            R.Call_sspRK_pitched(None, None, None, pitch, nx, ny, nz, step, dt, du, duo, dLu)
            O.adv_b5s(0, u, Lu, dx)
            O.adv_b5s(1, u, Lu, dy)
            O.adv_b5s(2, u, Lu, dz)
            O.diff13_b5(u, Lu, kx, ky, kz)
            O.ssprk_b5(u, uo, Lu, step, dt)
    check(du, u, pu, nx)
    check(dLu, Lu, np.zeros_like(pu), nx)


# ---------------------------------------------------------------------------
# SingleGPU/Diffusion3d_Blocking
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape,pitch", [((24, 20, 18), 32), ((7, 7, 6), 8), ((70, 40, 37), 72), ((9, 33, 17), 9)])
def test_blocking_lu3d(hipdev, shape, pitch):
    nx, ny, nz = shape
    kx, ky, kz = 0.05, 0.04, 0.03
    R.CopyToConstantMemory(kx, ky, kz)
    u, Lu = field(shape, 11), field(shape, 12)
    pu, pLu = padded(u, pitch, 13), padded(Lu, pitch, 14)
    du, dLu = DeviceArray.from_host(pu), DeviceArray.from_host(pLu)
    R.Call_Lu3d_blocking(None, None, None, pitch, nx, ny, nz, du, dLu)
    ref = Lu.copy()
    O.lu3d_blk(u, ref, kx, ky, kz)
    check(dLu, ref, pLu, nx)


def test_blocking_driver_freezes_the_second_ring(hipdev):
    """Diffusion3d_Blocking/main.cpp:101-122: Lu zeroed once, then per step
    uo <- u and per stage Call_Lu3d + Call_RK3d.  Call_RK3d updates nodes
    2..n-3 but Call_Lu3d writes only x and y nodes 3..n-4 and planes
    3..nz-3, so on the rest of the updated region (x or y index 2 or n-3,
    plane 2) the stages see Lu = 0 and return their input up to the rounding
    of 0.75 u + 0.25 u and (u + 2u)/3."""
    nx, ny, nz, pitch = 30, 26, 22, 32
    h = 1.0 / 29
    kx = ky = kz = 1.0 / (12 * h * h)
    dt = 0.1 * h * h
    R.CopyToConstantMemory(kx, ky, kz)
    u = field((nx, ny, nz), 21, 0.0, 1.0)
    u0 = u.copy()
    pu = padded(u, pitch, 22)
    du, duo, dLu = DeviceArray.from_host(pu), DeviceArray.from_host(pu), DeviceArray.from_host(np.zeros_like(pu))
    uo, Lu = u.copy(), np.zeros_like(u)
    for _ in range(4):
        duo.copy_from(du)
        uo[...] = u
        for step in (1, 2, 3):
            R.Call_Lu3d_blocking(None, None, None, pitch, nx, ny, nz, du, dLu)
            R.Call_RK3d_pitched(None, None, None, step, pitch, nx, ny, nz, dt, du, duo, dLu)
            O.lu3d_blk(u, Lu, kx, ky, kz)
            O.rk3d(u, uo, Lu, step, dt)
    check(du, u, pu, nx)
    ring = np.zeros(u.shape, bool)
    ring[2:-2, 2:-2, 2:-2] = True    # Call_RK3d's region
    ring[3:-2, 3:-3, 3:-3] = False   # Call_Lu3d's (planes 3..nz-3)
    assert np.max(np.abs(u[ring] - u0[ring]) / np.abs(u0[ring])) <= 1e-14
    assert np.max(np.abs(u[3:-3, 3:-3, 3:-3] - u0[3:-3, 3:-3, 3:-3])) > 1e-3


# ---------------------------------------------------------------------------
# MultiGPU/Diffusion3d_Baseline
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("kstart,kstop", [(3, 6), (6, 20), (20, 23), (3, 23), (7, 7)])
def test_md3_diff(hipdev, order, kstart, kstop):
    nx, ny, nz, pitch = 40, 24, 26, 48
    dxc, dyc, dzc = 0.7, 0.4, 0.25
    u, Lu = field((nx, ny, nz), 31), field((nx, ny, nz), 32)
    pu, pLu = padded(u, pitch, 33), padded(Lu, pitch, 34)
    du, dLu = DeviceArray.from_host(pu), DeviceArray.from_host(pLu)
    call = R.Call_Diff_mpi if order == 4 else R.Call_Diff_O2_mpi
    call(None, None, None, pitch, nx, ny, nz, kstart, kstop, dxc, dyc, dzc, du, dLu)
    ref = pLu.copy()
    O.md3_lap(order, pu, ref, kstart, kstop, dxc, dyc, dzc, nx=nx)
    assert np.array_equal(dLu.download(), ref)


def test_md3_diff_rejects_planes_it_cannot_read(hipdev):
    du, dLu = DeviceArray((20, 12, 12)), DeviceArray((20, 12, 12))
    with pytest.raises(_capi.MgadError):
        R.Call_Diff_mpi(None, None, None, 12, 12, 12, 20, 1, 10, 1.0, 1.0, 1.0, du, dLu)  # O4 reads k-2
    with pytest.raises(_capi.MgadError):
        R.Call_Diff_O2_mpi(None, None, None, 12, 12, 12, 20, 3, 20, 1.0, 1.0, 1.0, du, dLu)  # reads k+1


def test_md3_rank_stage_sequence(hipdev):
    """One rank's stage of MultiGPU/Diffusion3d_Baseline/main.c:197-300
    without neighbours: Lu over the slab's planes RADIUS.._Nz+RADIUS-1 in
    three calls (two boundary regions and the interior), then the RK stage
    on every plane; three steps against the oracle."""
    nx, ny, nzs, pitch, R3 = 32, 20, 18, 36, 3
    NZ = nzs + 2 * R3
    dxc = 0.05
    u = field((nx, ny, NZ), 41, 0.0, 1.0)
    pu = padded(u, pitch, 42)
    du, duo, dLu = DeviceArray.from_host(pu), DeviceArray.from_host(pu), DeviceArray.from_host(np.zeros_like(pu))
    hu, huo, hLu = pu.copy(), pu.copy(), np.zeros_like(pu)
    for _ in range(3):
        duo.copy_from(du)
        huo[...] = hu
        for step in (1, 2, 3):
            for ks, ke in ((nzs, nzs + R3), (R3, 2 * R3), (2 * R3, nzs)):
                R.Call_Diff_mpi(None, None, None, pitch, nx, ny, NZ, ks, ke, dxc, dxc, dxc, du, dLu)
                O.md3_lap(4, hu, hLu, ks, ke, dxc, dxc, dxc, nx=nx)
            R.Call_sspRK_mpi(None, None, None, step, pitch, nx, ny, NZ, 1e-3, du, duo, dLu)
            O.md3_rk(hu, huo, hLu, step, 1e-3, nx=nx)
    assert np.array_equal(du.download(), hu)
    assert np.array_equal(dLu.download(), hLu)
