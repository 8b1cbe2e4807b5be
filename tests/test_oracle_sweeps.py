"""Pins the oracle's sweeps, RK stages and heat operators with invariants and
closed forms (conservation, permutation symmetry, eigenmode decay, RK3
amplification, polynomial exactness).  Reference: SingleGPU/Burgers3d_WENO5/
kernels.cu:197-530, SingleGPU/Diffusion3d_baselineCode/kernels.cu:37-142.
"""
import numpy as np
import pytest

import oracle as O


def rnd(shape, seed, lo=-1.0, hi=2.0):
    nx, ny, nz = shape
    return np.random.default_rng(seed).uniform(lo, hi, size=(nz, ny, nx))


# ---------------------------------------------------------------- drop-in sweeps
def test_adv_constant_field():
    """Constant u: every face flux equals f(c) so Lu = 0 on the written set,
    except the x sweep's first node i=2 where F_{3/2} := 0 (kernels.cu:222)."""
    n, c = 12, 0.8
    u = np.full((n, n, n), c)
    Lu = np.full_like(u, 7.0)
    O.adv(0, u, Lu, 0.1, **O.REFERENCE)
    f = O.face(np.full(5, O.split(c)[0]), np.full(5, O.split(c)[1]), **O.REFERENCE)
    inner = Lu[3:n - 3, 3:n - 3, 3:n - 3]
    assert np.all(inner == 0.0)
    assert np.all(Lu[3:n - 3, 3:n - 3, 2] == -(f - 0) / 0.1)
    assert np.all(Lu[:, :, 0:2] == 7.0) and np.all(Lu[:, :, n - 3:] == 7.0)
    assert np.all(Lu[0:3] == 7.0) and np.all(Lu[:, 0:3] == 7.0)


@pytest.mark.parametrize("kw", [O.REFERENCE, O.JS_GLOBAL | {"alpha": 2.0}])
def test_adv_permutation(kw):
    """Transposing the data transposes the sweeps: Call_Adv_y on u^T equals
    -(Call_Adv_x on u) transposed, on a cubic grid."""
    n = 14
    u = rnd((n, n, n), 4)
    Lx = np.zeros_like(u)
    O.adv(0, u, Lx, 0.05, **kw)
    uy = np.ascontiguousarray(np.swapaxes(u, 1, 2))  # x <-> y
    Ly = np.zeros_like(u)
    O.adv(1, uy, Ly, 0.05, **kw)
    assert np.array_equal(np.swapaxes(Ly, 1, 2), Lx)
    uz = np.ascontiguousarray(np.swapaxes(u, 0, 2))  # x <-> z
    Lz = np.zeros_like(u)
    O.adv(2, uz, Lz, 0.05, **kw)
    assert np.array_equal(np.swapaxes(Lz, 0, 2), Lx)


def test_adv_telescoping():
    """dx * sum_i Lu over one row telescopes to -(F_{nx-4+1/2}) (F_{3/2} = 0)."""
    n = 16
    u = rnd((n, n, n), 5)
    Lu = np.zeros_like(u)
    dx = 0.1
    O.adv(0, u, Lu, dx, **O.REFERENCE)
    k, j = 5, 7
    row = u[k, j]
    fp = np.array([O.split(x)[0] for x in row])
    fm = np.array([O.split(x)[1] for x in row])
    i = n - 4
    F_last = O.face(fp[i - 2:i + 3], fm[i - 1:i + 4], **O.REFERENCE)
    s = np.sum(Lu[k, j, 2:n - 3]) * dx
    assert s == pytest.approx(-F_last, rel=1e-12, abs=1e-12)


def test_diff13_polynomial_exact():
    """The 13-point operator is exact for polynomials of degree <= 5 per axis:
    Lu = 12 dx^2 kx u_xx + ... = C lap(u) with k = C/(12 dx^2)."""
    nx, ny, nz = 16, 14, 12
    dx, dy, dz = 0.1, 0.13, 0.07
    C = 1.0
    kx, ky, kz = C / (12 * dx * dx), C / (12 * dy * dy), C / (12 * dz * dz)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = i * dx, j * dy, k * dz
    u = x ** 5 - 2 * x ** 3 + y ** 4 + 3 * y ** 2 * 1.0 + z ** 5 + x * y * z
    lap = 20 * x ** 3 - 12 * x + 12 * y ** 2 + 6 + 20 * z ** 3
    L1 = np.zeros_like(u)
    O.lu3d_13(u, L1, kx, ky, kz)
    sl = (slice(2, nz - 2), slice(2, ny - 2), slice(2, nx - 2))
    assert np.allclose(L1[sl], C * lap[sl], rtol=1e-9, atol=1e-9)
    L2 = np.zeros_like(u)
    O.diff13_b5(u, L2, kx, ky, kz)
    sl3 = (slice(3, nz - 2), slice(3, ny - 3), slice(3, nx - 3))
    assert np.allclose(L2[sl3], L1[sl3], rtol=1e-12, atol=1e-11)
    # untouched outside the written set
    assert np.all(L2[:3] == 0) and np.all(L2[nz - 2:] == 0)


def test_lu3d_o2_quadratic_exact():
    nx, ny, nz = 10, 9, 8
    dx, dy, dz = 0.1, 0.2, 0.3
    C = 2.0
    kx, ky, kz = C / (12 * dx * dx), C / (12 * dy * dy), C / (12 * dz * dz)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = i * dx, j * dy, k * dz
    u = 3 * x * x - y * y + 0.5 * z * z + x * y
    Lu = np.zeros_like(u)
    O.lu3d_o2(u, Lu, kx, ky, kz)
    assert np.allclose(Lu[1:-1, 1:-1, 1:-1], C * (6 - 2 + 1.0), rtol=1e-11)
    assert np.all(Lu[0] == 0) and np.all(Lu[:, :, -1] == 0)


# ---------------------------------------------------------------- RK
@pytest.mark.parametrize("z", [-0.3, -1.0, 0.2])
def test_rk3_amplification(z):
    """u' = lambda u: three SSP-RK3 stages give (1 + z + z^2/2 + z^3/6) u0."""
    lam, dt = z / 0.01, 0.01
    n = 9
    u0 = rnd((n, n, n), 6, 0.5, 1.5)
    for fn in ("all", "b5", "d3"):
        u, uo = u0.copy(), u0.copy()
        for stage in (1, 2, 3):
            Lu = lam * u
            if fn == "all":
                O.rk_stage_all(u, uo, Lu, stage, dt)
            elif fn == "b5":
                O.ssprk_b5(u, uo, Lu, stage, dt)
            else:
                O.rk3d(u, uo, Lu, stage, dt)
        amp = 1 + z + z * z / 2 + z ** 3 / 6
        lo = {"all": 0, "b5": 3, "d3": 2}[fn]
        sl = (slice(lo, n - lo),) * 3
        assert np.allclose(u[sl], amp * u0[sl], rtol=1e-14)
        if lo:
            assert np.array_equal(u[:lo], u0[:lo])


# ---------------------------------------------------------------- periodic RHS
def test_periodic_conservation_and_1d_reduction():
    n = (24, 16, 20)
    u = rnd(n, 7)
    for kw in (O.Z2_GLOBAL, O.JS_GLOBAL):
        a = O.maxabs(u)
        Lu = O.rhs_periodic(u, (0.1, 0.2, 0.3), **kw, alpha=a)
        assert abs(Lu.sum()) <= 1e-11 * np.abs(Lu).sum()
    # data varying only in x: y and z differences vanish exactly
    nx = 32
    x1 = np.random.default_rng(8).uniform(0, 1, nx)
    u = np.broadcast_to(x1, (6, 5, nx)).copy()
    Lu = O.rhs_periodic(u, 0.1, **O.Z2_GLOBAL, alpha=1.0)
    L1 = O.rhs_1d_periodic(x1, 0.1, **O.Z2_GLOBAL, alpha=1.0)
    assert np.array_equal(Lu[3, 2], L1)


def test_periodic_permutation():
    n = 12
    u = rnd((n, n, n), 9)
    La = O.rhs_periodic(u, 0.1, **O.JS_GLOBAL, alpha=2.0)
    ut = np.ascontiguousarray(np.transpose(u, (1, 2, 0)))
    Lb = O.rhs_periodic(ut, 0.1, **O.JS_GLOBAL, alpha=2.0)
    assert np.allclose(np.transpose(La, (1, 2, 0)), Lb, rtol=1e-13, atol=1e-12)


def test_burgers_run_conserves_mean():
    from multigpu_advectiondiffusion_amd import solvers

    u = solvers.ic_fourier(16, seed=2)
    m0 = u.sum()
    steps, t = O.burgers_run(u, 2.0 / 16, 0.4, 0.0, 0.5, **O.Z2_GLOBAL)
    assert t == 0.5 and steps > 5
    assert abs(u.sum() - m0) <= 1e-12 * np.abs(u).sum()


# ---------------------------------------------------------------- heat
def test_heat7_eigenmode_decay():
    """A Dirichlet sine mode is an eigenvector of the 7-point operator:
    after n forward-Euler steps it is scaled by exactly (1 - dt*lambda)^n."""
    N = 20
    dx = 1.0 / (N - 1)
    C = 1.0
    kx = ky = kz = C / (12 * dx * dx)
    dt = 0.15 * dx * dx
    m = (1, 2, 3)
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    u = np.sin(np.pi * m[0] * i * dx) * np.sin(np.pi * m[1] * j * dx) * np.sin(np.pi * m[2] * k * dx)
    u[0], u[-1], u[:, 0], u[:, -1], u[:, :, 0], u[:, :, -1] = 0, 0, 0, 0, 0, 0
    lam = sum(12 * kx * 4 * np.sin(np.pi * md * dx / 2) ** 2 for md in m)
    v = u.copy()
    O.heat7_fe(v, kx, ky, kz, dt, 100)
    assert np.allclose(v, (1 - dt * lam) ** 100 * u, rtol=1e-11, atol=1e-13)


def test_heat7_max_principle():
    from multigpu_advectiondiffusion_amd import solvers

    u = solvers.ic_gauss_noise(24, seed=3)
    m0 = np.abs(u).max()
    dx = 1.0 / 23
    k = 1.0 / (12 * dx * dx)
    O.heat7_fe(u, k, k, k, 0.15 * dx * dx, 50)
    assert np.abs(u).max() <= m0


# ------------------------------------- MultiGPU/Burgers3d_Baseline slab sweeps
def test_mb3_xy_equal_single_grid_sweeps():
    """Compute_dF/dG over planes 3..nz-4 at pitch nx are the single-GPU sweeps
    with the same scheme (MB3/Kernels.cu:225-371 vs B5/kernels.cu:197-327)."""
    nx, ny, nz = 13, 11, 12
    u = rnd((nx, ny, nz), 21)
    for axis in (0, 1):
        L0 = rnd((nx, ny, nz), 22)
        L1 = L0.copy()
        O.adv(axis, u, L0, 0.07, **O.MB3)
        O.mb3_adv(axis, u, L1, 3, nz - 3, 0.07)
        assert np.array_equal(L0, L1)


def test_mb3_pitch_padding_never_read():
    """Rows padded to pitch nx+5 with NaN: the same values on the grid, the
    padding of Lu untouched, for all three sweeps and the RK stage."""
    nx, ny, nz, px = 12, 10, 14, 17
    u = rnd((nx, ny, nz), 23)
    up = np.full((nz, ny, px), np.nan)
    up[:, :, :nx] = u
    for axis in (0, 1, 2):
        L = rnd((nx, ny, nz), 24)
        Lp = np.full((nz, ny, px), -5.0)
        Lp[:, :, :nx] = L
        O.mb3_adv(axis, u, L, 3, nz - 3, 0.1)
        O.mb3_adv(axis, up, Lp, 3, nz - 3, 0.1, nx=nx)
        assert np.array_equal(Lp[:, :, :nx], L) and np.all(Lp[:, :, nx:] == -5.0)
    qo, Lq = rnd((nx, ny, nz), 25), rnd((nx, ny, nz), 26)
    q = u.copy()
    O.mb3_rk(q, qo, Lq, 2, 0.01)
    qp = np.full((nz, ny, px), 3.0)
    qp[:, :, :nx] = u
    qop, Lqp = qp.copy(), qp.copy()
    qop[:, :, :nx], Lqp[:, :, :nx] = qo, Lq
    O.mb3_rk(qp, qop, Lqp, 2, 0.01, nx=nx)
    assert np.array_equal(qp[:, :, :nx], q) and np.all(qp[:, :, nx:] == 3.0)


def test_mb3_z_equals_single_grid_sweep_and_splits():
    """Compute_dH over planes 3..nz-4 gives the single-GPU z sweep there
    (both faces of each plane; only the single-GPU sweep also writes plane 2
    from F = 0), and two adjacent plane ranges give the whole range."""
    nx, ny, nz = 11, 12, 16
    u = rnd((nx, ny, nz), 27)
    L0 = rnd((nx, ny, nz), 28)
    L1, L2 = L0.copy(), L0.copy()
    Lb = L0.copy()
    O.adv(2, u, Lb, 0.09, **O.MB3)
    O.mb3_adv(2, u, L1, 3, nz - 3, 0.09)
    assert np.array_equal(L1[3:nz - 3], Lb[3:nz - 3])
    assert np.array_equal(L1[:3], L0[:3]) and np.array_equal(L1[nz - 3:], L0[nz - 3:])
    O.mb3_adv(2, u, L2, 3, 7, 0.09)
    O.mb3_adv(2, u, L2, 7, nz - 3, 0.09)
    assert np.array_equal(L1, L2)


def test_mb3_rk_every_node():
    """Compute_RK of the MB3 slab (Kernels.cu:457-490) is the single-GPU stage
    on the interior and advances the ghost nodes too."""
    n = 10
    q0, qo, Lq = rnd((n, n, n), 29), rnd((n, n, n), 30), rnd((n, n, n), 31)
    for step in (1, 2, 3):
        a, b = q0.copy(), q0.copy()
        O.ssprk_b5(a, qo, Lq, step, 0.02)
        O.mb3_rk(b, qo, Lq, step, 0.02)
        assert np.array_equal(a[3:n - 3, 3:n - 3, 3:n - 3], b[3:n - 3, 3:n - 3, 3:n - 3])
        assert np.all(b[0] != q0[0])
