"""Pins the oracle's 13-point heat operator and RK3 stage (orc_lu3d_13,
orc_rk3d: SingleGPU/Diffusion3d_baselineCode/kernels.cu Compute_Laplace3d and
Compute_RK) to the reference's only known answers, the 3-D accuracy table of
Matlab_Prototipes/DiffusionNd/TestingAccuracy.log (see matlab_heat_log.py for
the run that reproduces it).  A dropped or mis-signed stencil term, a wrong
RK weight or a wrong coefficient moves the norms by far more than the last
printed digit."""
import numpy as np
import pytest

import oracle as O
from matlab_heat_log import LOG_3D, case, norms, printed


@pytest.mark.parametrize("n", [10, 20, 40, 80])
def test_heat13_rk3_reproduces_the_matlab_accuracy_log(n):
    u0, ue, dx, kx, dt, steps = case(n)
    u, Lu = u0.copy(), np.zeros_like(u0)
    for _ in range(steps):
        uo = u.copy()
        for stage in (1, 2, 3):
            O.lu3d_13(u, Lu, kx, kx, kx)
            O.rk3d(u, uo, Lu, stage, dt)
    got = tuple(printed(v) for v in norms(u, ue, dx))
    assert got == LOG_3D[n], (n, steps, got)


def test_the_log_is_sensitive_to_the_operator():
    """The same run with the 7-point second difference in place of the
    13-point one, or a 5 % larger coefficient, misses the log's digits."""
    n = 20
    u0, ue, dx, kx, dt, steps = case(n)
    for variant in ("o2", "coef"):
        u, Lu = u0.copy(), np.zeros_like(u0)
        for _ in range(steps):
            uo = u.copy()
            for stage in (1, 2, 3):
                if variant == "o2":
                    O.lu3d_o2(u, Lu, kx, kx, kx)  # 12 kx (u_e - 2u + u_w) per axis
                else:
                    O.lu3d_13(u, Lu, 1.05 * kx, 1.05 * kx, 1.05 * kx)
                O.rk3d(u, uo, Lu, stage, dt)
        got = tuple(printed(v) for v in norms(u, ue, dx))
        assert got != LOG_3D[n], variant


# Matlab_Prototipes/DiffusionNd/TestingAccuracy.log:50-64 (L1, Linf) with the
# L2 of the run logs (:1-48): n -> (L1, L2, Linf)
LOG_1D = {10: (1.81e-01, 7.90e-02, 4.97e-02), 20: (4.87e-02, 1.96e-02, 1.34e-02),
          40: (9.89e-03, 3.77e-03, 2.62e-03), 80: (1.50e-03, 7.43e-04, 8.34e-04)}
LOG_2D = {10: (2.97e-01, 6.79e-02, 2.65e-02), 20: (3.42e-02, 5.97e-03, 2.96e-03),
          40: (5.14e-03, 8.55e-04, 3.59e-04), 80: (2.99e-03, 5.39e-04, 2.64e-04)}


def _matlab_nd(nd, n, tfinal=1.0, factor=0.9):
    """diffusion1dTest.m / diffusion2dTest.m (the 3-D script in 1 or 2
    dimensions: 13-point second difference on nodes 3..n-2 (1-based), SSP-RK3,
    boundary nodes set to 0 after each step, dt clipped only for the time
    counter), in numpy; returns the printed (L1, L2, Linf)."""
    L, t0 = 10.0, 0.1
    dx = L / (n - 1)
    x = -L / 2 + dx * np.arange(n)
    r2 = sum(g ** 2 for g in np.meshgrid(*([x] * nd), indexing="ij"))
    u = np.exp(-r2 / (4 * t0))
    ue = (t0 / tfinal) ** (nd / 2) * np.exp(-r2 / (4 * tfinal))
    Dx = 1.0 / dx ** 2
    inner = (slice(2, -2),) * nd

    def lap(v):
        out = np.zeros_like(v)
        acc = 0.0
        for ax in range(nd):
            def sh(o, ax=ax):
                sl = [slice(2, -2)] * nd
                sl[ax] = slice(2 + o, n - 2 + o)
                return v[tuple(sl)]
            acc = acc + Dx / 12 * (-sh(2) + 16 * sh(1) - 30 * sh(0) + 16 * sh(-1) - sh(-2))
        out[inner] = acc
        return out

    dt = 1 / (2 * (nd / dx ** 2)) * factor
    t = t0
    while t < tfinal:
        uo = u.copy()
        u = uo + dt * lap(u)
        u = 0.75 * uo + 0.25 * (u + dt * lap(u))
        u = (uo + 2 * (u + dt * lap(u))) / 3
        for ax in range(nd):
            for end in (0, -1):
                sl = [slice(None)] * nd
                sl[ax] = end
                u[tuple(sl)] = 0
        if t + dt > tfinal:
            dt = tfinal - t
        t = t + dt
    err = np.abs(ue - u)
    vol = dx ** nd
    return tuple(printed(v) for v in (vol * err.sum(), np.sqrt(vol * (err ** 2).sum()), err.max()))


@pytest.mark.parametrize("nd,table", [(1, LOG_1D), (2, LOG_2D)])
def test_the_reading_of_the_log_reproduces_its_1d_and_2d_tables(nd, table):
    """The run parameters the 3-D pin relies on (points 10..80, tFinal 1.0,
    factor 0.9) also give the 1-D and 2-D tables of the same log."""
    for n, want in table.items():
        assert _matlab_nd(nd, n) == want, (nd, n)
