"""The reference's only known answers for the heat path: the 3-D table of
Matlab_Prototipes/DiffusionNd/TestingAccuracy.log:66-72 (L1 and L-infinity
error of the 13-point SSP-RK3 heat solver against the exact Gaussian), with
the third norm (L2) of the run logs above it (:9-11, :21-23, :33-35, :45-47).

The log reproduces, every norm to its 3 printed digits, from the committed
diffusion3dTest.m (Matlab_Prototipes/DiffusionNd/diffusion3dTest.m:11-76:
D = 1, t0 = 0.1, L = W = H = 10, u0 = exp(-r^2/(4 t0)), exact solution
(t0/t)^1.5 exp(-r^2/(4 t)), dt = factor / (2 D sum 1/dx^2), SSP-RK3 on the
13-point Laplacian, boundary faces set to 0 after each step) called with
nx = ny = nz = 10, 20, 40, 80 points, tFinal = 1.0 and factor 0.9 - an
earlier TestingAccuracy.m than the committed one (which passes
nc = 9, 17, 33, 65, tEnd = 0.5 and prints nc - 1).  The time loop takes
every step with the full dt: it clips dt only after the step, for the time
counter (diffusion3dTest.m:66-70), so the run is `steps` full steps.

The reference's C heat step (Call_Lu3d + Call_RK3d) freezes the two outer
node layers instead of zeroing the outermost after each step; the
Gaussian is below 1e-16 there for every grid here, so the printed digits are
the same.
"""
import numpy as np

# n -> (L1, L2, Linf) as printed
LOG_3D = {10: (3.99e-01, 3.45e-02, 7.31e-03), 20: (2.75e-02, 1.84e-03, 5.04e-04),
          40: (1.35e-02, 9.80e-04, 3.29e-04), 80: (4.70e-03, 3.38e-04, 8.33e-05)}

D, T0, L, TFINAL, FACTOR = 1.0, 0.1, 10.0, 1.0, 0.9


def case(n):
    """(u0, exact solution, dx, kx, dt, steps) of the n^3 run, arrays (z, y, x)."""
    dx = L / (n - 1)
    x = -L / 2 + dx * np.arange(n)
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    r2 = x3 ** 2 + y3 ** 2 + z3 ** 2
    u0 = np.exp(-r2 / (4 * 1.0 * T0))
    ue = (T0 / TFINAL) ** 1.5 * np.exp(-r2 / (4 * 1.0 * TFINAL))
    Dx = D / dx ** 2
    kx = Dx / 12
    dt = 1 / (2 * D * (3 / dx ** 2)) * FACTOR
    # diffusion3dTest.m:44-71: steps of dt until the (clipped) time counter reaches tFinal
    t, h, steps = T0, dt, 0
    while t < TFINAL:
        steps += 1
        if t + h > TFINAL:
            h = TFINAL - t
        t = t + h
    return u0, ue, dx, kx, dt, steps


def norms(u, ue, dx):
    """diffusion3dTest.m:73-76."""
    err = np.abs(ue - u)
    return dx ** 3 * np.sum(err), float(np.sqrt(dx ** 3 * np.sum(err ** 2))), float(np.max(err))


def printed(v):
    """The value as the log prints it (%1.2e)."""
    return float(f"{v:1.2e}")
