"""Size-independent properties at the benchmark sizes, where the oracle would
take minutes: symmetry under exchanging x and y (bit for bit, since every
face is the same function of its stencil and the x and y contributions enter
as a sum), the decay factor of a discrete heat eigenmode, and the same
x/y symmetry for the WENO7 scheme."""
import numpy as np

from multigpu_advectiondiffusion_amd import solvers
from multigpu_advectiondiffusion_amd.hip import DeviceArray

import pytest

pytestmark = pytest.mark.gpu


def swap_xy(u):
    return np.ascontiguousarray(np.swapaxes(u, 1, 2))


def test_256_burgers_xy_exchange_bitwise(hipdev):
    """Lu = ((-dF) - dG) - dH: exchanging x and y exchanges dF and dG, and
    (-a) - b = (-b) - a in IEEE arithmetic, so the run on the transposed field
    is the transposed run, bit for bit, over 10 steps of the 256^3 benchmark."""
    n = 256
    u0 = solvers.ic_fourier(n, seed=3)
    a = solvers.PeriodicBurgers(n, scheme="Z2")
    a.set_initial(u0, tend=10.0)
    a.steps(10)
    b = solvers.PeriodicBurgers(n, scheme="Z2")
    b.set_initial(swap_xy(u0), tend=10.0)
    b.steps(10)
    assert np.array_equal(swap_xy(a.solution()), b.solution())


def test_512_heat_eigenmode_decay(hipdev):
    """u = sin(pi x) sin(2 pi y) sin(3 pi z) on the 512^3 Dirichlet grid is a
    discrete eigenmode of the 7-point Laplacian: one forward-Euler step scales
    it by 1 - dt*lambda with lambda = sum 4/h^2 sin^2(m pi h / 2)."""
    n = 512
    h = solvers.Heat7(n)
    x = np.arange(n) * h.dx
    m = (1, 2, 3)
    u0 = (np.sin(m[0] * np.pi * x)[None, None, :] * np.sin(m[1] * np.pi * x)[None, :, None] *
          np.sin(m[2] * np.pi * x)[:, None, None])
    u0[:, :, [0, -1]] = 0.0
    u0[:, [0, -1], :] = 0.0
    u0[[0, -1], :, :] = 0.0
    h.set_initial(u0)
    h.steps(1, graph=False)
    got = h.solution()
    lam = sum(4.0 / h.dx ** 2 * np.sin(mm * np.pi * h.dx / 2) ** 2 for mm in m)
    C = 12 * h.kx * h.dx ** 2  # the diffusivity behind kx = C/(12 dx^2)
    factor = 1.0 - h.dt * C * lam
    assert np.max(np.abs(got - factor * u0)) <= 1e-12 * np.max(np.abs(u0))


def test_128_weno7_xy_exchange_bitwise(hipdev):
    """L = (dF + dG) + dH is symmetric in x and y: the transposed field gives
    the transposed result over 5 low-storage RK steps."""
    n = 128
    a = solvers.Weno7Burgers(n)
    u0 = a.gaussian() + 0.1 * np.sin(3 * np.arange(n))[None, None, :]
    a.set_initial(u0, tend=10.0)
    a.steps(5)
    b = solvers.Weno7Burgers(n)
    b.set_initial(swap_xy(u0), tend=10.0)
    b.steps(5)
    assert np.array_equal(swap_xy(a.solution()), b.solution())
