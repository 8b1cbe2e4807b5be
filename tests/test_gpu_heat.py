"""7-point forward-Euler heat on the GPU vs the CPU oracle (bitwise)."""
import numpy as np
import pytest

import oracle as O
from multigpu_advectiondiffusion_amd import solvers

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graph", [False, True])
def test_heat_64_100_steps(hipdev, graph):
    """64^3, r = 0.15, 100 steps, Gaussian sigma 0.1 + 1e-3 noise (seed 1)."""
    h = solvers.Heat7(64)
    u0 = solvers.ic_gauss_noise(64, sigma=0.1, noise=1e-3, seed=1)
    h.set_initial(u0)
    h.steps(100, graph=graph)
    g = h.solution()
    ref = u0.copy()
    O.heat7_fe(ref, h.kx, h.ky, h.kz, h.dt, 100)
    assert np.array_equal(g, ref)


def test_heat_graph_replays(hipdev):
    h = solvers.Heat7((40, 30, 20))
    u0 = solvers.ic_gauss_noise((40, 30, 20), seed=5)
    h.set_initial(u0)
    ref = u0.copy()
    for _ in range(3):
        h.steps(7, graph=True)
        O.heat7_fe(ref, h.kx, h.ky, h.kz, h.dt, 7)
        assert np.array_equal(h.solution(), ref)


# Partial edge tiles in x and y, grids thinner than a tile, and z chunks of 1,
# 2, 20 and 83 planes (the last two more than one pass of the kernel's
# unrolled plane loop, the last with a remainder).
@pytest.mark.parametrize("shape", [(3, 3, 3), (67, 21, 35), (130, 17, 9), (256, 256, 27), (512, 512, 40),
                                   (1024, 512, 83)])
def test_heat_shapes_bitwise(hipdev, shape):
    h = solvers.Heat7(shape)
    u0 = solvers.ic_gauss_noise(shape, sigma=0.2, noise=1e-3, seed=7)
    h.set_initial(u0)
    h.steps(3, graph=False)
    ref = u0.copy()
    O.heat7_fe(ref, h.kx, h.ky, h.kz, h.dt, 3)
    got = h.solution()
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]
