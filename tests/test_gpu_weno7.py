"""The WENO7 Burgers prototype on the GPU (mgad_weno7_*) against the oracle's
restatement of Matlab_Prototipes/InviscidBurgersNd/LFWENO7FDM3d.m, bit for
bit: the right-hand side on ragged grids, and full low-storage RK steps to
past t_end."""
import numpy as np
import pytest

import oracle as O
from multigpu_advectiondiffusion_amd import solvers
from multigpu_advectiondiffusion_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu


# partial tiles in x and y (tiles are 16 x 16), single-node and thin axes
@pytest.mark.parametrize("shape", [(20, 20, 20), (37, 19, 11), (70, 33, 5), (1, 7, 9), (64, 48, 40)])
def test_rhs_bitwise(hipdev, shape):
    s = solvers.Weno7Burgers(shape)
    rng = np.random.default_rng(sum(shape))
    u = s.gaussian() + 0.05 * rng.standard_normal(s.gaussian().shape)
    du, dL = DeviceArray.from_host(u), DeviceArray(u.shape)
    s.rhs(du, dL)
    s.stream.synchronize()
    got = dL.download()
    ref = O.weno7_rhs(u, (s.dx, s.dy, s.dz))
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]


# The weights take the shortened division where every divisor (eps + B_k)^2
# lies in [2^-370, 2^370]: "huge" (|u| near 2^44 on half the planes, divisors
# near 2^386 there) sends those nodes to IEEE division, "zeros" gives
# divisors eps^2 on the zero half.
@pytest.mark.parametrize("case", ["huge", "zeros"])
def test_rhs_bitwise_division_paths(hipdev, case):
    shape = (24, 20, 18)
    s = solvers.Weno7Burgers(shape)
    rng = np.random.default_rng(11)
    u = s.gaussian() + 0.05 * rng.standard_normal(s.gaussian().shape)
    if case == "huge":
        u[: shape[2] // 2] *= 2.0 ** 44
    else:
        u[: shape[2] // 2] = 0.0
    du, dL = DeviceArray.from_host(u), DeviceArray(u.shape)
    s.rhs(du, dL)
    s.stream.synchronize()
    got = dL.download()
    ref = O.weno7_rhs(u, (s.dx, s.dy, s.dz))
    assert np.isfinite(ref).all()
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]


@pytest.mark.parametrize("graph", [False, True])
def test_steps_bitwise(hipdev, graph):
    shape = (40, 36, 28)
    s = solvers.Weno7Burgers(shape, cfl=0.4)
    u0 = s.gaussian()
    s.set_initial(u0, tend=0.05)
    s.steps(3, graph=graph)
    s.steps(20, graph=graph)  # runs to t_end; the steps after it are no-ops
    t, dt, n = s.state()
    ref, res = u0.copy(), np.zeros_like(u0)
    steps, tref = O.weno7_run(ref, res, (s.dx, s.dy, s.dz), 0.4, 0.0, 0.05)
    assert n == steps and t == tref
    got = s.solution()
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]
    assert np.array_equal(s.res.download(), res)
