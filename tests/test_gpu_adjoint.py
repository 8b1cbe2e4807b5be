"""Discrete adjoint of the periodic Burgers run on the GPU.

Checked against (a) central finite differences of the GPU forward run along
seeded unit directions v, formed point-wise as
sum 1/2 (u+ - u-)(u+ + u- - 2 target) / (2h) (the same quotient as
(J(u0+hv) - J(u0-hv))/(2h) without cancelling two large sums), and (b) the
CPU reference adjoint (oracle), itself pinned to finite differences in
test_oracle_adjoint.py.

The finite-difference quotient has its own error: rounding in the forward run
(amplified by the eps = 1e-40 weights) over 2h, which at h = 1e-6 reaches a
few 1e-7 of |grad| on these grids and shrinks ten-fold at h = 1e-5.  So the
check is 1e-6 relative to the projection <g, v> at h = 1e-5, and 1e-6
relative to |g| at h = 1e-6.
"""
import numpy as np
import pytest

import oracle as O
from multigpu_advectiondiffusion_amd import solvers

pytestmark = pytest.mark.gpu

ORC = {"Z2": O.Z2_GLOBAL, "JS": O.JS_GLOBAL}


def forward(n, scheme, u0, tend):
    s = solvers.PeriodicBurgers(n, scheme=scheme)
    s.set_initial(u0, tend=tend)
    s.run(batch=64)
    return s.solution()


def fd_directional(n, scheme, u0, target, v, tend, h=1e-6):
    up = forward(n, scheme, u0 + h * v, tend)
    um = forward(n, scheme, u0 - h * v, tend)
    return float(np.sum(0.5 * (up - um) * (up + um - 2 * target)) / (2 * h))


@pytest.mark.parametrize("scheme", ["Z2", "JS"])
@pytest.mark.parametrize("tend", [0.2, 0.5])
def test_adjoint_vs_oracle_and_fd(hipdev, scheme, tend):
    n = 32
    u0 = solvers.ic_fourier(n, seed=5)
    target = solvers.ic_fourier(n, seed=5) + 1e-2 * solvers.ic_fourier(n, seed=105)
    adj = solvers.BurgersAdjoint(n, scheme=scheme, max_steps=200)
    J, g, steps = adj(u0, target, tend)
    Jo, go, so = O.burgers_adjoint(u0, target, 2.0 / n, 0.4, 0.0, tend, **ORC[scheme])
    assert steps == so
    assert J == pytest.approx(Jo, rel=1e-12)
    assert np.max(np.abs(g - go)) <= 1e-9 * np.max(np.abs(go))
    rng = np.random.default_rng(7)
    gnorm = float(np.linalg.norm(g))
    for _ in range(4):
        v = rng.standard_normal(u0.shape)
        v /= np.linalg.norm(v)
        ad = float(np.sum(g * v))
        fd5 = fd_directional(n, scheme, u0, target, v, tend, h=1e-5)
        assert abs(fd5 - ad) <= 1e-6 * abs(ad), (fd5, ad)
        fd6 = fd_directional(n, scheme, u0, target, v, tend, h=1e-6)
        assert abs(fd6 - ad) <= 1e-6 * gnorm, (fd6, ad, gnorm)


# (48, 32, 8): one z chunk, whose lower and upper boundaries are the same
# periodic plane; (48, 32, 40): chunks of 5 planes, the shortest allowed.
# x has the finest spacing, which sets the CFL step.
@pytest.mark.parametrize("shape", [(48, 32, 8), (48, 32, 40)])
def test_adjoint_ragged_grids_vs_oracle(hipdev, shape):
    u0 = solvers.ic_fourier(shape, seed=11)
    target = solvers.ic_fourier(shape, seed=11) + 1e-2 * solvers.ic_fourier(shape, seed=12)
    adj = solvers.BurgersAdjoint(shape, scheme="Z2", max_steps=100)
    J, g, steps = adj(u0, target, 0.25)
    Jo, go, so = O.burgers_adjoint(u0, target, (adj.dx, adj.dy, adj.dz), 0.4, 0.0, 0.25, **ORC["Z2"])
    assert steps == so
    assert J == pytest.approx(Jo, rel=1e-12)
    assert np.max(np.abs(g - go)) <= 1e-9 * np.max(np.abs(go))


@pytest.mark.parametrize("every", [1, 3, 7, 1000])
def test_checkpoint_interval_does_not_change_gradient(hipdev, every):
    n = 32
    u0 = solvers.ic_fourier(n, seed=8)
    target = 0.9 * solvers.ic_fourier(n, seed=9)
    ref = solvers.BurgersAdjoint(n, scheme="Z2", max_steps=200, every=200)(u0, target, 0.3)
    got = solvers.BurgersAdjoint(n, scheme="Z2", max_steps=200, every=every)(u0, target, 0.3)
    assert got[0] == ref[0] and got[2] == ref[2]
    assert np.array_equal(got[1], ref[1])


def test_adjoint_too_many_steps_is_an_error(hipdev):
    from multigpu_advectiondiffusion_amd import _capi

    n = 32
    u0 = solvers.ic_fourier(n, seed=8)
    adj = solvers.BurgersAdjoint(n, scheme="Z2", max_steps=5)
    with pytest.raises(_capi.MgadError):
        adj(u0, u0, 0.5)


def test_autograd_function_gradcheck(hipdev):
    """The adjoint behind a torch.autograd.Function (CPU tensors): gradients
    from torch.autograd match directional finite differences."""
    from multigpu_advectiondiffusion_amd import _capi, hip

    _capi.lib()
    hip.lib()
    import torch

    n = 32
    target = solvers.ic_fourier(n, seed=12)
    adj = solvers.BurgersAdjoint(n, scheme="Z2", max_steps=200)

    class BurgersMisfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0):
            J, g, _ = adj(u0.detach().numpy(), target, 0.25)
            ctx.save_for_backward(torch.from_numpy(g))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            (g,) = ctx.saved_tensors
            return gJ * g

    u0 = torch.tensor(solvers.ic_fourier(n, seed=13), requires_grad=True)
    J = BurgersMisfit.apply(u0)
    (gu,) = torch.autograd.grad(J, u0)
    rng = np.random.default_rng(3)
    for _ in range(3):
        v = rng.standard_normal((n, n, n))
        v /= np.linalg.norm(v)
        fd = fd_directional(n, "Z2", u0.detach().numpy(), target, v, 0.25, h=1e-5)
        ad = float(torch.sum(gu * torch.from_numpy(v)))
        assert abs(fd - ad) <= 1e-6 * abs(ad)
