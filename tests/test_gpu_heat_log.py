"""The fused 13-point SSP-RK3 heat step (mgad_heat13_rk3_steps, the
reference's Call_Lu3d + Call_RK3d sequence in one launch per stage) on the
runs of Matlab_Prototipes/DiffusionNd/TestingAccuracy.log's 3-D table: bit for
bit against the oracle's Call_Lu3d / Call_RK3d sequence, and every error norm
equal to the log's printed digits (matlab_heat_log.py describes the run)."""
import numpy as np
import pytest

import oracle as O
from matlab_heat_log import LOG_3D, case, norms, printed
from multigpu_advectiondiffusion_amd import reference_api as R
from multigpu_advectiondiffusion_amd import solvers

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [10, 20, 40, 80])
def test_fused_heat_reproduces_the_matlab_accuracy_log(hipdev, n):
    u0, ue, dx, kx, dt, steps = case(n)
    R.CopyToConstantMemory(kx, kx, kx)
    h = solvers.ReferenceHeat(n)
    h.set_initial(u0)
    h.steps(steps, dt, graph=True)
    got = h.solution()
    u, Lu = u0.copy(), np.zeros_like(u0)
    for _ in range(steps):
        uo = u.copy()
        for stage in (1, 2, 3):
            O.lu3d_13(u, Lu, kx, kx, kx)
            O.rk3d(u, uo, Lu, stage, dt)
    assert np.array_equal(got, u), np.argwhere(got != u)[:5]
    assert tuple(printed(v) for v in norms(got, ue, dx)) == LOG_3D[n]
