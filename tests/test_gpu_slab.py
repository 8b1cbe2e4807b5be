"""The periodic Burgers run on z slabs with ghost planes (mgad_burgers_slab_*)
against the unsplit run: the oracle and the single-grid GPU path, bit for
bit, for 1 to 8 slabs; and the exchange's ghost planes against the
neighbours' boundary planes."""
import numpy as np
import pytest

import oracle as O
from multigpu_advectiondiffusion_amd import _capi, solvers
from multigpu_advectiondiffusion_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu

ORACLE_SCHEME = {"Z2": O.Z2_GLOBAL, "JS": O.JS_GLOBAL}


@pytest.mark.parametrize("nslab", [1, 2, 4, 8])
@pytest.mark.parametrize("scheme", ["Z2", "JS"])
@pytest.mark.parametrize("graph", [False, True])
def test_slabs_match_unsplit_run(hipdev, nslab, scheme, graph):
    shape = (32, 32, 32)
    u0 = solvers.ic_fourier(shape, seed=4)
    s = solvers.SlabBurgers(shape, nslab, scheme=scheme, cfl=0.4)
    s.set_initial(u0, tend=10.0)
    s.steps(4, graph=graph)
    s.steps(3, graph=graph)
    got = s.solution()
    ref = u0.copy()
    O.burgers_run(ref, (s.dx, s.dy, s.dz), 0.4, 0.0, 10.0, nsteps=7, **ORACLE_SCHEME[scheme])
    assert np.array_equal(got, ref), np.argwhere(got != ref)[:5]
    p = solvers.PeriodicBurgers(shape, scheme=scheme, cfl=0.4)
    p.set_initial(u0, tend=10.0)
    p.steps(7, graph=graph)
    assert np.array_equal(got, p.solution())


@pytest.mark.parametrize("nslab,nzs", [(1, 16), (2, 16), (4, 8), (8, 4)])
def test_slabs_full_plane_match_unsplit_run(hipdev, nslab, nzs):
    """256 x 256 planes: 1, 2 and 4 slabs take the default one-round launch of
    32x32 workgroups (one per CU); 8 slabs of 4 planes have too many tiles
    for one round and take 32x16 tiles.  Bit for bit against the unsplit
    grid."""
    shape = (256, 256, nslab * nzs)
    u0 = solvers.ic_fourier(shape, seed=12)
    s = solvers.SlabBurgers(shape, nslab, scheme="Z2", cfl=0.4)
    s.set_initial(u0, tend=10.0)
    s.steps(3, graph=True)
    p = solvers.PeriodicBurgers(shape, scheme="Z2", cfl=0.4)
    p.set_initial(u0, tend=10.0)
    p.steps(3, graph=True)
    assert np.array_equal(s.solution(), p.solution())


def test_slabs_uneven_shape_and_t_end(hipdev):
    """48 x 16 x 30 in 5 slabs of 6 planes, 16-wide tiles, run past t_end: the
    clipped last step and the no-op steps after it as on the unsplit grid."""
    shape = (48, 16, 30)
    u0 = solvers.ic_fourier(shape, seed=6)
    s = solvers.SlabBurgers(shape, 5, scheme="Z2", cfl=0.4)
    s.set_initial(u0, tend=0.05)
    s.steps(40, graph=True)
    p = solvers.PeriodicBurgers(shape, scheme="Z2", cfl=0.4)
    p.set_initial(u0, tend=0.05)
    p.steps(40, graph=True)
    st = p.state()
    assert st.steps < 40 and abs(st.t - 0.05) < 1e-15
    assert np.array_equal(s.solution(), p.solution())


def test_exchange_fills_ghosts(hipdev):
    nx, ny, nzs, nslab = 16, 16, 5, 3
    rng = np.random.default_rng(8)
    slabs = rng.standard_normal((nslab, nzs + 6, ny, nx))
    d = DeviceArray.from_host(slabs)
    rc = _capi.lib().mgad_slab_exchange(None, nx, ny, nzs, nslab, d.ptr)
    _capi.check(rc, "mgad_slab_exchange")
    got = d.download()
    for s in range(nslab):
        below, above = (s - 1) % nslab, (s + 1) % nslab
        assert np.array_equal(got[s, 0:3], slabs[below, nzs:nzs + 3])
        assert np.array_equal(got[s, nzs + 3:nzs + 6], slabs[above, 3:6])
        assert np.array_equal(got[s, 3:nzs + 3], slabs[s, 3:nzs + 3])


def test_slab_arguments_rejected(hipdev):
    d = DeviceArray((1, 8, 16, 16))
    assert _capi.lib().mgad_slab_exchange(None, 16, 16, 2, 1, d.ptr) == _capi.MGAD_ERR_INVALID_ARG
    assert _capi.last_error() == _capi.MGAD_ERR_INVALID_ARG
