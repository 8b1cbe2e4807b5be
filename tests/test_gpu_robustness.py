"""Failure paths and grid shapes of the fused periodic path on the GPU.

- A NaN or an infinity in the field stops the run: the stage that reads it
  sets dt = 0, u keeps its last finite state, t and the step count stay, and
  mgad_burgers_read_ctrl reports MGAD_ERR_NONFINITE.  In a batch only that
  problem stops.
- Grids whose nx or ny is not a multiple of the tile width run with tiles
  that reach past the grid, bit for bit equal to the oracle.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from multigpu_advectiondiffusion_amd import _capi, solvers

pytestmark = pytest.mark.gpu


def _ctrl(s):
    st = _capi.BurgersState()
    rc = _capi.lib().mgad_burgers_read_ctrl(s.stream.handle, s.ctrl.ptr, ctypes.byref(st))
    _capi.last_error()
    return rc, st


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
@pytest.mark.parametrize("graph", [False, True])
def test_nonfinite_value_in_the_field_stops_the_run(hipdev, bad, graph):
    n = 32
    u0 = solvers.ic_fourier(n, seed=7)
    u0[5, 6, 7] = bad
    s = solvers.PeriodicBurgers(n, scheme="Z2")
    s.set_initial(u0, tend=1.0)
    s.steps(3, graph=graph)
    rc, st = _ctrl(s)
    assert rc == _capi.MGAD_ERR_NONFINITE
    assert st.dt == 0.0 and st.t == 0.0 and st.steps == 0
    assert np.array_equal(s.solution(), u0, equal_nan=True)
    with pytest.raises(_capi.MgadError, match="non-finite"):
        s.state()


def test_overflow_inside_a_step_stops_before_the_update(hipdev):
    """|u| = 1e200 is finite, so stage 1 runs, but its flux u^2/2 overflows:
    stage 1's output holds infinities or NaNs, stage 2 finds them in its
    input's max|u| and stops the step, and stage 3 (the only one that
    writes u) never runs."""
    n = 32
    u0 = solvers.ic_fourier(n, seed=8)
    u0[10, 3, 4] = 1e200
    s = solvers.PeriodicBurgers(n, scheme="JS")
    s.set_initial(u0, tend=1.0)
    s.steps(2, graph=False)
    rc, st = _ctrl(s)
    assert rc == _capi.MGAD_ERR_NONFINITE
    assert st.dt == 0.0 and st.t == 0.0 and st.steps == 0
    assert np.array_equal(s.solution(), u0)


def test_nonfinite_problem_stops_alone_in_a_batch(hipdev):
    sizes = [32, 48, 16]
    u0s = [solvers.ic_fourier(sz, seed=30 + i) for i, sz in enumerate(sizes)]
    u0s[1][3, 4, 5] = np.nan
    b = solvers.BurgersBatch(sizes, scheme="Z2")
    b.set_initial(u0s, tend=10.0, kchunk=_capi.BATCH_SPLIT_WIDE)
    b.steps(3, graph=True)
    got = b.solutions()
    for i, sz in enumerate(sizes):
        st = _capi.BurgersState()
        rc = _capi.lib().mgad_burgers_read_ctrl(b.stream.handle, b.fields[i]["ctrl"].ptr, ctypes.byref(st))
        _capi.last_error()
        if i == 1:
            assert rc == _capi.MGAD_ERR_NONFINITE and st.steps == 0
            assert np.array_equal(got[i], u0s[i], equal_nan=True)
        else:
            assert rc == _capi.MGAD_OK and st.steps == 3
            ref = u0s[i].copy()
            O.burgers_run(ref, 2.0 / sz, 0.4, 0.0, 10.0, nsteps=3, **O.Z2_GLOBAL)
            assert np.array_equal(got[i], ref), i


# nx, ny not multiples of 16 (and of 32): the last tile of a row or column
# reaches past the grid; 4 is the smallest extent the halo wrap allows.
SHAPES = [(40, 40, 40), (20, 36, 12), (100, 24, 10), (4, 4, 4), (33, 17, 9), (70, 50, 8), (16, 20, 6),
          (48, 5, 7)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("scheme,kw", [("Z2", O.Z2_GLOBAL), ("JS", O.JS_GLOBAL)])
def test_any_grid_size_matches_oracle(hipdev, shape, scheme, kw):
    nx, ny, nz = shape
    u0 = solvers.ic_fourier(shape, seed=nx + 3 * ny + 7 * nz)
    s = solvers.PeriodicBurgers(shape, scheme=scheme)
    s.set_initial(u0, tend=10.0)
    s.steps(3, graph=True)
    got, st = s.solution(), s.state()
    ref = u0.copy()
    steps, t = O.burgers_run(ref, (2.0 / nx, 2.0 / ny, 2.0 / nz), 0.4, 0.0, 10.0, nsteps=3, **kw)
    assert st.steps == steps == 3 and st.t == t
    assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))


@pytest.mark.parametrize("tile,kchunk", [(16, 4), (32, 0), (32, 5), (16, 0)])
def test_partial_tiles_every_launch_shape(hipdev, tile, kchunk):
    shape = (72, 40, 24)
    u0 = solvers.ic_fourier(shape, seed=12)
    ref = u0.copy()
    O.burgers_run(ref, (2.0 / 72, 2.0 / 40, 2.0 / 24), 0.4, 0.0, 10.0, nsteps=2, **O.Z2_GLOBAL)
    solvers.set_launch("Z2", shape, tile, kchunk)
    try:
        s = solvers.PeriodicBurgers(shape, scheme="Z2")
        s.set_initial(u0, tend=10.0)
        s.steps(2, graph=False)
        got = s.solution()
    finally:
        solvers.set_launch("Z2", shape, 0, 0)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("scheme,kw", [("Z2", O.Z2_GLOBAL), ("JS", O.JS_GLOBAL)])
def test_square_tiles_past_the_grid_match_oracle(hipdev, scheme, kw):
    """94 x 94 x 112: 3 x 3 tiles of 32 x 32 reach 2 columns and rows past the
    grid (4.3 % more columns), and 28 bands of 4 planes fill 252 CUs in one
    round, so the default launch runs 32 x 32 workgroups that reach past the
    grid, one per CU."""
    shape = (94, 94, 112)
    u0 = solvers.ic_fourier(shape, seed=17)
    s = solvers.PeriodicBurgers(shape, scheme=scheme)
    s.set_initial(u0, tend=10.0)
    s.steps(2, graph=True)
    ref = u0.copy()
    O.burgers_run(ref, (2.0 / 94, 2.0 / 94, 2.0 / 112), 0.4, 0.0, 10.0, nsteps=2, **kw)
    assert np.array_equal(s.solution(), ref)


def test_partial_tiles_fp32_and_autotune(hipdev):
    shape = (40, 24, 16)
    u0 = solvers.ic_fourier(shape, seed=13)
    s = solvers.PeriodicBurgers(shape, scheme="Z2", precision="f32")
    s.set_initial(u0, tend=10.0)
    tile, kc = solvers.autotune(s)
    # (0, 0): the default launch; -1: shrinking z bands
    assert (tile, kc) == (0, 0) or (tile in (16, 32) and (kc >= 4 or kc in (0, -1)))
    s.set_initial(u0, tend=10.0)
    s.steps(2)
    got = s.solution()
    solvers.set_launch("Z2", shape, 0, 0, precision="f32")
    ref = solvers.PeriodicBurgers(shape, scheme="Z2", precision="f32")
    ref.set_initial(u0, tend=10.0)
    ref.steps(2, graph=False)
    assert got.dtype == np.float32 and np.array_equal(got, ref.solution())
