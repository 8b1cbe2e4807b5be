"""Batched persistent stages and launch configurations on the GPU.

Every problem of a batch must end bit-identical to the same problem advanced
alone (and to the CPU oracle), whatever tile width / z-chunk depth the launch
uses, because each node's arithmetic does not depend on the decomposition.
"""
import numpy as np
import pytest

import oracle as O
from multigpu_advectiondiffusion_amd import _capi, solvers

pytestmark = pytest.mark.gpu

SIZES = [16, 32, 48, (32, 16, 24), 16, 64, (48, 32, 8)]


def alone(shape, u0, scheme, nsteps, tend=10.0):
    s = solvers.PeriodicBurgers(shape, scheme=scheme)
    s.set_initial(u0, tend=tend)
    s.steps(nsteps, graph=False)
    return s.solution(), s.state()


# kchunk -1 / -2 are the split lists: SIZES hold 1712 16x16 tile planes, so
# the ~1024 shares of a full device are one or two planes long and most pieces
# start or end inside a column.  With 32x16 tiles, the 16- and 48-wide problems
# end each tile row in a tile that reaches past nx.
@pytest.mark.parametrize("scheme", ["Z2", "JS"])
@pytest.mark.parametrize("kchunk", [0, 8, 16, _capi.BATCH_SPLIT, _capi.BATCH_SPLIT_WIDE])
def test_batch_matches_single_problems(hipdev, scheme, kchunk):
    b = solvers.BurgersBatch(SIZES, scheme=scheme)
    u0s = [solvers.ic_fourier(s, seed=40 + i) for i, s in enumerate(SIZES)]
    b.set_initial(u0s, tend=10.0, kchunk=kchunk)
    assert b.batch.kchunk in (_capi.BATCH_SPLIT_WIDE, _capi.BATCH_SPLIT, 8, 16, 32) and b.batch.nitems > 0
    if kchunk:
        assert b.batch.kchunk == kchunk
    b.steps(4, graph=True)
    got = b.solutions()
    sts = b.states()
    for i, (shape, u0) in enumerate(zip(SIZES, u0s)):
        ref, st = alone(shape, u0, scheme, 4)
        assert np.array_equal(got[i], ref), (i, shape)
        assert sts[i].steps == 4 and sts[i].t == st.t


@pytest.mark.parametrize("kchunk", [16, _capi.BATCH_SPLIT, _capi.BATCH_SPLIT_WIDE])
def test_batch_fp32_matches_single_problems(hipdev, kchunk):
    """FP32 storage: the batched stages equal each FP32 problem advanced alone."""
    sizes = [32, 48, (16, 32, 24)]
    b = solvers.BurgersBatch(sizes, scheme="Z2", precision="f32")
    u0s = [solvers.ic_fourier(s, seed=80 + i) for i, s in enumerate(sizes)]
    b.set_initial(u0s, tend=10.0, kchunk=kchunk)
    b.steps(3, graph=True)
    got = b.solutions()
    for i, (shape, u0) in enumerate(zip(sizes, u0s)):
        s = solvers.PeriodicBurgers(shape, scheme="Z2", precision="f32")
        s.set_initial(u0, tend=10.0)
        s.steps(3, graph=False)
        ref = s.solution()
        assert got[i].dtype == np.float32 and np.array_equal(got[i], ref), (i, shape)


def test_batch_serial_path_and_oracle(hipdev):
    b = solvers.BurgersBatch(SIZES[:4], scheme="Z2")
    u0s = [solvers.ic_fourier(s, seed=60 + i) for i, s in enumerate(SIZES[:4])]
    b.set_initial(u0s, tend=10.0)
    b.steps_serial(3, graph=True)
    got = b.solutions()
    for i, (shape, u0) in enumerate(zip(SIZES[:4], u0s)):
        ref = u0.copy()
        nx, ny, nz = (shape,) * 3 if np.isscalar(shape) else shape
        O.burgers_run(ref, (2.0 / nx, 2.0 / ny, 2.0 / nz), 0.4, 0.0, 10.0, nsteps=3, **O.Z2_GLOBAL)
        assert np.array_equal(got[i], ref), i


@pytest.mark.parametrize("kchunk", [16, _capi.BATCH_SPLIT, _capi.BATCH_SPLIT_WIDE])
def test_batch_stops_each_problem_at_tend(hipdev, kchunk):
    sizes = [16, 32, 48]
    b = solvers.BurgersBatch(sizes, scheme="JS")
    u0s = [solvers.ic_fourier(s, seed=70 + i) for i, s in enumerate(sizes)]
    b.set_initial(u0s, tend=0.2, kchunk=kchunk)
    b.steps(200, graph=True)
    sts = b.states()
    got = b.solutions()
    for i, s in enumerate(sizes):
        assert sts[i].t == 0.2
        ref = u0s[i].copy()
        steps, t = O.burgers_run(ref, 2.0 / s, 0.4, 0.0, 0.2, **O.JS_GLOBAL)
        assert sts[i].steps == steps
        assert np.array_equal(got[i], ref)


@pytest.mark.parametrize("tile,kchunk", [(16, 4), (32, 8), (16, 64), (32, 0), (0, 5), (32, -1), (16, -1)])
def test_launch_configs_bit_identical(hipdev, tile, kchunk):
    shape = (64, 32, 64)
    u0 = solvers.ic_fourier(shape, seed=9)
    ref, _ = alone(shape, u0, "Z2", 2)
    solvers.set_launch("Z2", shape, tile, kchunk)
    try:
        got, _ = alone(shape, u0, "Z2", 2)
    finally:
        solvers.set_launch("Z2", shape, 0, 0)
    assert np.array_equal(got, ref)


def test_autotune(hipdev):
    shape = 64
    s = solvers.PeriodicBurgers(shape, scheme="Z2")
    u0 = solvers.ic_fourier(shape, seed=11)
    s.set_initial(u0, tend=10.0)
    tile, kc = solvers.autotune(s)
    # (0, 0): the default launch; -1: shrinking z bands
    assert (tile, kc) == (0, 0) or (tile in (16, 32) and (kc in (0, -1) or 4 <= kc <= 64))
    s.set_initial(u0, tend=10.0)
    s.steps(2)
    ref, _ = alone(shape, u0, "Z2", 2)
    assert np.array_equal(s.solution(), ref)
    solvers.set_launch("Z2", shape, 0, 0)
