"""The allocation-free batch set-up, mgad_batch_init_static, against
mgad_batch_init with a workspace allocated by the caller.

Both write the same work list (one in the library's static device scratch,
one in the caller's buffer), so every problem must end bit-identical.  Two
static set-ups run back to back, the second with a smaller batch, so the
second run reads the scratch after a larger list was left in it.
"""
import ctypes

import numpy as np
import pytest

from multigpu_advectiondiffusion_amd import _capi, solvers

pytestmark = pytest.mark.gpu

FIRST = [48, 32, (32, 16, 24), 64, 16, (48, 32, 8)]
SECOND = [32, (16, 16, 8)]


def _run(sizes, seed0, kchunk, static, nsteps=3):
    b = solvers.BurgersBatch(sizes, scheme="Z2")
    u0s = [solvers.ic_fourier(s, seed=seed0 + i) for i, s in enumerate(sizes)]
    b.set_initial(u0s, tend=10.0, kchunk=kchunk, static_scratch=static)
    b.steps(nsteps, graph=False)
    return b.solutions(), b.states(), (b.batch.kchunk, b.batch.nitems, b.batch.grid), b


@pytest.mark.parametrize("kchunk", [_capi.BATCH_SPLIT_WIDE, _capi.BATCH_SPLIT, 16])
def test_static_batch_matches_workspace_batch(hipdev, kchunk):
    got1, st1, shape1, b1 = _run(FIRST, 90, kchunk, True)
    assert b1.workspace is None and b1.batch.workspace_bytes == _capi.BATCH_SCRATCH_BYTES
    got2, st2, shape2, _ = _run(SECOND, 97, kchunk, True)  # smaller list over the first one's
    for sizes, seed0, got, sts, shape in ((FIRST, 90, got1, st1, shape1), (SECOND, 97, got2, st2, shape2)):
        ref, rst, rshape, rb = _run(sizes, seed0, kchunk, False)
        assert rb.workspace is not None
        assert shape == rshape and shape[0] == kchunk
        for i in range(len(sizes)):
            assert got[i].shape == ref[i].shape and np.array_equal(got[i], ref[i]), (sizes[i], i)
            assert sts[i].steps == rst[i].steps == 3 and sts[i].t == rst[i].t


def test_static_batch_rejects_lists_above_the_scratch(hipdev):
    """One 512^3 problem in items of 4 planes is a 131072-item list (1 MiB of
    item arrays alone): rejected before any device work, so the field
    pointers are never touched."""
    L = _capi.lib()
    fake = _capi.Problem(0x1000, 0x2000, 0x3000, 0x4000, 512, 512, 512, 2 / 512, 2 / 512, 2 / 512)
    probs = (_capi.Problem * 1)(fake)
    sc = solvers.scheme_struct("Z2")
    out = _capi.Batch()
    rc = L.mgad_batch_init_static(None, _capi.F64, ctypes.byref(sc), 1, probs, 4, 0.0, 1.0, 0.4,
                                  ctypes.byref(out))
    assert rc == _capi.MGAD_ERR_SCRATCH_FULL
    assert L.mgad_get_last_error() == _capi.MGAD_ERR_SCRATCH_FULL
    assert "scratch" in _capi.error_string(rc)
    # kchunk 0 (autotune) needs HIP events: not offered without allocation
    rc = L.mgad_batch_init_static(None, _capi.F64, ctypes.byref(sc), 1, probs, 0, 0.0, 1.0, 0.4,
                                  ctypes.byref(out))
    assert rc == _capi.MGAD_ERR_INVALID_ARG
    L.mgad_get_last_error()
