"""Pins of the oracle's restatements of the driver variants: the blocked heat
code's Call_Lu3d (orc_lu3d_blk), the shared-memory Burgers sweeps
(orc_adv_b5s) and the MPI heat code's Laplacians and RK stage (orc_md3_lap,
orc_md3_rk).

Each is checked by what the mathematics fixes (polynomial exactness with
integer data, where every operation is exact), by agreement with already
pinned restatements on the nodes they share, by a brute-force restatement of
the zero pad on tiny grids, and by sentinels on every node outside its update
set."""
import numpy as np
import pytest

import oracle as O

SENT = 7.25  # sentinel: nodes outside an update set must keep it


def _poly_field(shape, degs, seed=0):
    """Integer-valued sum of x^a, y^b, z^c monomials on index coordinates, and
    its Laplacian (also integer): every stencil sum on it is exact."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz, dtype=float), np.arange(ny, dtype=float), np.arange(nx, dtype=float),
                          indexing="ij")
    rng = np.random.default_rng(seed)
    u = np.zeros(shape)
    lap = np.zeros(shape)
    for var, deg in zip((x, y, z), degs):
        c = float(rng.integers(1, 4))
        u += c * var ** deg
        lap += c * deg * (deg - 1) * var ** max(deg - 2, 0)
    return u, lap


# ---------------------------------------------------------------------------
# orc_lu3d_blk
# ---------------------------------------------------------------------------
def test_lu3d_blk_region_and_exactness():
    shape = (9, 10, 11)  # nz, ny, nx
    u, lap = _poly_field(shape, (5, 4, 3), seed=1)
    Lu = np.full(shape, SENT)
    O.lu3d_blk(u, Lu, 1.0, 1.0, 1.0)
    nz, ny, nx = shape
    inside = np.zeros(shape, bool)
    inside[3:nz - 2, 3:ny - 3, 3:nx - 3] = True
    # -u_{-2} + 16u_{-1} - 30u + 16u_{+1} - u_{+2} = 12 f'' for degree <= 5
    assert np.array_equal(Lu[inside], 12.0 * lap[inside])
    assert np.all(Lu[~inside] == SENT)


def test_lu3d_blk_equals_call_diff_operator_on_zero_lu():
    """Same region and sums as Compute_Laplace_Async (orc_diff13_b5, pinned by
    polynomial exactness), stored instead of added to a zero Lu."""
    rng = np.random.default_rng(2)
    u = rng.uniform(-1, 1, (12, 11, 13))
    a = np.zeros_like(u)
    O.lu3d_blk(u, a, 0.31, 0.27, 0.22)
    b = np.zeros_like(u)
    O.diff13_b5(u, b, 0.31, 0.27, 0.22)
    assert np.array_equal(a + 0.0, b)


# ---------------------------------------------------------------------------
# orc_adv_b5s
# ---------------------------------------------------------------------------
def _brute_b5s(u, d, axis, Lu):
    """Every node of every line, the face below node i from nodes i-3..i+2 of
    the line with 0 outside, through the pinned face function."""
    v = np.moveaxis(u, 2 - axis, -1)  # lines along the last axis
    out = np.moveaxis(Lu.copy(), 2 - axis, -1)
    n = v.shape[-1]
    for idx in np.ndindex(v.shape[:-1]):
        line = np.concatenate([np.zeros(3), v[idx], np.zeros(3)])  # node q at q+3
        fp = np.array([O.split(x, 0.0, O.LF_LOCAL)[0] for x in line])
        fm = np.array([O.split(x, 0.0, O.LF_LOCAL)[1] for x in line])
        # face below node l: f+ of l-3..l+1 (line[l..l+4]), f- of l-2..l+2 (line[l+1..l+5])
        F = [O.face(fp[l:l + 5], fm[l + 1:l + 6], **O.B5S) for l in range(n + 1)]
        for i in range(n):
            q = (F[i + 1] - F[i]) / d
            out[idx + (i,)] = -q if axis == 0 else out[idx + (i,)] - q
    return np.ascontiguousarray(np.moveaxis(out, -1, 2 - axis))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_adv_b5s_brute_force_with_zero_pad(axis):
    rng = np.random.default_rng(3 + axis)
    u = rng.uniform(-1.5, 2.0, (5, 6, 7))
    Lu0 = rng.uniform(-1, 1, u.shape)
    got = Lu0.copy()
    O.adv_b5s(axis, u, got, 0.125)
    assert np.array_equal(got, _brute_b5s(u, 0.125, axis, Lu0))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_adv_b5s_interior_equals_base_sweep_with_b5s_faces(axis):
    """On nodes 3..n-4 of lines 3..n-4 the faces read no padding: the values
    equal the base code's line march (orc_adv_*) with the shared-memory
    rounding selected."""
    rng = np.random.default_rng(10 + axis)
    u = rng.uniform(-1, 2, (12, 13, 14))
    a = np.zeros_like(u)
    O.adv_b5s(axis, u, a, 0.1)
    b = np.zeros_like(u)
    O.adv(axis, u, b, 0.1, **O.B5S)
    inner = (slice(3, -3),) * 3
    assert np.array_equal(a[inner] + 0.0, b[inner] + 0.0)


def test_adv_b5s_constant_field_and_xy_exchange():
    u = np.full((8, 9, 10), 0.75)
    Lu = np.zeros_like(u)
    O.adv_b5s(0, u, Lu, 0.2)
    assert np.all(Lu[:, :, 3:-3] == 0.0)  # constant data: F is f(c) on every full stencil
    assert np.any(Lu[:, :, :3] != 0.0)    # the zero pad makes the edge faces differ
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, (6, 9, 9))
    a = np.zeros_like(u)
    O.adv_b5s(0, u, a, 0.3)
    ut = np.ascontiguousarray(np.swapaxes(u, 1, 2))
    b = np.zeros_like(ut)
    O.adv_b5s(1, ut, b, 0.3)
    assert np.array_equal(a + 0.0, np.swapaxes(b, 1, 2) + 0.0)


# ---------------------------------------------------------------------------
# orc_md3_lap, orc_md3_rk
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("order,degs", [(2, (3, 2, 3)), (4, (5, 4, 5))])
def test_md3_lap_exact_on_polynomials_and_plane_range(order, degs):
    nz, ny, nx, px = 12, 9, 10, 13
    u, lap = _poly_field((nz, ny, px), degs, seed=order)
    Lu = np.full(u.shape, SENT)
    kstart, kstop = 3, 8
    # O2: dx_ (u_w - 2u + u_e) ...: f'' exact for cubics; O4: dx_/12 (..) with
    # dx_ = 12 is 12 f'' summed over the axes, exact for degree <= 5
    c = 1.0 if order == 2 else 12.0
    O.md3_lap(order, u, Lu, kstart, kstop, c, c, c, nx=nx)
    inside = np.zeros(u.shape, bool)
    inside[kstart:kstop, 3:ny - 3, 3:nx - 3] = True
    want = (1.0 if order == 2 else 12.0) * lap
    assert np.array_equal(Lu[inside], want[inside])
    assert np.all(Lu[~inside] == SENT)


def test_md3_o2_agrees_with_7_point_operator():
    """12 kx (u_e - 2u + u_w) summed east first (orc_lu3d_o2) and the
    per-axis form summed west first agree to rounding on random data."""
    rng = np.random.default_rng(7)
    u = rng.uniform(-1, 1, (10, 11, 12))
    a = np.zeros_like(u)
    O.md3_lap(2, u, a, 1, 9, 1.5, 2.5, 3.5)
    b = np.zeros_like(u)
    O.lu3d_o2(u, b, 1.5 / 12, 2.5 / 12, 3.5 / 12)
    inner = (slice(1, 9), slice(3, -3), slice(3, -3))
    assert np.max(np.abs(a[inner] - b[inner])) <= 1e-14 * np.max(np.abs(b[inner]))
    assert not np.array_equal(a[inner], b[inner])  # different sums: last bits differ somewhere


def test_md3_o4_uses_only_diff_x():
    rng = np.random.default_rng(8)
    u = rng.uniform(-1, 1, (10, 9, 9))
    a = np.zeros_like(u)
    O.md3_lap(4, u, a, 2, 8, 0.7, 1.0, 2.0)
    b = np.zeros_like(u)
    O.md3_lap(4, u, b, 2, 8, 0.7, -5.0, 9.0)
    assert np.array_equal(a, b)


def test_md3_rk_region_every_plane():
    rng = np.random.default_rng(9)
    shape = (8, 9, 12)  # nz, ny, px with nx = 10
    q0, qo, Lq = (rng.uniform(-1, 1, shape) for _ in range(3))
    for step in (1, 2, 3):
        got = q0.copy()
        O.md3_rk(got, qo, Lq, step, 0.01, nx=10)
        every = q0.copy()
        O.mb3_rk(every, qo, Lq, step, 0.01, nx=12)  # the stage on every node (pinned)
        inside = np.zeros(shape, bool)
        inside[:, 3:9 - 3, 3:10 - 3] = True
        assert np.array_equal(got[inside], every[inside])
        assert np.array_equal(got[~inside], q0[~inside])
