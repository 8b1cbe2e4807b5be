"""Host-side file format, reference initial conditions and error norms,
against numpy restatements of the reference macros (no GPU)."""
import numpy as np
import pytest

from multigpu_advectiondiffusion_amd import fileio


def grid(shape, dx, dy, dz):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return i, j, k, i * dx, j * dy, k * dz


def test_save_binary3d_layout(tmp_path):
    u = np.random.default_rng(0).standard_normal((5, 4, 3))
    p = tmp_path / "r.bin"
    fileio.save_binary3d(u, p)
    raw = p.read_bytes()
    assert len(raw) == 4 * u.size  # floats, no header
    assert raw == u.astype("<f4").tobytes()  # x fastest (C order of (nz, ny, nx)), little endian
    back = fileio.load_binary3d(p, (3, 4, 5))
    assert np.array_equal(back, u.astype(np.float32).astype(np.float64))


def test_save_binary3d_default_name(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    fileio.save_binary3d(np.ones((2, 2, 2)))
    assert (tmp_path / "result.bin").stat().st_size == 32


def test_load_rejects_wrong_size(tmp_path):
    p = tmp_path / "r.bin"
    fileio.save_binary3d(np.ones((2, 2, 2)), p)
    with pytest.raises(Exception):
        fileio.load_binary3d(p, (3, 2, 2))
    with pytest.raises(Exception):
        fileio.load_binary3d(p, (1, 2, 2))


def test_init3d_burgers():
    shape, dx = (20, 18, 16), 2.0 / 15
    i, j, k, x, y, z = grid(shape, dx, dx, dx)
    nz, ny, nx = shape
    u1 = fileio.init3d_burgers(1, (nx, ny, nz), dx, dx, dx)
    ref1 = ((i > 0.4 * nx) & (i < 0.6 * nx) & (j > 0.4 * ny) & (j < 0.6 * ny) & (k > 0.4 * nz) & (k < 0.6 * nz))
    assert np.array_equal(u1, ref1.astype(float))
    assert np.array_equal(fileio.init3d_burgers(2, (nx, ny, nz), dx, dx, dx), np.zeros(shape))
    u3 = fileio.init3d_burgers(3, (nx, ny, nz), dx, dx, dx)
    ref3 = np.exp(-((x - 1.0) ** 2 + (y - 1.0) ** 2 + (z - 1.0) ** 2) / 0.1)
    assert np.allclose(u3, ref3, rtol=1e-14, atol=0) and u3.max() <= 1.0


def test_init3d_heat():
    shape, dx = (12, 11, 10), 10.0 / 9
    nz, ny, nx = shape
    i, j, k, x, y, z = grid(shape, dx, dx, dx)
    bnd = (i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1) | (k == 0) | (k == nz - 1)
    u1 = fileio.init3d_heat(1, (nx, ny, nz), dx, dx, dx)
    assert np.array_equal(u1, np.isin(k, [0, 1, nz - 2, nz - 1]).astype(float))
    u2 = fileio.init3d_heat(2, (nx, ny, nz), dx, dx, dx)
    ref2 = np.where(bnd, 0.0, np.sin(np.pi * i * dx) * np.sin(np.pi * j * dx) * np.sin(np.pi * k * dx))
    assert np.allclose(u2, ref2, rtol=1e-13, atol=1e-15)
    u3 = fileio.init3d_heat(3, (nx, ny, nz), dx, dx, dx)
    r2 = (x - 5) ** 2 + (y - 5) ** 2 + (z - 5) ** 2
    assert np.allclose(u3, np.where(bnd, 0.0, np.exp(-r2 / 0.4)), rtol=1e-14, atol=0)


def test_heat_error_norms():
    shape, dx, t = (9, 8, 7), 10.0 / 6, 0.3
    i, j, k, x, y, z = grid(shape, dx, dx, dx)
    exact = (0.1 / t) ** 1.5 * np.exp(-((x - 5) ** 2 + (y - 5) ** 2 + (z - 5) ** 2) / (4 * t))
    assert np.allclose(fileio.heat_error_norms(exact, t, dx, dx, dx), 0.0, atol=1e-15)
    u = exact + np.random.default_rng(1).uniform(-1e-3, 1e-3, exact.shape)
    err = exact - u
    l1, l2, linf = fileio.heat_error_norms(u, t, dx, dx, dx)
    assert l1 == pytest.approx(dx ** 3 * np.abs(err).sum(), rel=1e-12)
    assert l2 == pytest.approx(np.sqrt(dx ** 3 * (err ** 2).sum()), rel=1e-12)
    assert linf == pytest.approx(np.abs(err).max(), rel=1e-12)
    # the node index matters (the reference compared every node against u[i])
    u2 = u.copy()
    u2[3, 2, 1] += 1.0
    assert fileio.heat_error_norms(u2, t, dx, dx, dx)[2] >= 1.0 - 2e-3


def test_subdomain_split_and_merge_round_trip():
    """Init_subdomain / Merge_domains (MB3/Tools.c:181-223) against their
    index formulas idx_3d = i + Nx j + Nx Ny (k + n _Nz), idx_sd = i + Nx j +
    Nx Ny k, written out here as plane slices: slabs overlap by 6 planes, and
    merging every slab's own planes rebuilds the interior of the field."""
    P, nz_own, ny, nx = 3, 5, 4, 7
    g = np.random.default_rng(4).standard_normal((P * nz_own + 6, ny, nx))
    slabs = [fileio.init_subdomain(g, n, nz_own) for n in range(P)]
    for n, s in enumerate(slabs):
        assert s.shape == (nz_own + 6, ny, nx)
        assert np.array_equal(s, g[n * nz_own:n * nz_own + nz_own + 6])
    for n in range(P - 1):  # neighbours share 6 planes: the ghosts of one are owned by the other
        assert np.array_equal(slabs[n][-6:], slabs[n + 1][:6])
    back = np.full_like(g, np.nan)
    for n, s in enumerate(slabs):
        fileio.merge_domains(s, back, n)
    assert np.array_equal(back[3:-3], g[3:-3])
    assert np.isnan(back[:3]).all() and np.isnan(back[-3:]).all()  # the outer ghosts are not written


def test_subdomain_bounds_are_errors():
    g = np.zeros((16, 2, 2))
    with pytest.raises(ValueError):
        fileio.init_subdomain(g, 2, 5)  # planes 10..20 of 16
    from multigpu_advectiondiffusion_amd import _capi
    with pytest.raises(_capi.MgadError):
        fileio.init_subdomain(np.zeros((6, 2, 2)), 0, 0)  # no own planes


def test_subdomain_reference_names():
    """The reference's own call shapes: Init_subdomain(h_q, h_s_q, rank, Nx,
    Ny, _Nz) and Merge_domains(h_s_recvbuff, h_u, i, Nx, Ny, _Nz)
    (MB3/main.c:98, 347), on flat host buffers."""
    from multigpu_advectiondiffusion_amd import reference_api as R
    Nx, Ny, _Nz, P = 6, 3, 4, 2
    h_u = np.arange(Nx * Ny * (P * _Nz + 6), dtype=np.float64)
    h_s_u = np.zeros(Nx * Ny * (_Nz + 6))
    R.Init_subdomain(h_u, h_s_u, 1, Nx, Ny, _Nz)
    assert np.array_equal(h_s_u, h_u[Nx * Ny * _Nz:Nx * Ny * (2 * _Nz + 6)])
    out = np.zeros_like(h_u)
    R.Merge_domains(h_s_u, out, 1, Nx, Ny, _Nz)
    lo, hi = Nx * Ny * (_Nz + 3), Nx * Ny * (2 * _Nz + 3)
    assert np.array_equal(out[lo:hi], h_u[lo:hi]) and not out[:lo].any() and not out[hi:].any()
    with pytest.raises(ValueError):
        R.Init_subdomain(h_u, h_s_u, 2, Nx, Ny, _Nz)  # past the end of h_u
