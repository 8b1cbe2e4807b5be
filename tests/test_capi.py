"""CPU-side checks of the C ABI library: it loads, exports every symbol that
include/mgad.h declares, validates arguments without touching the device,
and its host initial-condition generators are deterministic."""
import os
import re

import numpy as np
import pytest

from multigpu_advectiondiffusion_amd import _capi, reference_api, solvers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mgad.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|char|size_t)\s*\*?\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_header():
    lib = _capi.lib()
    names = declared_functions()
    assert "Call_Adv_x" in names and "mgad_burgers_stage" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_capi.EXPORTS)
    assert lib.mgad_version().startswith(b"mgad")


def test_reference_names_mirrored():
    for n in ("CopyToConstantMemory", "Call_Adv_x", "Call_Adv_y", "Call_Adv_z", "Call_Diff_", "Call_sspRK",
              "Call_Lu3d", "Call_RK3d"):
        assert callable(getattr(reference_api, n))


def test_invalid_arguments_are_reported_without_device():
    L = _capi.lib()
    assert _capi.last_error() == 0
    L.Call_Adv_x(_capi.Dim3(1, 1, 1), _capi.Dim3(1, 1, 1), None, 6, 10, 10, 0.1, None, None)
    assert _capi.last_error() == _capi.MGAD_ERR_INVALID_ARG
    assert _capi.last_error() == 0  # cleared on read
    L.Call_sspRK(_capi.Dim3(), _capi.Dim3(), None, 10, 10, 10, 4, 0.1, None, None, None)
    assert _capi.last_error() == _capi.MGAD_ERR_INVALID_ARG
    sc = _capi.Scheme(0, 2, 1e-40)
    import ctypes

    rc = L.mgad_burgers_stage(None, ctypes.byref(sc), 32, 32, 32, 0.1, 0.1, 0.1, 4, None, None, None, None)
    assert rc == _capi.MGAD_ERR_INVALID_ARG
    rc = L.mgad_burgers_stage(None, ctypes.byref(sc), 32, 32, 2, 0.1, 0.1, 0.1, 1, 1, 1, 2, 3)
    assert rc == _capi.MGAD_ERR_INVALID_ARG  # nz < 4
    assert _capi.last_error() == _capi.MGAD_ERR_INVALID_ARG
    assert L.mgad_set_adv_scheme(5, 1, 1e-6, 0, 0.0) == _capi.MGAD_ERR_INVALID_ARG
    assert _capi.last_error() == _capi.MGAD_ERR_INVALID_ARG
    # launch rules: bands of kchunk > 0 planes, 0 (default) or shrinking bands; nothing else
    assert L.mgad_burgers_set_launch(_capi.F64, ctypes.byref(sc), 40, 40, 40, 32, _capi.BANDS_SHRINK) == 0
    assert L.mgad_burgers_set_launch(_capi.F64, ctypes.byref(sc), 40, 40, 40, 32, -2) == _capi.MGAD_ERR_INVALID_ARG
    assert L.mgad_burgers_set_launch(_capi.F64, ctypes.byref(sc), 40, 40, 40, 8, 0) == _capi.MGAD_ERR_INVALID_ARG
    assert L.mgad_burgers_set_launch(_capi.F64, ctypes.byref(sc), 40, 40, 40, 0, 0) == 0
    _capi.last_error()
    assert _capi.error_string(_capi.MGAD_ERR_UNSUPPORTED) == "unsupported grid shape"


def test_ic_fourier_deterministic():
    a = solvers.ic_fourier((16, 12, 8), seed=4)
    b = solvers.ic_fourier((16, 12, 8), seed=4)
    c = solvers.ic_fourier((16, 12, 8), seed=5)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # integer wavenumbers on the periodic grid average to zero: mean is the offset
    assert a.mean() == pytest.approx(0.5, abs=1e-12)
    assert a.shape == (8, 12, 16)


def test_ic_gauss_noise():
    u = solvers.ic_gauss_noise(20, sigma=0.1, noise=1e-3, seed=1)
    assert np.all(u[0] == 0) and np.all(u[-1] == 0) and np.all(u[:, :, 0] == 0) and np.all(u[:, -1] == 0)
    assert np.array_equal(u, solvers.ic_gauss_noise(20, sigma=0.1, noise=1e-3, seed=1))
    nx = 20
    i = np.arange(nx) / (nx - 1) - 0.5
    g = np.exp(-(i[None, None, :] ** 2 + i[None, :, None] ** 2 + i[:, None, None] ** 2) / 0.02)
    inner = (slice(1, -1),) * 3
    assert np.max(np.abs(u[inner] - g[inner])) <= 1e-3 + 1e-15


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_capi, "_lib", None)
    monkeypatch.setattr(_capi, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(_capi.MgadError):
        _capi.lib()


def test_static_batch_init_rejects_without_device():
    """mgad_batch_init_static checks its arguments and the work list's size
    against the static scratch before any device call: kchunk 0 (autotune) is
    not offered, and a 512^3 problem cut into items of 4 planes needs more
    than MGAD_BATCH_SCRATCH_BYTES (the field pointers are never touched)."""
    import ctypes

    L = _capi.lib()
    fake = _capi.Problem(0x1000, 0x2000, 0x3000, 0x4000, 512, 512, 512, 2 / 512, 2 / 512, 2 / 512)
    probs = (_capi.Problem * 1)(fake)
    sc = solvers.scheme_struct("Z2")
    out = _capi.Batch()
    args = (None, _capi.F64, ctypes.byref(sc), 1, probs)
    tail = (0.0, 1.0, 0.4, ctypes.byref(out))
    assert L.mgad_batch_init_static(*args, 0, *tail) == _capi.MGAD_ERR_INVALID_ARG
    assert _capi.last_error() == _capi.MGAD_ERR_INVALID_ARG
    assert L.mgad_batch_init_static(*args, 4, *tail) == _capi.MGAD_ERR_SCRATCH_FULL
    assert _capi.last_error() == _capi.MGAD_ERR_SCRATCH_FULL
    assert "scratch" in _capi.error_string(_capi.MGAD_ERR_SCRATCH_FULL)
