"""Reference file format, initial conditions and error norms (host side).

save_binary3d / load_binary3d: SaveBinary3D of SingleGPU/Burgers3d_WENO5/
tools.cpp:59-87 (32-bit floats, x fastest, no header).  init3d_burgers and
init3d_heat: Call_Init3d (Burgers tools.cpp:92-140) and Call_Init3D
(SingleGPU/Diffusion3d_baselineCode/tools.cpp:92-143).  heat_error_norms:
CalcError (Diffusion3d tools.cpp:156-180) with its node index corrected.
Arrays are numpy (nz, ny, nx), x fastest as in the reference.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _capi


def _shape(u):
    nz, ny, nx = u.shape
    return nx, ny, nz


def save_binary3d(u: np.ndarray, path: str | os.PathLike | None = None) -> None:
    u = np.ascontiguousarray(u, dtype=np.float64)
    p = None if path is None else os.fsencode(path)
    _capi.check(_capi.lib().mgad_save_binary3d(u.ctypes.data_as(ctypes.c_void_p), *_shape(u), p),
                "mgad_save_binary3d")


def load_binary3d(path: str | os.PathLike, shape) -> np.ndarray:
    nx, ny, nz = (shape, shape, shape) if np.isscalar(shape) else shape
    u = np.empty((nz, ny, nx))
    _capi.check(_capi.lib().mgad_load_binary3d(u.ctypes.data_as(ctypes.c_void_p), nx, ny, nz, os.fsencode(path)),
                "mgad_load_binary3d")
    return u


def _init(name, ic, shape, dx, dy, dz):
    nx, ny, nz = (shape, shape, shape) if np.isscalar(shape) else shape
    u = np.empty((nz, ny, nx))
    _capi.check(getattr(_capi.lib(), name)(ic, u.ctypes.data_as(ctypes.c_void_p), dx, dy, dz, nx, ny, nz), name)
    return u


def init3d_burgers(ic: int, shape, dx: float, dy: float, dz: float) -> np.ndarray:
    return _init("mgad_init3d_burgers", ic, shape, dx, dy, dz)


def init3d_heat(ic: int, shape, dx: float, dy: float, dz: float) -> np.ndarray:
    return _init("mgad_init3d_heat", ic, shape, dx, dy, dz)


def heat_error_norms(u: np.ndarray, t: float, dx: float, dy: float, dz: float):
    """(L1, L2, Linf) of u against the heat kernel at time t."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(3)
    _capi.check(_capi.lib().mgad_heat_error_norms(u.ctypes.data_as(ctypes.c_void_p), t, dx, dy, dz, *_shape(u),
                                                  out.ctypes.data_as(ctypes.c_void_p)), "mgad_heat_error_norms")
    return tuple(out)


def init_subdomain(h_q: np.ndarray, n: int, nz_own: int) -> np.ndarray:
    """Init_subdomain (MultiGPU/Burgers3d_Baseline/Tools.c:181-199): slab n,
    planes n*nz_own .. n*nz_own + nz_own + 5 of the global field h_q
    (shape (P*nz_own + 6, Ny, Nx))."""
    h_q = np.ascontiguousarray(h_q, dtype=np.float64)
    _, ny, nx = h_q.shape
    if (n + 1) * nz_own + 6 > h_q.shape[0]:
        raise ValueError("slab n reaches past the global field")
    out = np.empty((nz_own + 6, ny, nx))
    _capi.check(_capi.lib().mgad_init_subdomain(h_q.ctypes.data_as(ctypes.c_void_p),
                                                out.ctypes.data_as(ctypes.c_void_p), n, nx, ny, nz_own),
                "mgad_init_subdomain")
    return out


def merge_domains(h_s_q: np.ndarray, h_q: np.ndarray, n: int) -> None:
    """Merge_domains (MultiGPU/Burgers3d_Baseline/Tools.c:204-223): writes
    the own planes 3 .. nz_own+2 of slab n into planes n*nz_own+3 .. of h_q,
    in place (h_q must be C-contiguous float64)."""
    h_s_q = np.ascontiguousarray(h_s_q, dtype=np.float64)
    nz_own = h_s_q.shape[0] - 6
    if h_q.dtype != np.float64 or not h_q.flags.c_contiguous or h_q.shape[1:] != h_s_q.shape[1:]:
        raise ValueError("h_q must be a C-contiguous float64 field of the slab's Nx, Ny")
    if nz_own < 1 or (n + 1) * nz_own + 6 > h_q.shape[0]:
        raise ValueError("slab n reaches past the global field")
    _, ny, nx = h_q.shape
    _capi.check(_capi.lib().mgad_merge_domains(h_s_q.ctypes.data_as(ctypes.c_void_p),
                                               h_q.ctypes.data_as(ctypes.c_void_p), n, nx, ny, nz_own),
                "mgad_merge_domains")
