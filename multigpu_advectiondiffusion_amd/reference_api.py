"""Python mirror of the reference's device-wrapper interface.

Same function names and argument order as the reference's extern "C" block
(SingleGPU/Burgers3d_WENO5/Burgers.h:70-83 and
SingleGPU/Diffusion3d_baselineCode/acoustics3d.h:55-60); buffers are
DeviceArray objects (or raw device addresses), the stream a hip.Stream or
None.  Each call goes straight to libmgad.so; like the reference the calls
return nothing, and an argument error raises MgadError here (the C entry
point records it in mgad_get_last_error()).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from .hip import DeviceArray, Stream

_DIM = _capi.Dim3(1, 1, 1)


def _ptr(a):
    if isinstance(a, DeviceArray):
        return a.ptr
    return ctypes.c_void_p(int(a))


def _stream(s):
    if s is None:
        return None
    if isinstance(s, Stream):
        return s.handle
    return ctypes.c_void_p(int(s))


def _dim(d):
    if isinstance(d, _capi.Dim3):
        return d
    if d is None:
        return _DIM
    d = tuple(d) + (1,) * (3 - len(tuple(d)))
    return _capi.Dim3(*d)


def CopyToConstantMemory(kx: float, ky: float, kz: float) -> None:
    _capi.lib().CopyToConstantMemory(kx, ky, kz)


def _adv(name, numBlocks, threadsPerBlock, aStream, Nx, Ny, Nz, d, u, Lu):
    getattr(_capi.lib(), name)(_dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), Nx, Ny, Nz, d,
                               _ptr(u), _ptr(Lu))
    _capi.raise_last(name)


def Call_Adv_x(numBlocks, threadsPerBlock, aStream, Nx, Ny, Nz, dx, u, Lu) -> None:
    _adv("Call_Adv_x", numBlocks, threadsPerBlock, aStream, Nx, Ny, Nz, dx, u, Lu)


def Call_Adv_y(numBlocks, threadsPerBlock, aStream, Nx, Ny, Nz, dy, v, Lv) -> None:
    _adv("Call_Adv_y", numBlocks, threadsPerBlock, aStream, Nx, Ny, Nz, dy, v, Lv)


def Call_Adv_z(numBlocks, threadsPerBlock, aStream, Nx, Ny, Nz, dz, w, Lw) -> None:
    _adv("Call_Adv_z", numBlocks, threadsPerBlock, aStream, Nx, Ny, Nz, dz, w, Lw)


def Call_Diff_(numBlocks, threadsPerBlock, aStream, Nx, Ny, Nz, q, Lq) -> None:
    _capi.lib().Call_Diff_(_dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), Nx, Ny, Nz, _ptr(q), _ptr(Lq))
    _capi.raise_last("Call_Diff_")


def Call_sspRK(numBlocks, threadsPerBlock, aStream, Nx, Ny, Nz, step, dt, u, uo, Lu) -> None:
    _capi.lib().Call_sspRK(_dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), Nx, Ny, Nz, step, dt,
                           _ptr(u), _ptr(uo), _ptr(Lu))
    _capi.raise_last("Call_sspRK")


def Call_Lu3d(numBlocks, threadsPerBlock, aStream, nx, ny, nz, u, Lu) -> None:
    _capi.lib().Call_Lu3d(_dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), nx, ny, nz, _ptr(u), _ptr(Lu))
    _capi.raise_last("Call_Lu3d")


def Call_RK3d(numBlocks, threadsPerBlock, aStream, step, nx, ny, nz, dt, u, uo, Lu) -> None:
    _capi.lib().Call_RK3d(_dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), step, nx, ny, nz, dt,
                          _ptr(u), _ptr(uo), _ptr(Lu))
    _capi.raise_last("Call_RK3d")


# Pitched-memory variants (SingleGPU/Burgers3d_WENO5_PitchedMem/Burgers.h:73-77,
# SingleGPU/Diffusion3d_PitchedMem/acoustics3d.h:58-59): the reference's names
# and argument order, pitch = row length in elements.

def _call(name, *args):
    getattr(_capi.lib(), name)(*args)
    _capi.raise_last(name)


def Call_Adv_x_pitched(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, Nz, dx, u, Lu) -> None:
    _call("Call_Adv_x_pitched", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, Nz, dx,
          _ptr(u), _ptr(Lu))


def Call_Adv_y_pitched(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, Nz, dy, v, Lv) -> None:
    _call("Call_Adv_y_pitched", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, Nz, dy,
          _ptr(v), _ptr(Lv))


def Call_Adv_z_pitched(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, Nz, dz, w, Lw) -> None:
    _call("Call_Adv_z_pitched", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, Nz, dz,
          _ptr(w), _ptr(Lw))


def Call_Diff_pitched(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, Nz, q, Lq) -> None:
    _call("Call_Diff_pitched", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, Nz, _ptr(q),
          _ptr(Lq))


def Call_sspRK_pitched(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, Nz, step, dt, u, uo, Lu) -> None:
    _call("Call_sspRK_pitched", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, Nz, step,
          dt, _ptr(u), _ptr(uo), _ptr(Lu))


def Call_Lu3d_pitched(numBlocks, threadsPerBlock, aStream, pitch, nx, ny, nz, u, Lu) -> None:
    _call("Call_Lu3d_pitched", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, nx, ny, nz, _ptr(u),
          _ptr(Lu))


def Call_RK3d_pitched(numBlocks, threadsPerBlock, aStream, step, pitch, nx, ny, nz, dt, u, uo, Lu) -> None:
    _call("Call_RK3d_pitched", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), step, pitch, nx, ny, nz,
          dt, _ptr(u), _ptr(uo), _ptr(Lu))


# SingleGPU/Burgers3d_WENO5_SharedMem (Burgers.h:79-81): sweeps over every
# node with a zero pad beyond the array and the *(1/6 * (1/sum)) rounding.

def Call_Adv_x_sharedmem(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, Nz, dx, u, Lu) -> None:
    _call("Call_Adv_x_sharedmem", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, Nz, dx,
          _ptr(u), _ptr(Lu))


def Call_Adv_y_sharedmem(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, Nz, dy, v, Lv) -> None:
    _call("Call_Adv_y_sharedmem", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, Nz, dy,
          _ptr(v), _ptr(Lv))


def Call_Adv_z_sharedmem(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, Nz, dz, w, Lw) -> None:
    _call("Call_Adv_z_sharedmem", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, Nz, dz,
          _ptr(w), _ptr(Lw))


# SingleGPU/Diffusion3d_Blocking (acoustics3d.h:59): Lu on nodes 3..n-4,
# planes 3..nz-3, each axis summed from its low side.

def Call_Lu3d_blocking(numBlocks, threadsPerBlock, aStream, pitch, nx, ny, nz, u, Lu) -> None:
    _call("Call_Lu3d_blocking", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, nx, ny, nz,
          _ptr(u), _ptr(Lu))


# MultiGPU/Diffusion3d_Baseline (DiffusionMPICUDA.h:118-123): one rank's slab.

def Call_Diff_mpi(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, _Nz, kstart, kstop, diff_x, diff_y, diff_z,
                  q, Lq) -> None:
    _call("Call_Diff_mpi", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, _Nz, kstart,
          kstop, diff_x, diff_y, diff_z, _ptr(q), _ptr(Lq))


def Call_Diff_O2_mpi(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, _Nz, kstart, kstop, diff_x, diff_y,
                     diff_z, q, Lq) -> None:
    _call("Call_Diff_O2_mpi", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, _Nz, kstart,
          kstop, diff_x, diff_y, diff_z, _ptr(q), _ptr(Lq))


def Call_sspRK_mpi(numBlocks, threadsPerBlock, aStream, step, pitch, Nx, Ny, _Nz, dt, q, qo, Lq) -> None:
    _call("Call_sspRK_mpi", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), step, pitch, Nx, Ny, _Nz, dt,
          _ptr(q), _ptr(qo), _ptr(Lq))


def set_adv_scheme(weno: int = _capi.WENO_Z, p: int = 1, eps: float = 1e-6, lf: int = _capi.LF_LOCAL,
                   alpha: float = 0.0) -> None:
    """Extra Call_Adv_* parameters (defaults = the reference's compile-time choice)."""
    _capi.check(_capi.lib().mgad_set_adv_scheme(weno, p, eps, lf, alpha), "mgad_set_adv_scheme")


def lu3d_o2(u, Lu, nx, ny, nz, stream=None) -> None:
    """Compute_Laplace3d_O2 (7-point) with the CopyToConstantMemory coefficients."""
    _capi.check(_capi.lib().mgad_lu3d_o2(_stream(stream), nx, ny, nz, _ptr(u), _ptr(Lu)), "mgad_lu3d_o2")


def burgers_reference_step(u, uo, Lu, Nx, Ny, Nz, dx, dy, dz, dt, stream=None, viscous=True) -> None:
    """One time step in the reference driver's call order (SingleGPU/
    Burgers3d_WENO5/main.cpp:139-149): uo <- u, then for each RK stage
    Adv_x, Adv_y, Adv_z, Diff_, sspRK."""
    uo.copy_from(u, stream)
    for step in (1, 2, 3):
        Call_Adv_x(None, None, stream, Nx, Ny, Nz, dx, u, Lu)
        Call_Adv_y(None, None, stream, Nx, Ny, Nz, dy, u, Lu)
        Call_Adv_z(None, None, stream, Nx, Ny, Nz, dz, u, Lu)
        if viscous:
            Call_Diff_(None, None, stream, Nx, Ny, Nz, u, Lu)
        Call_sspRK(None, None, stream, Nx, Ny, Nz, step, dt, u, uo, Lu)


def heat_reference_step(u, uo, Lu, nx, ny, nz, dt, stream=None) -> None:
    """One time step of SingleGPU/Diffusion3d_baselineCode/main.cpp:108-121."""
    uo.copy_from(u, stream)
    for step in (1, 2, 3):
        Call_Lu3d(None, None, stream, nx, ny, nz, u, Lu)
        Call_RK3d(None, None, stream, step, nx, ny, nz, dt, u, uo, Lu)


# ---- MultiGPU/Burgers3d_Baseline device wrappers (BurgersMPI.h:112-132) ----

def Compute_Adv_x_Async(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, _Nz, kstart, kstop, dx, u, Lu) -> None:
    _call("Compute_Adv_x_Async", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, _Nz,
          kstart, kstop, dx, _ptr(u), _ptr(Lu))


def Compute_Adv_y_Async(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, _Nz, kstart, kstop, dy, u, Lu) -> None:
    _call("Compute_Adv_y_Async", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, _Nz,
          kstart, kstop, dy, _ptr(u), _ptr(Lu))


def Compute_Adv_z_Async(numBlocks, threadsPerBlock, aStream, pitch, Nx, Ny, _Nz, kstart, kstop, dz, u, Lu) -> None:
    _call("Compute_Adv_z_Async", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), pitch, Nx, Ny, _Nz,
          kstart, kstop, dz, _ptr(u), _ptr(Lu))


def Compute_sspRK3Async(numBlocks, threadsPerBlock, aStream, step, pitch, Nx, Ny, _Nz, dt, q, qo, Lq) -> None:
    _call("Compute_sspRK3Async", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream), step, pitch, Nx, Ny, _Nz,
          dt, _ptr(q), _ptr(qo), _ptr(Lq))


def CopyBoundaryRegionToGhostCellAsync(numBlocks, threadsPerBlock, aStream, d_s_q, d_send_buffer, Nx, Ny, _Nz,
                                       pitch, gc_pitch, side) -> None:
    _call("CopyBoundaryRegionToGhostCellAsync", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream),
          _ptr(d_s_q), _ptr(d_send_buffer), Nx, Ny, _Nz, pitch, gc_pitch, side)


def CopyGhostCellToBoundaryRegionAsync(numBlocks, threadsPerBlock, aStream, d_s_q, d_recv_buffer, Nx, Ny, _Nz,
                                       pitch, gc_pitch, side) -> None:
    _call("CopyGhostCellToBoundaryRegionAsync", _dim(numBlocks), _dim(threadsPerBlock), _stream(aStream),
          _ptr(d_s_q), _ptr(d_recv_buffer), Nx, Ny, _Nz, pitch, gc_pitch, side)


def _host(a: np.ndarray, n: int) -> int:
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous or a.size < n:
        raise ValueError(f"need a C-contiguous float64 host array of at least {n} values")
    return a.ctypes.data


def Init_subdomain(h_q, h_s_q, n, Nx, Ny, _Nz) -> None:
    """MultiGPU/Burgers3d_Baseline/Tools.c:181-199, host arrays: h_s_q gets
    planes n*_Nz .. n*_Nz+_Nz+5 of h_q."""
    xy = Nx * Ny
    _host(h_q, xy * ((n + 1) * _Nz + 6))
    _host(h_s_q, xy * (_Nz + 6))
    _call("mgad_init_subdomain", h_q.ctypes.data, h_s_q.ctypes.data, n, Nx, Ny, _Nz)


def Merge_domains(h_s_q, h_q, n, Nx, Ny, _Nz) -> None:
    """MultiGPU/Burgers3d_Baseline/Tools.c:204-223, host arrays: planes
    3.._Nz+2 of h_s_q go to planes n*_Nz+3.. of h_q."""
    xy = Nx * Ny
    _host(h_s_q, xy * (_Nz + 6))
    _host(h_q, xy * ((n + 1) * _Nz + 6))
    _call("mgad_merge_domains", h_s_q.ctypes.data, h_q.ctypes.data, n, Nx, Ny, _Nz)
