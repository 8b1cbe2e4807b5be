"""Solver drivers over the C ABI: periodic WENO5 Burgers with global LF
splitting and SSP-RK3 (on-device dt), and the 7-point forward-Euler heat
equation.  Initial conditions are generated on the host by the library's
own generators so every consumer starts from bit-identical fields.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi
from .hip import DeviceArray, Stream

# Scheme presets: (weno, p, eps)
SCHEMES = {
    "Z2": (_capi.WENO_Z, 2, 1e-40),   # WENO5-Z, p = 2, eps = 1e-40 (Borges et al. 2008)
    "JS": (_capi.WENO_JS, 2, 1e-6),   # WENO5-JS, eps = 1e-6 (MultiGPU/Burgers3d_Baseline/Kernels.cu:112-220)
    "Z1": (_capi.WENO_Z, 1, 1e-6),    # the reference's WENO5-Z (SingleGPU/Burgers3d_WENO5/kernels.cu:142-192)
}


def scheme_struct(name: str) -> _capi.Scheme:
    w, p, e = SCHEMES[name]
    return _capi.Scheme(w, p, e)


def ic_fourier(n, seed: int, nmodes: int = 8, kmax: int = 4, amp: float = 0.5, offset: float = 0.5,
               lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """0.5 + sum of nmodes seeded Fourier modes on the periodic cube [lo,hi)^3."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    u = np.empty((nz, ny, nx))
    rc = _capi.lib().mgad_ic_fourier(u.ctypes.data_as(ctypes.c_void_p), nx, ny, nz, lo, hi, seed, nmodes, kmax,
                                     amp, offset)
    _capi.check(rc, "mgad_ic_fourier")
    return u


def ic_gauss_noise(n, sigma: float = 0.1, noise: float = 1e-3, seed: int = 1) -> np.ndarray:
    """Centred Gaussian bump on [0,1]^3 plus seeded uniform noise, boundary 0."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    u = np.empty((nz, ny, nx))
    rc = _capi.lib().mgad_ic_gauss_noise(u.ctypes.data_as(ctypes.c_void_p), nx, ny, nz, sigma, noise, seed)
    _capi.check(rc, "mgad_ic_gauss_noise")
    return u


@dataclass
class BurgersState:
    t: float
    dt: float
    tend: float
    cfl: float
    amax: float
    steps: int


class PeriodicBurgers:
    """u_t + div(u^2/2, u^2/2, u^2/2) = 0 on a periodic box, WENO5 + global
    LF + SSP-RK3, dt = cfl*dx/max|u^n| clipped to tend (MLB/LFWENO5FDM3d.m:71)."""

    def __init__(self, shape, scheme: str = "Z2", cfl: float = 0.4, lo: float = -1.0, hi: float = 1.0,
                 stream: Stream | None = None):
        self.nx, self.ny, self.nz = (shape, shape, shape) if np.isscalar(shape) else shape
        self.dx = (hi - lo) / self.nx
        self.dy = (hi - lo) / self.ny
        self.dz = (hi - lo) / self.nz
        self.cfl = cfl
        self.scheme = scheme_struct(scheme)
        self.stream = stream if stream is not None else Stream()
        shp = (self.nz, self.ny, self.nx)
        self.u = DeviceArray(shp)
        self.w1 = DeviceArray(shp)
        self.w2 = DeviceArray(shp)
        self.ctrl = DeviceArray((_capi.CTRL_BYTES,), np.uint8)

    @property
    def npts(self) -> int:
        return self.nx * self.ny * self.nz

    def set_initial(self, u0: np.ndarray, t0: float = 0.0, tend: float = float("inf")) -> None:
        self.u.upload(u0)
        rc = _capi.lib().mgad_burgers_init(self.stream.handle, self.nx, self.ny, self.nz, self.u.ptr,
                                           self.ctrl.ptr, t0, tend, self.cfl)
        _capi.check(rc, "mgad_burgers_init")

    def stage(self, stage: int, u_in: DeviceArray, u_n: DeviceArray, u_out: DeviceArray) -> None:
        rc = _capi.lib().mgad_burgers_stage(self.stream.handle, ctypes.byref(self.scheme), self.nx, self.ny,
                                            self.nz, self.dx, self.dy, self.dz, stage, u_in.ptr, u_n.ptr,
                                            u_out.ptr, self.ctrl.ptr)
        _capi.check(rc, "mgad_burgers_stage")

    def steps(self, nsteps: int, graph: bool = True) -> None:
        rc = _capi.lib().mgad_burgers_steps(self.stream.handle, ctypes.byref(self.scheme), self.nx, self.ny,
                                            self.nz, self.dx, self.dy, self.dz, self.u.ptr, self.w1.ptr,
                                            self.w2.ptr, self.ctrl.ptr, nsteps, 1 if graph else 0)
        _capi.check(rc, "mgad_burgers_steps")

    def state(self) -> BurgersState:
        s = _capi.BurgersState()
        rc = _capi.lib().mgad_burgers_read_ctrl(self.stream.handle, self.ctrl.ptr, ctypes.byref(s))
        _capi.check(rc, "mgad_burgers_read_ctrl")
        return BurgersState(s.t, s.dt, s.tend, s.cfl, s.amax, int(s.steps))

    def run(self, batch: int = 16, graph: bool = True, max_steps: int = 10**7) -> BurgersState:
        """Advance until t >= tend (steps after the end are no-ops on device)."""
        st = self.state()
        while st.t < st.tend and st.steps < max_steps:
            self.steps(min(batch, max_steps - st.steps), graph)
            st = self.state()
        return st

    def solution(self) -> np.ndarray:
        self.stream.synchronize()
        return self.u.download()


class Heat7:
    """u_t = C lap u, 7-point Laplacian (Compute_Laplace3d_O2) + forward Euler
    on the Dirichlet interior; kx = C/(12 dx^2) as SingleGPU/Diffusion3d_
    baselineCode/main.cpp:59-61 defines it."""

    def __init__(self, shape, C: float = 1.0, L: float = 1.0, r: float = 0.15, stream: Stream | None = None):
        self.nx, self.ny, self.nz = (shape, shape, shape) if np.isscalar(shape) else shape
        self.dx, self.dy, self.dz = L / (self.nx - 1), L / (self.ny - 1), L / (self.nz - 1)
        self.kx = C / (12 * self.dx * self.dx)
        self.ky = C / (12 * self.dy * self.dy)
        self.kz = C / (12 * self.dz * self.dz)
        self.dt = r * self.dx * self.dx / C
        self.stream = stream if stream is not None else Stream()
        shp = (self.nz, self.ny, self.nx)
        self.u = DeviceArray(shp)
        self.w = DeviceArray(shp)

    def set_initial(self, u0: np.ndarray) -> None:
        self.u.upload(u0)

    def steps(self, nsteps: int, graph: bool = True) -> None:
        rc = _capi.lib().mgad_heat7_fe_steps(self.stream.handle, self.nx, self.ny, self.nz, self.kx, self.ky,
                                             self.kz, self.dt, self.u.ptr, self.w.ptr, nsteps, 1 if graph else 0)
        _capi.check(rc, "mgad_heat7_fe_steps")

    def solution(self) -> np.ndarray:
        self.stream.synchronize()
        return self.u.download()
