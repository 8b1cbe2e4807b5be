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
                 stream: Stream | None = None, precision: str = "f64"):
        self.nx, self.ny, self.nz = (shape, shape, shape) if np.isscalar(shape) else shape
        self.dx = (hi - lo) / self.nx
        self.dy = (hi - lo) / self.ny
        self.dz = (hi - lo) / self.nz
        self.cfl = cfl
        self.scheme = scheme_struct(scheme)
        self.stream = stream if stream is not None else Stream()
        self.precision = {"f64": _capi.F64, "f32": _capi.F32}[precision]
        dtype = np.float64 if self.precision == _capi.F64 else np.float32
        shp = (self.nz, self.ny, self.nx)
        self.u = DeviceArray(shp, dtype)
        self.w1 = DeviceArray(shp, dtype)
        self.w2 = DeviceArray(shp, dtype)
        self.ctrl = DeviceArray((_capi.CTRL_BYTES,), np.uint8)

    @property
    def npts(self) -> int:
        return self.nx * self.ny * self.nz

    def set_initial(self, u0: np.ndarray, t0: float = 0.0, tend: float = float("inf")) -> None:
        """Upload u0 (rounded once to float for FP32 storage) and start a run at t0."""
        self.u.upload(u0, self.stream)
        rc = _capi.lib().mgad_burgers_init_p(self.precision, self.stream.handle, self.nx, self.ny, self.nz,
                                             self.u.ptr, self.ctrl.ptr, t0, tend, self.cfl)
        _capi.check(rc, "mgad_burgers_init")

    def stage(self, stage: int, u_in: DeviceArray, u_n: DeviceArray, u_out: DeviceArray) -> None:
        rc = _capi.lib().mgad_burgers_stage(self.stream.handle, ctypes.byref(self.scheme), self.nx, self.ny,
                                            self.nz, self.dx, self.dy, self.dz, stage, u_in.ptr, u_n.ptr,
                                            u_out.ptr, self.ctrl.ptr)
        _capi.check(rc, "mgad_burgers_stage")

    def steps(self, nsteps: int, graph: bool = True) -> None:
        rc = _capi.lib().mgad_burgers_steps_p(self.precision, self.stream.handle, ctypes.byref(self.scheme),
                                              self.nx, self.ny, self.nz, self.dx, self.dy, self.dz, self.u.ptr,
                                              self.w1.ptr, self.w2.ptr, self.ctrl.ptr, nsteps, 1 if graph else 0)
        _capi.check(rc, "mgad_burgers_steps_p")

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
        return self.u.download(self.stream)


def autotune(solver: "PeriodicBurgers") -> tuple:
    """Pick the fastest tile width / z-chunk depth for this solver's grid
    (times stage launches on its buffers); re-initialise afterwards."""
    t, k = ctypes.c_int(), ctypes.c_int()
    rc = _capi.lib().mgad_burgers_autotune(solver.stream.handle, solver.precision, ctypes.byref(solver.scheme),
                                           solver.nx, solver.ny, solver.nz, solver.dx, solver.dy, solver.dz,
                                           solver.u.ptr, solver.w1.ptr, solver.ctrl.ptr, ctypes.byref(t),
                                           ctypes.byref(k))
    _capi.check(rc, "mgad_burgers_autotune")
    return t.value, k.value


def set_launch(scheme: str, shape, tile: int = 0, kchunk: int = 0, precision: str = "f64") -> None:
    nx, ny, nz = (shape, shape, shape) if np.isscalar(shape) else shape
    prec = {"f64": _capi.F64, "f32": _capi.F32}[precision]
    rc = _capi.lib().mgad_burgers_set_launch(prec, ctypes.byref(scheme_struct(scheme)), nx, ny, nz, tile, kchunk)
    _capi.check(rc, "mgad_burgers_set_launch")


class SlabBurgers:
    """PeriodicBurgers on z slabs with 3 ghost planes each (mgad_burgers_slab_*):
    the data layout and per-stage halo exchange of the z-slab decomposition
    (MultiGPU/Burgers3d_Baseline/main.c:64-66), all slabs on one device.  The
    own planes equal PeriodicBurgers' field bit for bit."""

    def __init__(self, shape, nslab: int, scheme: str = "Z2", cfl: float = 0.4, lo: float = -1.0,
                 hi: float = 1.0, stream: Stream | None = None):
        self.nx, self.ny, self.nz = (shape, shape, shape) if np.isscalar(shape) else shape
        if self.nz % nslab:
            raise ValueError(f"nz={self.nz} is not a multiple of nslab={nslab}")
        self.nslab, self.nzs = nslab, self.nz // nslab
        self.dx, self.dy, self.dz = (hi - lo) / self.nx, (hi - lo) / self.ny, (hi - lo) / self.nz
        self.cfl = cfl
        self.scheme = scheme_struct(scheme)
        self.stream = stream if stream is not None else Stream()
        shp = (self.nslab, self.nzs + 6, self.ny, self.nx)
        self.u, self.w1, self.w2 = DeviceArray(shp), DeviceArray(shp), DeviceArray(shp)
        self.ctrl = DeviceArray((_capi.CTRL_BYTES,), np.uint8)

    def set_initial(self, u0: np.ndarray, t0: float = 0.0, tend: float = float("inf")) -> None:
        """Scatter u0 (nz, ny, nx) into the slabs' own planes; the ghosts are
        filled on the device by the exchange in mgad_burgers_slab_init."""
        slabs = np.zeros((self.nslab, self.nzs + 6, self.ny, self.nx))
        slabs[:, 3:3 + self.nzs] = np.asarray(u0, dtype=np.float64).reshape(self.nslab, self.nzs, self.ny, self.nx)
        self.u.upload(slabs, self.stream)
        rc = _capi.lib().mgad_burgers_slab_init(self.stream.handle, self.nx, self.ny, self.nzs, self.nslab,
                                                self.u.ptr, self.ctrl.ptr, t0, tend, self.cfl)
        _capi.check(rc, "mgad_burgers_slab_init")

    def steps(self, nsteps: int, graph: bool = True) -> None:
        rc = _capi.lib().mgad_burgers_slab_steps(self.stream.handle, ctypes.byref(self.scheme), self.nx, self.ny,
                                                 self.nzs, self.nslab, self.dx, self.dy, self.dz, self.u.ptr,
                                                 self.w1.ptr, self.w2.ptr, self.ctrl.ptr, nsteps, 1 if graph else 0)
        _capi.check(rc, "mgad_burgers_slab_steps")

    def slabs(self) -> np.ndarray:
        """The whole slab set, ghosts included: (nslab, nzs + 6, ny, nx)."""
        return self.u.download(self.stream)

    def solution(self) -> np.ndarray:
        """The own planes gathered into (nz, ny, nx)."""
        return self.slabs()[:, 3:3 + self.nzs].reshape(self.nz, self.ny, self.nx)


class Weno7Burgers:
    """The WENO7 Burgers prototype (Matlab_Prototipes/InviscidBurgersNd/
    LFWENO7FDM3d.m) on the device: node grid on [lo, hi]^3 with dx = (hi-lo)/(n-1),
    local LF, WENO7 faces, repeated-end extension, low-storage RK45,
    dt = cfl*dx/max|u| clipped to tend."""

    def __init__(self, shape, cfl: float = 0.4, lo: float = -1.0, hi: float = 1.0,
                 stream: Stream | None = None):
        self.nx, self.ny, self.nz = (shape, shape, shape) if np.isscalar(shape) else shape
        self.dx, self.dy, self.dz = ((hi - lo) / max(n - 1, 1) for n in (self.nx, self.ny, self.nz))
        self.lo, self.cfl = lo, cfl
        self.stream = stream if stream is not None else Stream()
        shp = (self.nz, self.ny, self.nx)
        self.u, self.w1, self.w2, self.res = (DeviceArray(shp) for _ in range(4))
        self.ctrl = DeviceArray((_capi.CTRL_BYTES,), np.uint8)

    def gaussian(self) -> np.ndarray:
        """The prototype's initial field exp(-(x^2+y^2+z^2)/0.1) (LFWENO7FDM3d.m:53)."""
        x = self.lo + self.dx * np.arange(self.nx)
        y = self.lo + self.dy * np.arange(self.ny)
        z = self.lo + self.dz * np.arange(self.nz)
        return 1.0 * np.exp(-(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2) / 0.1)

    def set_initial(self, u0: np.ndarray, t0: float = 0.0, tend: float = float("inf")) -> None:
        self.u.upload(u0, self.stream)
        rc = _capi.lib().mgad_weno7_init(self.stream.handle, self.nx, self.ny, self.nz, self.u.ptr, self.res.ptr,
                                         self.ctrl.ptr, t0, tend, self.cfl)
        _capi.check(rc, "mgad_weno7_init")

    def steps(self, nsteps: int, graph: bool = True) -> None:
        rc = _capi.lib().mgad_weno7_steps(self.stream.handle, self.nx, self.ny, self.nz, self.dx, self.dy, self.dz,
                                          self.u.ptr, self.w1.ptr, self.w2.ptr, self.res.ptr, self.ctrl.ptr, nsteps,
                                          1 if graph else 0)
        _capi.check(rc, "mgad_weno7_steps")

    def rhs(self, u: DeviceArray, out: DeviceArray) -> None:
        rc = _capi.lib().mgad_weno7_rhs(self.stream.handle, self.nx, self.ny, self.nz, self.dx, self.dy, self.dz,
                                        u.ptr, out.ptr)
        _capi.check(rc, "mgad_weno7_rhs")

    def state(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_ulonglong()
        rc = _capi.lib().mgad_weno7_read_ctrl(self.stream.handle, self.ctrl.ptr, ctypes.byref(t), ctypes.byref(dt),
                                              ctypes.byref(n))
        _capi.check(rc, "mgad_weno7_read_ctrl")
        return t.value, dt.value, n.value

    def solution(self) -> np.ndarray:
        return self.u.download(self.stream)


class BurgersBatch:
    """Independent periodic Burgers problems of uneven cubic sizes advanced
    together (one persistent launch per SSP-RK3 stage), each with its own dt."""

    def __init__(self, sizes, scheme: str = "Z2", cfl: float = 0.4, precision: str = "f64", lo: float = -1.0,
                 hi: float = 1.0, stream: Stream | None = None):
        self.sizes = [tuple(s) if not np.isscalar(s) else (s, s, s) for s in sizes]
        self.scheme = scheme_struct(scheme)
        self.cfl = cfl
        self.precision = {"f64": _capi.F64, "f32": _capi.F32}[precision]
        self.dtype = np.float64 if self.precision == _capi.F64 else np.float32
        self.stream = stream if stream is not None else Stream()
        self.fields = []
        probs = []
        for (nx, ny, nz) in self.sizes:
            shp = (nz, ny, nx)
            f = dict(u=DeviceArray(shp, self.dtype), w1=DeviceArray(shp, self.dtype), w2=DeviceArray(shp, self.dtype),
                     ctrl=DeviceArray((_capi.CTRL_BYTES,), np.uint8))
            self.fields.append(f)
            probs.append(_capi.Problem(f["u"].ptr, f["w1"].ptr, f["w2"].ptr, f["ctrl"].ptr, nx, ny, nz,
                                       (hi - lo) / nx, (hi - lo) / ny, (hi - lo) / nz))
        self.probs = (_capi.Problem * len(probs))(*probs)
        self.batch = _capi.Batch()
        self.workspace = None

    @property
    def npts(self) -> int:
        return int(sum(nx * ny * nz for nx, ny, nz in self.sizes))

    def set_initial(self, u0s, t0: float = 0.0, tend: float = float("inf"), kchunk: int = 0,
                    static_scratch: bool = False) -> None:
        """Upload the initial fields and build the work list: in a workspace
        allocated here (mgad_batch_init; kchunk 0 autotunes), or, with
        static_scratch, in the library's static device scratch
        (mgad_batch_init_static, no device allocation; kchunk must name the
        list, and the batch is valid until the next such call on the device)."""
        for f, u0 in zip(self.fields, u0s):
            f["u"].upload(u0, self.stream)
        L = _capi.lib()
        if static_scratch:
            self.workspace = None
            rc = L.mgad_batch_init_static(self.stream.handle, self.precision, ctypes.byref(self.scheme),
                                          len(self.sizes), self.probs, kchunk, t0, tend, self.cfl,
                                          ctypes.byref(self.batch))
            _capi.check(rc, "mgad_batch_init_static")
            return
        nbytes = L.mgad_batch_workspace_bytes(len(self.sizes), self.probs)
        self.workspace = DeviceArray((nbytes,), np.uint8)
        rc = L.mgad_batch_init(self.stream.handle, self.precision, ctypes.byref(self.scheme), len(self.sizes),
                               self.probs, self.workspace.ptr, nbytes, kchunk, t0, tend, self.cfl,
                               ctypes.byref(self.batch))
        _capi.check(rc, "mgad_batch_init")

    def steps(self, nsteps: int, graph: bool = True) -> None:
        rc = _capi.lib().mgad_batch_steps(self.stream.handle, ctypes.byref(self.batch), nsteps, 1 if graph else 0)
        _capi.check(rc, "mgad_batch_steps")

    def steps_serial(self, nsteps: int, graph: bool = True) -> None:
        rc = _capi.lib().mgad_batch_steps_serial(self.stream.handle, ctypes.byref(self.batch), self.probs, nsteps,
                                                 1 if graph else 0)
        _capi.check(rc, "mgad_batch_steps_serial")

    def states(self):
        out = []
        for f in self.fields:
            s = _capi.BurgersState()
            rc = _capi.lib().mgad_burgers_read_ctrl(self.stream.handle, f["ctrl"].ptr, ctypes.byref(s))
            _capi.check(rc, "mgad_burgers_read_ctrl")
            out.append(BurgersState(s.t, s.dt, s.tend, s.cfl, s.amax, int(s.steps)))
        return out

    def solutions(self):
        return [f["u"].download(self.stream) for f in self.fields]


class BurgersAdjoint:
    """dJ/du0 of J = 1/2 |u(tend) - target|^2 for the periodic Burgers run
    (WENO5 + global LF + SSP-RK3, CFL dt) by the library's discrete adjoint
    with checkpoints every `every` steps (default ceil(sqrt(max_steps)))."""

    def __init__(self, shape, scheme: str = "Z2", cfl: float = 0.4, max_steps: int = 1000, every: int | None = None,
                 lo: float = -1.0, hi: float = 1.0, stream: Stream | None = None):
        self.nx, self.ny, self.nz = (shape, shape, shape) if np.isscalar(shape) else shape
        self.dx, self.dy, self.dz = (hi - lo) / self.nx, (hi - lo) / self.ny, (hi - lo) / self.nz
        self.scheme = scheme_struct(scheme)
        self.cfl = cfl
        self.max_steps = max_steps
        self.every = every if every else int(np.ceil(np.sqrt(max_steps)))
        self.stream = stream if stream is not None else Stream()
        shp = (self.nz, self.ny, self.nx)
        self.u0, self.target, self.grad = DeviceArray(shp), DeviceArray(shp), DeviceArray(shp)
        nbytes = _capi.lib().mgad_adjoint_workspace_bytes(self.nx, self.ny, self.nz, max_steps, self.every)
        if nbytes == 0:
            raise _capi.MgadError("mgad_adjoint_workspace_bytes: invalid arguments")
        self.workspace = DeviceArray((nbytes,), np.uint8)

    def __call__(self, u0: np.ndarray, target: np.ndarray, tend: float, t0: float = 0.0):
        """Returns (J, dJ/du0, steps)."""
        self.u0.upload(u0, self.stream)
        self.target.upload(target, self.stream)
        J, n = ctypes.c_double(), ctypes.c_int()
        rc = _capi.lib().mgad_burgers_adjoint(self.stream.handle, ctypes.byref(self.scheme), self.nx, self.ny,
                                              self.nz, self.dx, self.dy, self.dz, self.cfl, t0, tend, self.u0.ptr,
                                              self.target.ptr, self.grad.ptr, ctypes.byref(J), self.max_steps,
                                              self.every, self.workspace.ptr, self.workspace.nbytes, ctypes.byref(n))
        _capi.check(rc, "mgad_burgers_adjoint")
        return J.value, self.grad.download(self.stream), n.value


class ReferenceBurgers:
    """The reference driver's Burgers step (SingleGPU/Burgers3d_WENO5/main.cpp:
    139-149: local-LF WENO5 sweeps, optional 13-point viscous term with the
    CopyToConstantMemory coefficients, SSP-RK3 on the interior, frozen 3-cell
    ghosts) as one fused launch per stage."""

    def __init__(self, shape, dx, dy, dz, scheme: str = "Z1", viscous: bool = True, stream: Stream | None = None):
        self.nx, self.ny, self.nz = (shape, shape, shape) if np.isscalar(shape) else shape
        self.dx, self.dy, self.dz = dx, dy, dz
        self.scheme = scheme_struct(scheme)
        self.viscous = viscous
        self.stream = stream if stream is not None else Stream()
        shp = (self.nz, self.ny, self.nx)
        self.u, self.w1, self.w2 = DeviceArray(shp), DeviceArray(shp), DeviceArray(shp)

    def set_initial(self, u0: np.ndarray) -> None:
        self.u.upload(np.ascontiguousarray(u0, dtype=np.float64), self.stream)

    def steps(self, nsteps: int, dt: float, graph: bool = True) -> None:
        rc = _capi.lib().mgad_burgers_ref_steps(self.stream.handle, ctypes.byref(self.scheme),
                                                1 if self.viscous else 0, self.nx, self.ny, self.nz, self.dx,
                                                self.dy, self.dz, dt, self.u.ptr, self.w1.ptr, self.w2.ptr, nsteps,
                                                1 if graph else 0)
        _capi.check(rc, "mgad_burgers_ref_steps")

    def solution(self) -> np.ndarray:
        return self.u.download(self.stream)


class ReferenceHeat:
    """The reference driver's heat step (SingleGPU/Diffusion3d_baselineCode/
    main.cpp:108-121: 13-point Laplacian with the CopyToConstantMemory
    coefficients and SSP-RK3 on the interior, two frozen ghost layers) as one
    fused launch per stage."""

    def __init__(self, shape, stream: Stream | None = None):
        self.nx, self.ny, self.nz = (shape, shape, shape) if np.isscalar(shape) else shape
        self.stream = stream if stream is not None else Stream()
        shp = (self.nz, self.ny, self.nx)
        self.u, self.w1, self.w2 = DeviceArray(shp), DeviceArray(shp), DeviceArray(shp)

    def set_initial(self, u0: np.ndarray) -> None:
        self.u.upload(np.ascontiguousarray(u0, dtype=np.float64), self.stream)

    def steps(self, nsteps: int, dt: float, graph: bool = True) -> None:
        rc = _capi.lib().mgad_heat13_rk3_steps(self.stream.handle, self.nx, self.ny, self.nz, dt, self.u.ptr,
                                               self.w1.ptr, self.w2.ptr, nsteps, 1 if graph else 0)
        _capi.check(rc, "mgad_heat13_rk3_steps")

    def solution(self) -> np.ndarray:
        return self.u.download(self.stream)


def heat7_params(shape, C: float = 1.0, L: float = 1.0, r: float = 0.15) -> dict:
    """Node-inclusive spacing dx = L/(n-1) (SingleGPU/Diffusion3d_baselineCode/
    main.cpp:55-57), coefficients kx = C/(12 dx^2) (:59-61), dt = r dx^2 / C."""
    nx, ny, nz = (shape, shape, shape) if np.isscalar(shape) else shape
    dx, dy, dz = L / (nx - 1), L / (ny - 1), L / (nz - 1)
    return dict(nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, kx=C / (12 * dx * dx), ky=C / (12 * dy * dy),
                kz=C / (12 * dz * dz), dt=r * dx * dx / C)


class Heat7:
    """u_t = C lap u, 7-point Laplacian (Compute_Laplace3d_O2) + forward Euler
    on the Dirichlet interior; kx = C/(12 dx^2) as SingleGPU/Diffusion3d_
    baselineCode/main.cpp:59-61 defines it."""

    def __init__(self, shape, C: float = 1.0, L: float = 1.0, r: float = 0.15, stream: Stream | None = None):
        p = heat7_params(shape, C, L, r)
        self.nx, self.ny, self.nz = p["nx"], p["ny"], p["nz"]
        self.dx, self.dy, self.dz = p["dx"], p["dy"], p["dz"]
        self.kx, self.ky, self.kz, self.dt = p["kx"], p["ky"], p["kz"], p["dt"]
        self.stream = stream if stream is not None else Stream()
        shp = (self.nz, self.ny, self.nx)
        self.u = DeviceArray(shp)
        self.w = DeviceArray(shp)

    def set_initial(self, u0: np.ndarray) -> None:
        self.u.upload(u0, self.stream)

    def steps(self, nsteps: int, graph: bool = True) -> None:
        rc = _capi.lib().mgad_heat7_fe_steps(self.stream.handle, self.nx, self.ny, self.nz, self.kx, self.ky,
                                             self.kz, self.dt, self.u.ptr, self.w.ptr, nsteps, 1 if graph else 0)
        _capi.check(rc, "mgad_heat7_fe_steps")

    def solution(self) -> np.ndarray:
        return self.u.download(self.stream)
