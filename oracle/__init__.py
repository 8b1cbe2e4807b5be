"""CPU oracle: ctypes bindings to oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker.  The product package
never imports it.

Every function is a restatement of a reference routine (file:line cited in
oracle.c) evaluated in IEEE binary64 with the reference's expression order.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

WENO_Z, WENO_JS, WENO_JS_B5, WENO_Z_B5S = 0, 1, 2, 3
LF_LOCAL, LF_GLOBAL = 0, 1


class Scheme(ctypes.Structure):
    _fields_ = [
        ("weno", ctypes.c_int),
        ("p", ctypes.c_int),
        ("eps", ctypes.c_double),
        ("lf", ctypes.c_int),
        ("alpha", ctypes.c_double),
    ]


def scheme(weno=WENO_Z, p=1, eps=1e-6, lf=LF_LOCAL, alpha=0.0) -> Scheme:
    return Scheme(weno, p, eps, lf, alpha)


# Reference defaults (B5/Burgers.h:26, B5/kernels.cu:142-192, local LF).
REFERENCE = dict(weno=WENO_Z, p=1, eps=1e-6, lf=LF_LOCAL)
# WENO5-JS, eps 1e-6, global LF (periodic Burgers, JS configuration).
JS_GLOBAL = dict(weno=WENO_JS, p=2, eps=1e-6, lf=LF_GLOBAL)
# WENO5-Z p=2, eps 1e-40, global LF (periodic Burgers, Z configuration).
Z2_GLOBAL = dict(weno=WENO_Z, p=2, eps=1e-40, lf=LF_GLOBAL)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


_lib = None
_D = ctypes.c_double
_U = ctypes.c_uint
_P = ctypes.POINTER(ctypes.c_double)
_S = ctypes.POINTER(Scheme)


_NATIVE_PATH = os.path.join(_HERE, "_native", "liboracle_native.so")
_native = None


def native_lib():
    """The same oracle built -O3 -march=native for this machine (built here on
    first use; for timing the CPU baseline, never shipped)."""
    global _native
    if _native is None:
        subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True)
        _native = _setup(ctypes.CDLL(_NATIVE_PATH))
    return _native


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _setup(ctypes.CDLL(_LIB_PATH))
    return _lib


def _setup(L):
    """Declares the argument and result types of every oracle function."""
    L.orc_flux.restype = _D
    L.orc_flux.argtypes = [_D]
    L.orc_split.argtypes = [_D, _D, ctypes.c_int, _P, _P]
    L.orc_face.restype = _D
    L.orc_face.argtypes = [_P, _P, _S]
    L.orc_weno_plus.restype = _D
    L.orc_weno_plus.argtypes = [_P, _S]
    L.orc_weno_minus.restype = _D
    L.orc_weno_minus.argtypes = [_P, _S]
    for name in ("orc_adv_x", "orc_adv_y", "orc_adv_z"):
        getattr(L, name).argtypes = [_U, _U, _U, _D, _P, _P, _S]
    L.orc_diff13_b5.argtypes = [_U, _U, _U, _D, _D, _D, _P, _P]
    L.orc_ssprk_b5.argtypes = [_U, _U, _U, _U, _D, _P, _P, _P]
    L.orc_lu3d_13.argtypes = [_U, _U, _U, _D, _D, _D, _P, _P]
    L.orc_lu3d_o2.argtypes = [_U, _U, _U, _D, _D, _D, _P, _P]
    L.orc_rk3d.argtypes = [_U, _U, _U, _U, _D, _P, _P, _P]
    L.orc_weno7_face.argtypes = [_P, _P, _P, _P]
    L.orc_weno7_rhs.argtypes = [_U, _U, _U, _D, _D, _D, _P, _P]
    L.orc_lsrk_stage.argtypes = [ctypes.c_uint64, ctypes.c_int, _D, _P, _P, _P]
    L.orc_weno7_run.restype = ctypes.c_int
    L.orc_weno7_run.argtypes = [_U, _U, _U, _D, _D, _D, _D, _P, _D, ctypes.c_int, _P, _P]
    L.orc_mb3_adv.argtypes = [ctypes.c_int, _U, _U, _U, _U, _U, _U, _D, _P, _P, _S]
    L.orc_mb3_rk.argtypes = [_U, _U, _U, _U, _U, _D, _P, _P, _P]
    L.orc_lu3d_blk.argtypes = [_U, _U, _U, _D, _D, _D, _P, _P]
    L.orc_adv_b5s.argtypes = [ctypes.c_int, _U, _U, _U, _D, _P, _P, _S]
    L.orc_md3_lap.argtypes = [ctypes.c_int, _U, _U, _U, _U, _U, _U, _D, _D, _D, _P, _P]
    L.orc_md3_rk.argtypes = [_U, _U, _U, _U, _U, _D, _P, _P, _P]
    L.orc_maxabs.restype = _D
    L.orc_maxabs.argtypes = [ctypes.c_uint64, _P]
    L.orc_rhs_periodic.argtypes = [_U, _U, _U, _D, _D, _D, _P, _P, _S]
    L.orc_rk_stage_all.argtypes = [ctypes.c_uint64, _U, _D, _P, _P, _P]
    L.orc_burgers_step.restype = _D
    L.orc_burgers_step.argtypes = [_U, _U, _U, _D, _D, _D, _D, _D, _D, _S, _P, _P]
    L.orc_burgers_run.restype = ctypes.c_int
    L.orc_burgers_run.argtypes = [_U, _U, _U, _D, _D, _D, _D, _P, _D, ctypes.c_int, _S, _P]
    L.orc_heat7_fe.argtypes = [_U, _U, _U, _D, _D, _D, _D, ctypes.c_int, _P]
    L.orc_rhs_1d_periodic.argtypes = [_U, _D, _P, _P, _S]
    L.orc_side_grad.restype = _D
    L.orc_side_grad.argtypes = [_P, ctypes.c_int, _S, _P]
    L.orc_rhs_vjp.argtypes = [_U, _U, _U, _D, _D, _D, _P, _P, _P, _S, _P, _P]
    L.orc_burgers_adjoint.restype = _D
    L.orc_burgers_adjoint.argtypes = [_U, _U, _U, _D, _D, _D, _D, _D, _D, _S, _P, _P, _P,
                                      ctypes.POINTER(ctypes.c_int)]
    L.orc_num_threads.restype = ctypes.c_int
    return L


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def _shape(u: np.ndarray):
    """(nx, ny, nz) of an array stored x-fastest as u[k, j, i]."""
    nz, ny, nx = u.shape
    return nx, ny, nz


def num_threads(native: bool = False) -> int:
    return (native_lib() if native else lib()).orc_num_threads()


def flux(u: float) -> float:
    return lib().orc_flux(u)


def split(u: float, a: float = 0.0, lf: int = LF_LOCAL):
    fp, fm = ctypes.c_double(), ctypes.c_double()
    lib().orc_split(u, a, lf, ctypes.byref(fp), ctypes.byref(fm))
    return fp.value, fm.value


def face(v, w, **kw) -> float:
    v = np.ascontiguousarray(v, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    s = scheme(**kw)
    return lib().orc_face(_p(v), _p(w), ctypes.byref(s))


def weno_plus(v, **kw) -> float:
    v = np.ascontiguousarray(v, dtype=np.float64)
    s = scheme(**kw)
    return lib().orc_weno_plus(_p(v), ctypes.byref(s))


def weno_minus(w, **kw) -> float:
    w = np.ascontiguousarray(w, dtype=np.float64)
    s = scheme(**kw)
    return lib().orc_weno_minus(_p(w), ctypes.byref(s))


def adv(axis: int, u: np.ndarray, Lu: np.ndarray, d: float, **kw) -> None:
    """Reference Call_Adv_{x,y,z} sweep (in place on Lu)."""
    s = scheme(**kw)
    fn = (lib().orc_adv_x, lib().orc_adv_y, lib().orc_adv_z)[axis]
    fn(*_shape(u), d, _p(u), _p(Lu), ctypes.byref(s))


# MultiGPU/Burgers3d_Baseline (BurgersMPI.h:21-29, Kernels.cu:112-262):
# Reconstruct1d (WENO5-JS), EPS 1e-6, local LF.
MB3 = dict(weno=WENO_JS, p=2, eps=1e-6, lf=LF_LOCAL)


def mb3_adv(axis: int, u: np.ndarray, Lu: np.ndarray, kstart: int, kstop: int, d: float, pitch: int | None = None,
            nx: int | None = None) -> None:
    """Compute_Adv_{x,y,z}_Async of one rank's slab (in place on Lu); u and Lu
    are (nz, ny, pitch) arrays of which the first nx columns are the grid."""
    nz, ny, px = u.shape
    s = scheme(**MB3)
    lib().orc_mb3_adv(axis, px if pitch is None else pitch, px if nx is None else nx, ny, nz, kstart, kstop, d,
                      _p(u), _p(Lu), ctypes.byref(s))


def mb3_rk(q, qo, Lq, step, dt, nx: int | None = None) -> None:
    nz, ny, px = q.shape
    lib().orc_mb3_rk(step, px, px if nx is None else nx, ny, nz, dt, _p(q), _p(qo), _p(Lq))


# ---- WENO7 Burgers prototype (Matlab_Prototipes/InviscidBurgersNd) ----

def weno7_face(V, U):
    """(hn, hp) at face i+1/2 from the split fluxes V = f+, U = f- of nodes i-3..i+4."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    U = np.ascontiguousarray(U, dtype=np.float64)
    assert V.size == 8 and U.size == 8
    hn, hp = ctypes.c_double(), ctypes.c_double()
    lib().orc_weno7_face(_p(V), _p(U), ctypes.byref(hn), ctypes.byref(hp))
    return hn.value, hp.value


def weno7_rhs(u, d) -> np.ndarray:
    """dF + dG + dH of LFWENO7FDM3d.m (node grid, ends extrapolated)."""
    dx, dy, dz = (d, d, d) if np.isscalar(d) else d
    res = np.empty_like(u)
    lib().orc_weno7_rhs(*_shape(u), dx, dy, dz, _p(u), _p(res))
    return res


def lsrk_stage(u, res, L, s, dt) -> None:
    lib().orc_lsrk_stage(u.size, s, dt, _p(u), _p(res), _p(L))


def weno7_run(u, res, d, cfl, t, tend, nsteps=-1):
    """Advance u (and the carried residual res) in place; returns (steps, t)."""
    dx, dy, dz = (d, d, d) if np.isscalar(d) else d
    tt = ctypes.c_double(t)
    n = lib().orc_weno7_run(*_shape(u), dx, dy, dz, cfl, ctypes.byref(tt), tend, nsteps, _p(u), _p(res))
    return n, tt.value


def diff13_b5(u, Lu, kx, ky, kz) -> None:
    lib().orc_diff13_b5(*_shape(u), kx, ky, kz, _p(u), _p(Lu))


def ssprk_b5(u, uo, Lu, step, dt) -> None:
    lib().orc_ssprk_b5(*_shape(u), step, dt, _p(u), _p(uo), _p(Lu))


def lu3d_13(u, Lu, kx, ky, kz) -> None:
    lib().orc_lu3d_13(*_shape(u), kx, ky, kz, _p(u), _p(Lu))


def lu3d_o2(u, Lu, kx, ky, kz) -> None:
    lib().orc_lu3d_o2(*_shape(u), kx, ky, kz, _p(u), _p(Lu))


def rk3d(u, uo, Lu, step, dt) -> None:
    nx, ny, nz = _shape(u)
    lib().orc_rk3d(step, nx, ny, nz, dt, _p(u), _p(uo), _p(Lu))


# SingleGPU/Burgers3d_WENO5_SharedMem: WENO5Zreconstruction (p = 1,
# eps = 1e-6, Burgers.h:26), local LF, the *(1/6 * (1/sum)) rounding.
B5S = dict(weno=WENO_Z_B5S, p=1, eps=1e-6, lf=LF_LOCAL)


def lu3d_blk(u, Lu, kx, ky, kz) -> None:
    """Call_Lu3d of Diffusion3d_Blocking (Lu = on nodes 3..n-4, planes 3..nz-3)."""
    lib().orc_lu3d_blk(*_shape(u), kx, ky, kz, _p(u), _p(Lu))


def adv_b5s(axis: int, u, Lu, d, **kw) -> None:
    """Call_Adv_{x,y,z} of Burgers3d_WENO5_SharedMem over every node (in place on Lu)."""
    s = scheme(**(kw or B5S))
    lib().orc_adv_b5s(axis, *_shape(u), d, _p(u), _p(Lu), ctypes.byref(s))


def md3_lap(order: int, u, Lu, kstart, kstop, dxc, dyc, dzc, nx: int | None = None) -> None:
    """LaplaceO2_async (order 2) / LaplaceO4_async (order 4) of the MPI heat
    code on a (nz, ny, pitch) slab whose first nx columns are the grid."""
    nz, ny, px = u.shape
    lib().orc_md3_lap(order, px, px if nx is None else nx, ny, nz, kstart, kstop, dxc, dyc, dzc, _p(u), _p(Lu))


def md3_rk(q, qo, Lq, step, dt, nx: int | None = None) -> None:
    nz, ny, px = q.shape
    lib().orc_md3_rk(step, px, px if nx is None else nx, ny, nz, dt, _p(q), _p(qo), _p(Lq))


def maxabs(u) -> float:
    return lib().orc_maxabs(u.size, _p(u))


def rhs_periodic(u, d, **kw) -> np.ndarray:
    dx, dy, dz = (d, d, d) if np.isscalar(d) else d
    Lu = np.empty_like(u)
    s = scheme(**kw)
    lib().orc_rhs_periodic(*_shape(u), dx, dy, dz, _p(u), _p(Lu), ctypes.byref(s))
    return Lu


def rk_stage_all(u, uo, Lu, stage, dt) -> None:
    lib().orc_rk_stage_all(u.size, stage, dt, _p(u), _p(uo), _p(Lu))


def burgers_step(u, d, cfl, t, tend, native: bool = False, **kw) -> float:
    """One SSP-RK3 step in place; returns dt.  native: through native_lib()."""
    dx, dy, dz = (d, d, d) if np.isscalar(d) else d
    work = np.empty(2 * u.size)
    s = scheme(**kw)
    L = native_lib() if native else lib()
    return L.orc_burgers_step(*_shape(u), dx, dy, dz, cfl, t, tend, ctypes.byref(s), _p(u), _p(work))


def burgers_run(u, d, cfl, t, tend, nsteps=-1, **kw):
    """Advance u in place; returns (steps, t)."""
    dx, dy, dz = (d, d, d) if np.isscalar(d) else d
    tt = ctypes.c_double(t)
    s = scheme(**kw)
    n = lib().orc_burgers_run(*_shape(u), dx, dy, dz, cfl, ctypes.byref(tt), tend, nsteps, ctypes.byref(s), _p(u))
    return n, tt.value


def heat7_fe(u, kx, ky, kz, dt, nsteps) -> None:
    lib().orc_heat7_fe(*_shape(u), kx, ky, kz, dt, nsteps, _p(u))


def rhs_1d_periodic(u, dx, **kw) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    Lu = np.empty_like(u)
    s = scheme(**kw)
    lib().orc_rhs_1d_periodic(u.size, dx, _p(u), _p(Lu), ctypes.byref(s))
    return Lu


def side_grad(v, plus=True, **kw):
    """(h, dh/dv) of one reconstruction side."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    g = np.empty(5)
    sc = scheme(**kw)
    h = lib().orc_side_grad(_p(v), 1 if plus else 0, ctypes.byref(sc), _p(g))
    return h, g


def rhs_vjp(u, lam, d, alpha, **kw):
    """((dL/du)^T lam, <lam, L>, <lam, dL/dalpha>) for the periodic global-LF RHS."""
    dx, dy, dz = (d, d, d) if np.isscalar(d) else d
    ub = np.zeros_like(u)
    L, La = ctypes.c_double(), ctypes.c_double()
    sc = scheme(**kw, alpha=alpha)
    sc.lf = LF_GLOBAL
    lib().orc_rhs_vjp(*_shape(u), dx, dy, dz, _p(u), _p(np.ascontiguousarray(lam)), _p(ub), ctypes.byref(sc),
                      ctypes.byref(L), ctypes.byref(La))
    return ub, L.value, La.value


def burgers_adjoint(u0, target, d, cfl, t0, tend, **kw):
    """(J, dJ/du0, steps) of J = 1/2 |u(tend) - target|^2."""
    dx, dy, dz = (d, d, d) if np.isscalar(d) else d
    g = np.empty_like(u0)
    n = ctypes.c_int()
    sc = scheme(**kw)
    J = lib().orc_burgers_adjoint(*_shape(u0), dx, dy, dz, cfl, t0, tend, ctypes.byref(sc), _p(u0),
                                  _p(np.ascontiguousarray(target)), _p(g), ctypes.byref(n))
    return J, g, n.value
