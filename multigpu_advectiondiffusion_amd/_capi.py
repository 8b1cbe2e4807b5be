"""ctypes binding of libmgad.so (the C ABI declared in include/mgad.h).

The library is the only compute path: if libmgad.so is missing this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmgad.so")

MGAD_OK, MGAD_ERR_INVALID_ARG, MGAD_ERR_UNSUPPORTED, MGAD_ERR_HIP, MGAD_ERR_NONFINITE = 0, 1, 2, 3, 4
MGAD_ERR_SCRATCH_FULL = 5
BATCH_SCRATCH_BYTES = 1 << 20  # MGAD_BATCH_SCRATCH_BYTES
WENO_Z, WENO_JS = 0, 1
LF_LOCAL, LF_GLOBAL = 0, 1
CTRL_BYTES = 256
F64, F32 = 0, 1
BATCH_SPLIT, BATCH_SPLIT_WIDE = -1, -2  # kchunk of the split batch work lists (16x16 / 32x16 tiles)
BANDS_SHRINK = -1  # kchunk of a single-problem stage launch with z bands that shrink towards the top

# Every symbol include/mgad.h declares.
EXPORTS = (
    "mgad_get_last_error", "mgad_error_string", "mgad_version",
    "CopyToConstantMemory", "Call_Adv_x", "Call_Adv_y", "Call_Adv_z", "Call_Diff_",
    "Call_sspRK", "Call_Lu3d", "Call_RK3d",
    "Call_Adv_x_pitched", "Call_Adv_y_pitched", "Call_Adv_z_pitched", "Call_Diff_pitched", "Call_sspRK_pitched",
    "Call_Lu3d_pitched", "Call_RK3d_pitched",
    "Call_Adv_x_sharedmem", "Call_Adv_y_sharedmem", "Call_Adv_z_sharedmem", "Call_Lu3d_blocking",
    "Call_Diff_mpi", "Call_Diff_O2_mpi", "Call_sspRK_mpi",
    "mgad_set_adv_scheme", "mgad_lu3d_o2",
    "mgad_burgers_init", "mgad_burgers_stage", "mgad_burgers_steps", "mgad_burgers_read_ctrl",
    "mgad_burgers_init_p", "mgad_burgers_steps_p", "mgad_burgers_autotune", "mgad_burgers_set_launch",
    "mgad_batch_workspace_bytes", "mgad_batch_init", "mgad_batch_steps", "mgad_batch_steps_serial",
    "mgad_batch_init_static",
    "mgad_adjoint_workspace_bytes", "mgad_burgers_adjoint",
    "mgad_burgers_ref_steps", "mgad_heat7_fe_steps", "mgad_heat13_rk3_steps", "mgad_check_division", "mgad_ic_fourier", "mgad_ic_gauss_noise",
    "mgad_save_binary3d", "mgad_load_binary3d", "mgad_init3d_burgers", "mgad_init3d_heat", "mgad_heat_error_norms",
    "mgad_init_subdomain", "mgad_merge_domains",
    "mgad_slab_exchange", "mgad_burgers_slab_init", "mgad_burgers_slab_steps",
    "Compute_Adv_x_Async", "Compute_Adv_y_Async", "Compute_Adv_z_Async", "Compute_sspRK3Async",
    "CopyBoundaryRegionToGhostCellAsync", "CopyGhostCellToBoundaryRegionAsync",
    "mgad_weno7_init", "mgad_weno7_rhs", "mgad_weno7_steps", "mgad_weno7_read_ctrl",
)


class Dim3(ctypes.Structure):
    """HIP dim3: three 32-bit unsigned ints, passed by value."""
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32), ("z", ctypes.c_uint32)]


class Scheme(ctypes.Structure):
    """mgad_scheme."""
    _fields_ = [("weno", ctypes.c_int), ("p", ctypes.c_int), ("eps", ctypes.c_double)]


class BurgersState(ctypes.Structure):
    """mgad_burgers_state."""
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("tend", ctypes.c_double),
                ("cfl", ctypes.c_double), ("amax", ctypes.c_double), ("steps", ctypes.c_uint64)]


class Problem(ctypes.Structure):
    """mgad_problem."""
    _fields_ = [("u", ctypes.c_void_p), ("w1", ctypes.c_void_p), ("w2", ctypes.c_void_p), ("ctrl", ctypes.c_void_p),
                ("nx", ctypes.c_uint), ("ny", ctypes.c_uint), ("nz", ctypes.c_uint),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double)]


class Batch(ctypes.Structure):
    """mgad_batch."""
    _fields_ = [("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t), ("nprob", ctypes.c_int),
                ("nitems", ctypes.c_int), ("grid", ctypes.c_int), ("kchunk", ctypes.c_int),
                ("precision", ctypes.c_int), ("scheme", Scheme)]


class MgadError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MgadError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                            "(or python -m multigpu_advectiondiffusion_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        U, D, I, VP = ctypes.c_uint, ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.mgad_get_last_error.restype = I
        L.mgad_error_string.restype = ctypes.c_char_p
        L.mgad_error_string.argtypes = [I]
        L.mgad_version.restype = ctypes.c_char_p
        L.CopyToConstantMemory.argtypes = [D, D, D]
        L.CopyToConstantMemory.restype = None
        for name in ("Call_Adv_x", "Call_Adv_y", "Call_Adv_z"):
            f = getattr(L, name)
            f.argtypes = [Dim3, Dim3, VP, U, U, U, D, VP, VP]
            f.restype = None
        L.Call_Diff_.argtypes = [Dim3, Dim3, VP, U, U, U, VP, VP]
        L.Call_Diff_.restype = None
        L.Call_sspRK.argtypes = [Dim3, Dim3, VP, U, U, U, U, D, VP, VP, VP]
        L.Call_sspRK.restype = None
        L.Call_Lu3d.argtypes = [Dim3, Dim3, VP, U, U, U, VP, VP]
        L.Call_Lu3d.restype = None
        L.Call_RK3d.argtypes = [Dim3, Dim3, VP, U, U, U, U, D, VP, VP, VP]
        L.Call_RK3d.restype = None
        for name in ("Call_Adv_x_pitched", "Call_Adv_y_pitched", "Call_Adv_z_pitched"):
            f = getattr(L, name)
            f.argtypes = [Dim3, Dim3, VP, U, U, U, U, D, VP, VP]
            f.restype = None
        L.Call_Diff_pitched.argtypes = [Dim3, Dim3, VP, U, U, U, U, VP, VP]
        L.Call_Diff_pitched.restype = None
        L.Call_sspRK_pitched.argtypes = [Dim3, Dim3, VP, U, U, U, U, U, D, VP, VP, VP]
        L.Call_sspRK_pitched.restype = None
        L.Call_Lu3d_pitched.argtypes = [Dim3, Dim3, VP, U, U, U, U, VP, VP]
        L.Call_Lu3d_pitched.restype = None
        L.Call_RK3d_pitched.argtypes = [Dim3, Dim3, VP, U, U, U, U, U, D, VP, VP, VP]
        L.Call_RK3d_pitched.restype = None
        for name in ("Call_Adv_x_sharedmem", "Call_Adv_y_sharedmem", "Call_Adv_z_sharedmem"):
            f = getattr(L, name)
            f.argtypes = [Dim3, Dim3, VP, U, U, U, U, D, VP, VP]
            f.restype = None
        L.Call_Lu3d_blocking.argtypes = [Dim3, Dim3, VP, U, U, U, U, VP, VP]
        L.Call_Lu3d_blocking.restype = None
        for name in ("Call_Diff_mpi", "Call_Diff_O2_mpi"):
            f = getattr(L, name)
            f.argtypes = [Dim3, Dim3, VP, U, U, U, U, U, U, D, D, D, VP, VP]
            f.restype = None
        L.Call_sspRK_mpi.argtypes = [Dim3, Dim3, VP, U, U, U, U, U, D, VP, VP, VP]
        L.Call_sspRK_mpi.restype = None
        for name in ("Compute_Adv_x_Async", "Compute_Adv_y_Async", "Compute_Adv_z_Async"):
            f = getattr(L, name)
            f.argtypes = [Dim3, Dim3, VP, U, U, U, U, U, U, D, VP, VP]
            f.restype = None
        L.Compute_sspRK3Async.argtypes = [Dim3, Dim3, VP, U, U, U, U, U, D, VP, VP, VP]
        L.Compute_sspRK3Async.restype = None
        for name in ("CopyBoundaryRegionToGhostCellAsync", "CopyGhostCellToBoundaryRegionAsync"):
            f = getattr(L, name)
            f.argtypes = [Dim3, Dim3, VP, VP, VP, U, U, U, U, U, U]
            f.restype = None
        L.mgad_set_adv_scheme.argtypes = [I, I, D, I, D]
        L.mgad_set_adv_scheme.restype = I
        L.mgad_lu3d_o2.argtypes = [VP, U, U, U, VP, VP]
        L.mgad_lu3d_o2.restype = I
        L.mgad_burgers_init.argtypes = [VP, U, U, U, VP, VP, D, D, D]
        L.mgad_burgers_init.restype = I
        L.mgad_burgers_stage.argtypes = [VP, ctypes.POINTER(Scheme), U, U, U, D, D, D, I, VP, VP, VP, VP]
        L.mgad_burgers_stage.restype = I
        L.mgad_burgers_steps.argtypes = [VP, ctypes.POINTER(Scheme), U, U, U, D, D, D, VP, VP, VP, VP, I, I]
        L.mgad_burgers_steps.restype = I
        L.mgad_burgers_init_p.argtypes = [I, VP, U, U, U, VP, VP, D, D, D]
        L.mgad_burgers_init_p.restype = I
        L.mgad_burgers_steps_p.argtypes = [I, VP, ctypes.POINTER(Scheme), U, U, U, D, D, D, VP, VP, VP, VP, I, I]
        L.mgad_burgers_steps_p.restype = I
        IP = ctypes.POINTER(ctypes.c_int)
        L.mgad_burgers_autotune.argtypes = [VP, I, ctypes.POINTER(Scheme), U, U, U, D, D, D, VP, VP, VP, IP, IP]
        L.mgad_burgers_autotune.restype = I
        L.mgad_burgers_set_launch.argtypes = [I, ctypes.POINTER(Scheme), U, U, U, I, I]
        L.mgad_burgers_set_launch.restype = I
        PP, BP = ctypes.POINTER(Problem), ctypes.POINTER(Batch)
        L.mgad_batch_workspace_bytes.argtypes = [I, PP]
        L.mgad_batch_workspace_bytes.restype = ctypes.c_size_t
        L.mgad_batch_init.argtypes = [VP, I, ctypes.POINTER(Scheme), I, PP, VP, ctypes.c_size_t, I, D, D, D, BP]
        L.mgad_batch_init.restype = I
        L.mgad_batch_init_static.argtypes = [VP, I, ctypes.POINTER(Scheme), I, PP, I, D, D, D, BP]
        L.mgad_batch_init_static.restype = I
        L.mgad_batch_steps.argtypes = [VP, BP, I, I]
        L.mgad_batch_steps.restype = I
        L.mgad_batch_steps_serial.argtypes = [VP, BP, PP, I, I]
        L.mgad_batch_steps_serial.restype = I
        L.mgad_adjoint_workspace_bytes.argtypes = [U, U, U, I, I]
        L.mgad_adjoint_workspace_bytes.restype = ctypes.c_size_t
        L.mgad_burgers_adjoint.argtypes = [VP, ctypes.POINTER(Scheme), U, U, U, D, D, D, D, D, D, VP, VP, VP,
                                           ctypes.POINTER(D), I, I, VP, ctypes.c_size_t, IP]
        L.mgad_burgers_adjoint.restype = I
        L.mgad_burgers_read_ctrl.argtypes = [VP, VP, ctypes.POINTER(BurgersState)]
        L.mgad_burgers_read_ctrl.restype = I
        L.mgad_heat7_fe_steps.argtypes = [VP, U, U, U, D, D, D, D, VP, VP, I, I]
        L.mgad_heat7_fe_steps.restype = I
        L.mgad_burgers_ref_steps.argtypes = [VP, ctypes.POINTER(Scheme), I, U, U, U, D, D, D, D, VP, VP, VP, I, I]
        L.mgad_burgers_ref_steps.restype = I
        L.mgad_heat13_rk3_steps.argtypes = [VP, U, U, U, D, VP, VP, VP, I, I]
        L.mgad_heat13_rk3_steps.restype = I
        L.mgad_slab_exchange.argtypes = [VP, U, U, U, U, VP]
        L.mgad_slab_exchange.restype = I
        L.mgad_burgers_slab_init.argtypes = [VP, U, U, U, U, VP, VP, D, D, D]
        L.mgad_burgers_slab_init.restype = I
        L.mgad_burgers_slab_steps.argtypes = [VP, ctypes.POINTER(Scheme), U, U, U, U, D, D, D, VP, VP, VP, VP, I, I]
        L.mgad_burgers_slab_steps.restype = I
        L.mgad_weno7_init.argtypes = [VP, U, U, U, VP, VP, VP, D, D, D]
        L.mgad_weno7_init.restype = I
        L.mgad_weno7_rhs.argtypes = [VP, U, U, U, D, D, D, VP, VP]
        L.mgad_weno7_rhs.restype = I
        L.mgad_weno7_steps.argtypes = [VP, U, U, U, D, D, D, VP, VP, VP, VP, VP, I, I]
        L.mgad_weno7_steps.restype = I
        L.mgad_weno7_read_ctrl.argtypes = [VP, VP, ctypes.POINTER(D), ctypes.POINTER(D),
                                           ctypes.POINTER(ctypes.c_ulonglong)]
        L.mgad_weno7_read_ctrl.restype = I
        L.mgad_check_division.argtypes = [VP, I, VP, VP, VP, VP, ctypes.c_size_t]
        L.mgad_check_division.restype = I
        L.mgad_save_binary3d.argtypes = [VP, U, U, U, ctypes.c_char_p]
        L.mgad_save_binary3d.restype = I
        L.mgad_load_binary3d.argtypes = [VP, U, U, U, ctypes.c_char_p]
        L.mgad_load_binary3d.restype = I
        for name in ("mgad_init3d_burgers", "mgad_init3d_heat"):
            f = getattr(L, name)
            f.argtypes = [I, VP, D, D, D, U, U, U]
            f.restype = I
        L.mgad_heat_error_norms.argtypes = [VP, D, D, D, D, U, U, U, VP]
        L.mgad_heat_error_norms.restype = I
        for name in ("mgad_init_subdomain", "mgad_merge_domains"):
            f = getattr(L, name)
            f.argtypes = [VP, VP, U, U, U, U]
            f.restype = I
        L.mgad_ic_fourier.argtypes = [VP, U, U, U, D, D, ctypes.c_uint64, I, I, D, D]
        L.mgad_ic_fourier.restype = I
        L.mgad_ic_gauss_noise.argtypes = [VP, U, U, U, D, D, ctypes.c_uint64]
        L.mgad_ic_gauss_noise.restype = I
        _lib = L
    return _lib


def error_string(code: int) -> str:
    return lib().mgad_error_string(code).decode()


def check(rc: int, what: str) -> None:
    """Raise on a failed status; the failure is delivered, so the library's
    sticky copy of it is read and cleared too."""
    if rc != MGAD_OK:
        lib().mgad_get_last_error()
        raise MgadError(f"{what}: {error_string(rc)} (code {rc})")


def last_error() -> int:
    return lib().mgad_get_last_error()


def raise_last(what: str) -> None:
    """Raise if a void drop-in entry point recorded an error."""
    rc = last_error()
    if rc != MGAD_OK:
        raise MgadError(f"{what}: {error_string(rc)} (code {rc})")
