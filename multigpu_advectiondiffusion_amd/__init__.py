"""MI355X-native WENO5 Burgers / 7-point heat stencil library (libmgad.so).

The compute path is the HIP library built by build.py; this package is its
host side: ctypes bindings (_capi), HIP plumbing (hip), the reference's
device-wrapper interface (reference_api) and solver drivers (solvers).
"""
from . import _capi, hip, reference_api, solvers  # noqa: F401
from ._capi import MgadError  # noqa: F401

__version__ = "0.1.0"
