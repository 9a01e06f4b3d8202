"""Loader for the HIP engine's shared library (galpy_amd/libgalpy_hip.so).

Mirrors galpy's ctypes loader (galpy/util/_load_extension_libs.py:64-111):
the library is looked up next to the package; there is no fallback -- if the
library is missing or cannot be loaded every integration call raises.
"""

import ctypes
import os

import numpy
from numpy.ctypeslib import ndpointer

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgalpy_hip.so")

_lib = None
_load_error = None

_F = ("C_CONTIGUOUS", "WRITEABLE")

# name -> (restype, argtypes); the C declarations are in include/galpy_hip.h
SIGNATURES = {
    "integrateFullOrbit": (
        None,
        [
            ctypes.c_int,
            ndpointer(dtype=numpy.float64, flags=_F),
            ctypes.c_int,
            ndpointer(dtype=numpy.float64, flags=_F),
            ctypes.c_int,
            ctypes.c_int,
            ndpointer(dtype=numpy.int32, flags=_F),
            ndpointer(dtype=numpy.float64, flags=_F),
            ctypes.c_void_p,
            ctypes.c_double,
            ctypes.c_double,
            ctypes.c_double,
            ndpointer(dtype=numpy.float64, flags=_F),
            ndpointer(dtype=numpy.int32, flags=_F),
            ctypes.c_int,
            ctypes.c_void_p,
        ],
    ),
    "integrateFullOrbit_hip": (
        ctypes.c_int,
        [
            ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
        ],
    ),
    "gh_potential_upload": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)],
    ),
    "gh_planar_potential_upload": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)],
    ),
    "gh_potential_upload_tfuncs": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
         ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
    ),
    "gh_host_tfunc_eval": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "gh_potential_free": (ctypes.c_int, [ctypes.c_void_p]),
    "gh_potential_device_ptr": (ctypes.c_void_p, [ctypes.c_void_p]),
    "gh_energy_drift_hip": (
        ctypes.c_int,
        [
            ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ],
    ),
    "gh_eval_hip": (
        ctypes.c_int,
        [ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_void_p] * 5 + [ctypes.c_void_p],
    ),
    "gh_host_log_ge1": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "gh_host_sincos": (ctypes.c_int, [ctypes.c_int] + [ctypes.c_void_p] * 3),
    "galpy_hip_last_error": (ctypes.c_char_p, []),
    "galpy_hip_set_interrupt": (None, [ctypes.c_int]),
}


_DXDV_ONE = [
    ndpointer(dtype=numpy.float64, flags=_F),
    ctypes.c_int,
    ndpointer(dtype=numpy.float64, flags=_F),
    ctypes.c_int,
    ndpointer(dtype=numpy.int32, flags=_F),
    ndpointer(dtype=numpy.float64, flags=_F),
    ctypes.c_void_p,
    ctypes.c_double,
    ctypes.c_double,
    ctypes.c_double,
    ndpointer(dtype=numpy.float64, flags=_F),
    ndpointer(dtype=numpy.int32, flags=_F),
    ctypes.c_int,
]
SIGNATURES["integrateFullOrbit_dxdv"] = (None, _DXDV_ONE)
SIGNATURES["integratePlanarOrbit_dxdv"] = (None, _DXDV_ONE + [ctypes.c_void_p])
SIGNATURES["gh_integrate_dxdv"] = (
    None,
    [
        ctypes.c_int, ctypes.c_int, ndpointer(dtype=numpy.float64, flags=_F), ctypes.c_int,
        ndpointer(dtype=numpy.float64, flags=_F), ctypes.c_int, ctypes.c_int,
        ndpointer(dtype=numpy.int32, flags=_F), ndpointer(dtype=numpy.float64, flags=_F),
        ctypes.c_double, ctypes.c_double, ctypes.c_double,
        ndpointer(dtype=numpy.float64, flags=_F), ndpointer(dtype=numpy.int32, flags=_F),
        ctypes.c_int,
    ],
)
SIGNATURES["gh_integrate_dxdv_hip"] = (
    ctypes.c_int,
    [
        ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
    ],
)

# the planar entry points take the same arguments as the full ones
SIGNATURES["integratePlanarOrbit"] = SIGNATURES["integrateFullOrbit"]
SIGNATURES["integratePlanarOrbit_hip"] = SIGNATURES["integrateFullOrbit_hip"]
SIGNATURES["integrateFullOrbit_hip_noalloc"] = SIGNATURES["integrateFullOrbit_hip"]
SIGNATURES["gh_planar_energy_drift_hip"] = SIGNATURES["gh_energy_drift_hip"]
SIGNATURES["gh_orbit_summary_hip"] = (
    ctypes.c_int,
    [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
     ctypes.c_void_p, ctypes.c_double] + [ctypes.c_void_p] * 7,
)


def load(path=None):
    """Load (once) and return the ctypes handle; raises if unavailable."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"HIP engine library not built: {path} is missing (run `python -m galpy_amd.build`)"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().galpy_hip_last_error().decode()


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {last_error()}")
