"""The C-ABI library loads without a GPU and exports every function that
include/galpy_hip.h declares (no compute calls here)."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            src = open(os.path.join(inc, fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            src = re.sub(r"^\s*#.*$", "", src, flags=re.M)  # preprocessor lines
            names += re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", src)
    return sorted(set(names))


def test_library_exports_declared_symbols():
    from galpy_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert "integrateFullOrbit" in declared and "integrateFullOrbit_hip" in declared
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing


def test_loader_signatures_cover_declarations():
    from galpy_amd import _lib

    lib = _lib.load()
    for name in _lib.SIGNATURES:
        assert hasattr(lib, name)


def test_error_reporting_without_gpu_work():
    from galpy_amd import _lib

    lib = _lib.load()
    # decoding failures are host-side and need no device
    fn = lib.gh_host_bulge_lower_gamma
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                   ctypes.c_void_p, ctypes.c_void_p]
    import numpy

    pt = numpy.array([999], dtype=numpy.int32)  # no galpy potential has this code
    pa = numpy.zeros(4)
    x = numpy.ones(1)
    out = numpy.zeros(1)
    assert fn(1, pt.ctypes.data, pa.ctypes.data, 0, 1, x.ctypes.data, out.ctypes.data) != 0
    assert "pot_type 999" in _lib.last_error()
