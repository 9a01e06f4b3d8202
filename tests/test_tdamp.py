"""TimeDependentAmplitudeWrapperPotential (-9): the wrapped potential times
amp A(t), A a Python function that galpy calls from C through pot_tfuncs.
The oracle calls it the same way; pinned by the wrapper's closed form (the
forces of the wrapped potential scaled by amp A(t), which does not go
through the wrapper's restatement) and against the host class.  The GPU
side evaluates A from a piecewise Chebyshev table that the library builds
on the host from the same callback: the table is checked here, on the CPU,
through the host-side evaluation entry point."""

import ctypes

import numpy
import pytest

from helpers import oracle, oracle_forces, oracle_integrate, oracle_potential, oracle_tfuncs


def A(t):
    return 1.0 + 0.3 * numpy.sin(0.7 * t) + 0.05 * t


def _pot(amp=1.2):
    from galpy_amd import potential as P

    return [P.TimeDependentAmplitudeWrapperPotential(amp=amp, A=A, pot=P.MWPotential2014_factory())]


def test_packing_and_closed_form():
    from galpy_amd import potential as P

    pot = _pot()
    _, pt, pa, tf = P.parse_pot(pot)
    assert list(pt) == [-9, 15, 5, 9] and pa[0] == 3 and pa[-1] == 1.2 and tf == [A]
    _, mpt, mpa, _ = P.parse_pot(P.MWPotential2014_factory())
    rng = numpy.random.default_rng(5)
    n = 32
    R, z = rng.uniform(0.2, 2.5, n), rng.uniform(-1.0, 1.0, n)
    phi, t = rng.uniform(0.0, 2.0 * numpy.pi, n), rng.uniform(-5.0, 20.0, n)
    with oracle_tfuncs(tf):
        FR, Fz, Fp = oracle_forces(pt, pa, R, z, phi, t)
        V = oracle_potential(pt, pa, R, z, phi, t)
    wR, wz, _ = oracle_forces(mpt, mpa, R, z, phi, t)
    f = 1.2 * A(t)
    assert numpy.allclose(FR, f * wR, rtol=1e-15, atol=0) and numpy.allclose(Fz, f * wz, rtol=1e-15, atol=0)
    assert numpy.allclose(V, f * oracle_potential(mpt, mpa, R, z, phi, t), rtol=1e-15, atol=0)
    w = pot[0]
    assert numpy.allclose(w.Rforce(R, z, phi, t), FR, rtol=1e-13) and numpy.allclose(w(R, z, phi, t), V, rtol=1e-13)
    # without its time function the parse fails
    yo = numpy.array([[1.0, 0.1, 1.0, 0.0, 0.1, 0.0]])
    _, err, _, _ = oracle_integrate(pt, pa, yo, numpy.linspace(0, 1, 3), 6)
    assert err[0] != 0


def test_chebyshev_table_matches_the_function():
    """The library's host-side table of A (the one the device reads)
    reproduces A to rounding over the tabulated range."""
    from galpy_amd import _lib
    from galpy_amd.integrate import tfunc_array

    lib = _lib.load()
    fn = lib.gh_host_tfunc_eval
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                   ctypes.c_void_p]
    arr = tfunc_array([A, lambda t: numpy.exp(-0.5 * (t - 3.0) ** 2)])
    ts = numpy.linspace(-1.0, 101.0, 20001)
    out = numpy.empty_like(ts)
    for k, f in enumerate((A, lambda t: numpy.exp(-0.5 * (t - 3.0) ** 2))):
        assert fn(ctypes.cast(ctypes.byref(arr, k * ctypes.sizeof(ctypes.c_void_p)), ctypes.c_void_p), -1.0, 101.0,
                  len(ts), ts.ctypes.data, out.ctypes.data) == 0
        want = numpy.array([f(x) for x in ts])
        assert numpy.abs(out - want).max() <= 4e-15 * numpy.abs(want).max()
