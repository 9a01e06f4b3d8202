"""The oracle's spherical components that need special functions --
EinastoPotential (41), TwoPowerSphericalPotential (42), ExpTruncNFWPotential
(46) -- and those functions, pinned by what the mathematics fixes:

* e^x E1(x) (replacing GSL's gsl_sf_expint_E1) against mpmath.e1;
* 2F1(A, B; A+1; -x) (galpy calls the vendored xsf library's hyp2f1)
  against mpmath.hyp2f1 over alpha, beta and x in [1e-6, 1e4], including the
  integer 3 - beta cases where the large-x expansion has a logarithm, and
  against the reference's own hyp2f1 compiled from galpy/util/wrap_xsf.cpp
  into oracle/_ref when that is built;
* each component's radial force against the enclosed mass of galpy's
  documented density integrated by mpmath, and its potential against
  -M(r)/r - 4 pi int_r^inf rho u du (not for TwoPowerSpherical, whose Eval
  calls 2F1 outside that family and is not restated);
* Phi''(r) against Poisson's equation, 4 pi rho - 2 Phi'/r.
Einasto's force is restated with gamma(3n, x) where galpy forms
-Gamma(3n) (Q(3n, x) - 1), which loses every digit for x << 3n; the pin
holds at r = 1e-3 h where galpy's form is off by 1e-3."""

import ctypes
import os

import mpmath
import numpy
import pytest

from helpers import oracle, oracle_forces, oracle_hessian, oracle_potential

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle", "_ref", "libref_hyp2f1.so")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


X = numpy.array([1e-6, 0.01, 0.5, 1.0, 2.9, 3.0, 3.1, 10.0, 100.0, 1e4])
ALPHA_BETA = [(1.3, 3.6), (0.5, 3.0), (1.0, 3.0), (1.0, 4.0), (2.0, 4.0), (0.0, 4.0), (1.5, 2.5),
              (1.2, 7.5), (0.3, 2.0), (2.5, 3.0), (0.7, 5.0)]


def test_e1_scaled_against_mpmath():
    out = numpy.empty_like(X)
    oracle().oracle_e1_scaled(len(X), _p(X), _p(out))
    with mpmath.workdps(30):
        want = numpy.array([float(mpmath.e1(mpmath.mpf(v)) * mpmath.exp(mpmath.mpf(v))) for v in X])
    assert numpy.all(numpy.abs(out / want - 1.0) < 4e-15)


@pytest.mark.parametrize("alpha,beta", ALPHA_BETA)
def test_hyp2f1_family_against_mpmath_and_reference(alpha, beta):
    A, B = 3.0 - alpha, beta - alpha
    out = numpy.empty_like(X)
    oracle().oracle_hyp2f1_ap1(len(X), ctypes.c_double(A), ctypes.c_double(B), ctypes.c_double(3.0 - beta),
                               _p(X), _p(out))
    with mpmath.workdps(30):
        want = numpy.array([float(mpmath.hyp2f1(A, B, A + 1, -mpmath.mpf(v))) for v in X])
    assert numpy.all(numpy.abs(out / want - 1.0) < 4e-15), (alpha, beta)
    if os.path.exists(REF):
        ref = ctypes.CDLL(REF)
        ref.hyp2f1.restype = ctypes.c_double
        ref.hyp2f1.argtypes = [ctypes.c_double] * 4
        got = numpy.array([ref.hyp2f1(A, B, A + 1, -v) for v in X])
        assert numpy.all(numpy.abs(out / got - 1.0) < 4e-15), (alpha, beta)


def _cases():
    pi = mpmath.pi
    amp = 1.3
    return {
        "Einasto n=2.5": ([41], [amp, 1.7, 2.5], lambda u: amp * mpmath.exp(-(u / mpmath.mpf(1.7)) ** (1 / mpmath.mpf(2.5)))),
        "Einasto n=0.5": ([41], [amp, 0.9, 0.5], lambda u: amp * mpmath.exp(-(u / mpmath.mpf(0.9)) ** 2)),
        "ExpTruncNFW": ([46], [amp, 1.2, 3.0],
                        lambda u: amp * mpmath.exp(-u / 3) / (4 * pi * mpmath.mpf(1.2) ** 2 * u * (1 + u / mpmath.mpf(1.2)) ** 2)),
        "TwoPowerSpherical": ([42], [amp, 1.2, 1.3, 3.6],
                              lambda u: amp / (4 * pi * mpmath.mpf(1.2) ** 3) / ((u / mpmath.mpf(1.2)) ** mpmath.mpf(1.3)
                                                                             * (1 + u / mpmath.mpf(1.2)) ** mpmath.mpf(2.3))),
        "TwoPowerSpherical beta=3": ([42], [amp, 2.0, 0.5, 3.0],
                                     lambda u: amp / (4 * pi * 8) / ((u / 2) ** mpmath.mpf(0.5) * (1 + u / 2) ** mpmath.mpf(2.5))),
    }


R = numpy.array([1e-3, 0.2, 1.0, 3.0, 12.0, 40.0])
Z = numpy.array([5e-4, 0.1, 0.5, -1.0, 4.0, -9.0])


@pytest.mark.parametrize("name", list(_cases()))
def test_force_and_potential_against_enclosed_mass(name):
    pt, pa, rho = _cases()[name]
    pt = numpy.array(pt, dtype=numpy.int32)
    pa = numpy.array(pa, dtype=float)
    FR, Fz, Fphi = oracle_forces(pt, pa, R, Z)
    pot = oracle_potential(pt, pa, R, Z)
    assert numpy.all(Fphi == 0.0)
    with mpmath.workdps(30):
        for k in range(len(R)):
            r = mpmath.sqrt(mpmath.mpf(R[k]) ** 2 + mpmath.mpf(Z[k]) ** 2)
            M = 4 * mpmath.pi * mpmath.quad(lambda u: rho(u) * u**2, [0, r / 2, r])
            g = float(-M / r**3)
            # the ExpTruncNFW series below 1e-3 min(a, rc) is galpy's, good to ~(r/a)^2
            tol = 2e-12 if (name == "ExpTruncNFW" and k == 0) else 1e-13
            assert abs(FR[k] - g * R[k]) <= tol * abs(g * R[k]), (name, k)
            assert abs(Fz[k] - g * Z[k]) <= tol * abs(g * R[k]), (name, k)
            if name.startswith("TwoPowerSpherical"):
                assert numpy.isnan(pot[k])
                continue
            phi = -M / r - 4 * mpmath.pi * mpmath.quad(lambda u: rho(u) * u, [r, 2 * r, 10 * r, mpmath.inf])
            assert abs(pot[k] - float(phi)) <= 1e-13 * abs(float(phi)), (name, k)


@pytest.mark.parametrize("name", list(_cases()))
def test_second_derivative_is_poisson(name):
    pt, pa, rho = _cases()[name]
    pt = numpy.array(pt, dtype=numpy.int32)
    pa = numpy.array(pa, dtype=float)
    Rp = numpy.array([0.05, 0.7, 2.0, 9.0])
    zp = numpy.zeros_like(Rp)
    R2, z2, Rz = oracle_hessian(pt, pa, Rp, zp)  # in the plane: R2 = Phi'', z2 = Phi'/r
    with mpmath.workdps(30):
        for k in range(len(Rp)):
            want = float(4 * mpmath.pi * rho(mpmath.mpf(Rp[k]))) - 2.0 * z2[k]
            assert abs(R2[k] - want) <= 1e-12 * max(abs(want), abs(z2[k])), (name, k)
