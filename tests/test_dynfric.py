"""ChandrasekharDynamicalFrictionForce (-7; ChandrasekharDynamicalFrictionForce.c)
and the densities it needs.  Pinned by: each supported density equals
-div F / 4 pi of its potential's forces (Poisson); the force equals
Chandrasekhar's formula written with scipy's erf; a satellite on a circular
orbit in the singular isothermal sphere spirals in as r^2 = r0^2 - 2 GM
lnLambda [erf 1 - 2 e^-1 / sqrt(pi)] t (Binney & Tremaine eq. 8.13: vc = 1,
sigma = 1/sqrt 2), integrated by the oracle's DOP853."""

import numpy
import pytest
from scipy import special

from helpers import oracle_density, oracle_diss_forces, oracle_forces, oracle_integrate


def _densities():
    from galpy_amd import potential as P

    return {"hernquist": P.HernquistPotential(amp=1.3, a=0.8), "nfw": P.NFWPotential(amp=2.0, a=1.5),
            "plummer": P.PlummerPotential(amp=0.7, b=0.4), "mn": P.MiyamotoNagaiPotential(amp=1.1, a=0.5, b=0.2),
            "mn0": P.MiyamotoNagaiPotential(amp=1.1, a=0.0, b=0.3),
            "psc": P.PowerSphericalPotentialwCutoff(amp=1.0, alpha=1.8, rc=1.9),
            "loghalo": P.LogarithmicHaloPotential(amp=1.0, q=0.8, core=0.1),
            "logtriax": P.LogarithmicHaloPotential(amp=1.0, q=0.8, b=0.85, core=0.1),
            "jaffe": P.JaffePotential(amp=1.2, a=0.9), "isochrone": P.IsochronePotential(amp=1.0, b=0.6),
            "powerlaw": P.PowerSphericalPotential(amp=1.0, alpha=2.2), "burkert": P.BurkertPotential(amp=1.0, a=1.2)}


@pytest.mark.parametrize("name", list(_densities()))
def test_density_is_poisson(name):
    from galpy_amd import potential as P

    pot = _densities()[name]
    _, pt, pa, _ = P.parse_pot([pot])
    rng = numpy.random.default_rng(70)
    R, z, phi = rng.uniform(0.3, 2.5, 20), rng.uniform(-1.0, 1.0, 20), rng.uniform(0, 2 * numpy.pi, 20)
    h = 1e-4
    F = lambda RR, zz, pp: oracle_forces(pt, pa, RR, zz, pp)
    lap = -((F(R + h, z, phi)[0] - F(R - h, z, phi)[0]) / (2 * h) + F(R, z, phi)[0] / R
            + (F(R, z + h, phi)[1] - F(R, z - h, phi)[1]) / (2 * h)
            + (F(R, z, phi + h)[2] - F(R, z, phi - h)[2]) / (2 * h) / R**2)
    assert numpy.allclose(oracle_density(pt, pa, R, z, phi), lap / (4 * numpy.pi), rtol=1e-6, atol=1e-9)


def test_force_is_chandrasekhars_formula():
    from galpy_amd import potential as P

    mw = P.MWPotential2014_factory()
    sig = lambda r: 0.5 + 0.1 * numpy.exp(-r)
    f = P.ChandrasekharDynamicalFrictionForce(GMs=0.02, rhm=0.05, dens=mw, sigmar=sig, gamma=1.2)
    npot, pt, pa, _ = P.parse_pot([f])
    assert list(pt) == [-7, 15, 5, 9] and npot == 1
    _, mt, ma, _ = P.parse_pot(mw)
    rng = numpy.random.default_rng(71)
    n = 30
    R, z, phi = rng.uniform(0.3, 3.0, n), rng.uniform(-1.0, 1.0, n), rng.uniform(0, 6.28, n)
    vR, vT, vz = rng.normal(0, 0.3, n), rng.uniform(0.5, 1.2, n), rng.normal(0, 0.3, n)
    FR, Fz, Fp = oracle_diss_forces(pt, pa, R, z, phi, 0.0, vR, vT, vz)
    r2, v2 = R**2 + z**2, vR**2 + vT**2 + vz**2
    v = numpy.sqrt(v2)
    GMvs = 0.02 / v2
    D = numpy.where(GMvs < 0.05, 0.05, GMvs)
    lnL = 0.5 * numpy.log(1 + r2 / 1.2**2 / D**2)
    X = v / numpy.sqrt(2) / sig(numpy.sqrt(r2))
    Xf = special.erf(X) - 2 / numpy.sqrt(numpy.pi) * X * numpy.exp(-X * X)
    A = -4 * numpy.pi * 0.02 * Xf * lnL * oracle_density(mt, ma, R, z, phi) / v2 / v
    # sigma_r is a natural spline through 501 samples: agreement to its accuracy
    assert numpy.allclose(FR, A * vR, rtol=1e-7) and numpy.allclose(Fz, A * vz, rtol=1e-7)
    assert numpy.allclose(Fp, A * vT * R, rtol=1e-7)


def test_circular_orbit_decay_in_the_isothermal_sphere():
    from galpy_amd import potential as P

    sis = P.LogarithmicHaloPotential(normalize=1.0, q=1.0)
    GMs, lnL = 0.005, 3.0
    f = P.ChandrasekharDynamicalFrictionForce(GMs=GMs, const_lnLambda=lnL)  # the isothermal default
    _, pt, pa, _ = P.parse_pot([sis, f])
    t = numpy.linspace(0.0, 20.0, 201)
    r0 = 1.5
    res, err, *_ = oracle_integrate(pt, pa, numpy.array([[r0, 0.0, 1.0, 0.0, 0.0, 0.0]]), t, 6)
    assert err[0] == 0
    r2 = res[0, :, 0] ** 2 + res[0, :, 3] ** 2
    Xf1 = special.erf(1.0) - 2 / numpy.sqrt(numpy.pi) * numpy.exp(-1.0)
    # the mean decay rate of r^2 (the orbit, started circular, picks up a
    # small epicyclic wobble as friction slows it)
    slope = numpy.polyfit(t, r2, 1)[0]
    assert abs(slope / (-2 * GMs * lnL * Xf1) - 1.0) < 0.05


def test_cos_integral_and_fdm_factor():
    """The oracle's Ci against scipy's sici; the FDM force is the
    Chandrasekhar force times FDMfactor/CDMfactor (below one), written out
    with scipy in its three regimes."""
    import ctypes

    from helpers import _p, oracle

    from galpy_amd import potential as P

    x = numpy.concatenate([numpy.geomspace(1e-6, 4.0, 100), numpy.geomspace(4.0, 1e4, 100)])
    ci = numpy.empty_like(x)
    oracle().oracle_cos_integral(x.size, _p(x), _p(ci))
    assert numpy.abs(ci - special.sici(x)[1]).max() < 2e-15
    sig = lambda r: 0.5 + 0.0 * r
    kw = dict(GMs=0.02, rhm=0.05, sigmar=sig)
    ch = P.ChandrasekharDynamicalFrictionForce(**kw)
    rng = numpy.random.default_rng(73)
    n = 60
    R, z, phi = rng.uniform(0.3, 3.0, n), rng.uniform(-1.0, 1.0, n), rng.uniform(0, 6.28, n)
    vR, vT, vz = rng.normal(0, 0.3, n), rng.uniform(0.5, 1.2, n), rng.normal(0, 0.3, n)
    _, cpt, cpa, _ = P.parse_pot([ch])
    cFR = oracle_diss_forces(cpt, cpa, R, z, phi, 0.0, vR, vT, vz)[0]
    for mh in (0.2, 2.0, 40.0):  # zero-dispersion, blended and dispersion regimes
        f = P.FDMDynamicalFrictionForce(mhbar=mh, **kw)
        npot, pt, pa, _ = P.parse_pot([f])
        assert list(pt[:2]) == [-11, -7] and npot == 1
        FR = oracle_diss_forces(pt, pa, R, z, phi, 0.0, vR, vT, vz)[0]
        r, v = numpy.hypot(R, z), numpy.sqrt(vR**2 + vT**2 + vz**2)
        D = numpy.where(0.02 / v**2 < 0.05, 0.05, 0.02 / v**2)
        lnL = 0.5 * numpy.log(1 + r**2 / D**2)
        X = v / numpy.sqrt(2) / 0.5
        Xf = special.erf(X) - 2 / numpy.sqrt(numpy.pi) * X * numpy.exp(-X * X)
        kr, M = 2 * mh * v * r, v / 0.5
        zero = lambda k: numpy.euler_gamma + numpy.log(k) - special.sici(k)[1] + numpy.sin(k) / k - 1
        mu = (2 * M - kr / 2) / (2 * M - M / 2)
        fdm = numpy.where(kr < M, zero(kr), numpy.where(kr > 4 * M, numpy.log(kr / M) * Xf,
                                                         mu * zero(M) + (1 - mu) * numpy.log(4.0) * Xf))
        ratio = fdm / (lnL * Xf)
        want = numpy.where(ratio < 1, cFR * ratio, cFR)
        assert numpy.allclose(FR, want, rtol=1e-10, atol=1e-15)
