"""The spherical and Kuzmin-disk potentials on the host: galpy's packing
(integrateFullOrbit.py:82-92,176-213,358-365 and the planar packing,
integratePlanarOrbit.py), normalize, the Python classes' forces against the
oracle (whose forces are pinned to mpmath), the adjoint's differentiable
slots, and two closed forms the oracle's orbits must reproduce: the Kepler
period and the conservation of the angular-momentum vector in a spherical
potential."""

import numpy
import pytest

from helpers import cyl_to_rect, ics, oracle_forces, oracle_integrate, oracle_potential, spherical_a, spherical_b


def _classes():
    from galpy_amd import potential as P

    return [
        (P.PowerSphericalPotential(amp=2.0, alpha=1.7), 7, [2.0 * 4.0 * numpy.pi / 1.3, 1.7]),
        (P.KeplerPotential(amp=0.7), 7, [0.7, 3.0]),
        (P.JaffePotential(amp=1.1, a=0.8), 10, [1.1, 0.8]),
        (P.IsochronePotential(amp=1.2, b=0.7), 14, [1.2, 0.7]),
        (P.PlummerPotential(amp=0.9), 17, [0.9, 0.8]),
        (P.PseudoIsothermalPotential(amp=1.4, a=1.3), 18, [1.4, 1.3]),
        (P.KuzminDiskPotential(amp=1.0, a=0.6), 19, [1.0, 0.6]),
        (P.BurkertPotential(amp=0.8), 20, [0.8, 2.0]),
        (P.DehnenCoreSphericalPotential(amp=1.5, a=1.2), 33, [1.5, 1.2]),
        (P.DehnenSphericalPotential(amp=1.3, a=0.9), 34, [1.3, 0.9, 1.5]),
        (P.HomogeneousSpherePotential(amp=1.0), 35, [1.0, 1.1**2, 1.1**3]),
        (P.FlattenedPowerPotential(amp=1.2, alpha=0.5, q=0.9, core=0.1), 12, [1.2, 0.5, 0.81, 0.01]),
    ]


def test_packing_full_and_planar():
    from galpy_amd.potential import parse_pot, parse_pot_planar

    for pot, code, args in _classes():
        for parse in (parse_pot, parse_pot_planar):
            n, pt, pa, tf = parse([pot])
            assert n == 1 and list(pt) == [code] and tf == []
            want = [args[0], args[1], args[3]] if (code == 12 and parse is parse_pot_planar) else args
            numpy.testing.assert_allclose(pa, want, rtol=1e-15, atol=0)
    # NullPotential: dropped unless nothing else is left (integrateFullOrbit.py:51-56)
    from galpy_amd.potential import HernquistPotential, NullPotential

    n, pt, pa, _ = parse_pot([NullPotential(), HernquistPotential(), NullPotential()])
    assert n == 1 and list(pt) == [8] and list(pa) == [1.0, 1.0]
    n, pt, pa, _ = parse_pot([NullPotential()])
    assert n == 1 and list(pt) == [40] and len(pa) == 0
    # power-law amplitude: mass within r1 (PowerSphericalPotential.py:49-53)
    from galpy_amd.potential import PowerSphericalPotential

    p = PowerSphericalPotential(amp=1.0, alpha=2.5, r1=2.0)
    assert abs(p._amp - 2.0 ** (-0.5) * 4.0 * numpy.pi / 0.5) < 1e-13


def test_dehnen_alpha_range():
    from galpy_amd.potential import DehnenSphericalPotential

    for bad in (-0.1, 3.0, 3.5):
        with pytest.raises(OSError):
            DehnenSphericalPotential(alpha=bad)


@pytest.mark.parametrize("idx", range(12))
def test_class_forces_and_normalize_match_oracle(idx):
    """The numpy forces and potential of each class equal the oracle's
    (pinned to mpmath); normalize(f) makes -Rforce(1, 0) = f."""
    from galpy_amd.potential import parse_pot

    pot, code, _ = _classes()[idx]
    rng = numpy.random.default_rng(idx)
    R = rng.uniform(0.05, 4.0, 500)
    z = rng.normal(0.0, 0.5, 500)
    _, pt, pa, _ = parse_pot([pot])
    FR, Fz, _ = oracle_forces(pt, pa, R, z)
    pv = oracle_potential(pt, pa, R, z)
    for got, want in ((pot.Rforce(R, z), FR), (pot.zforce(R, z), Fz), (pot(R, z), pv)):
        assert numpy.abs(got - want).max() <= 1e-12 * numpy.abs(want).max()
    pot.normalize(0.37)
    assert abs(pot.Rforce(1.0, 0.0) + 0.37) < 1e-14
    _, pt, pa, _ = parse_pot([pot])
    assert abs(oracle_forces(pt, pa, numpy.array([1.0]), 0.0)[0][0] + 0.37) < 1e-14


def test_normalized_groups_have_unit_circular_velocity():
    from galpy_amd.potential import evaluateRforces

    for pot in (spherical_a(), spherical_b()):
        assert abs(evaluateRforces(pot, 1.0, 0.0) + 1.0) < 1e-13


def test_paramspec_differentiable_slots():
    from galpy_amd.adjoint import ParamSpec

    pot = spherical_a()
    spec = ParamSpec(pot, [(0, 0), (1, 0), (2, 0), (3, 0)])
    numpy.testing.assert_array_equal(spec.theta0(), [p._amp for p in pot])
    for bad in ((0, 1), (1, 1), (3, 1), (4, 0)):
        with pytest.raises(ValueError):
            ParamSpec(pot, [bad])
    from galpy_amd.potential import MWPotential2014

    ParamSpec(MWPotential2014, [(1, 2), (2, 1), (0, 0)])
    for bad in ((0, 1), (0, 2), (2, 2)):
        with pytest.raises(ValueError):
            ParamSpec(MWPotential2014, [bad])


def test_kepler_period_closed_form():
    """A Kepler orbit of semi-major axis a returns to its initial state
    after T = 2 pi a^1.5 / sqrt(GM) (DOP853 at rtol = atol = 1e-13)."""
    from galpy_amd.potential import KeplerPotential, parse_pot

    GM = 1.3
    _, pt, pa, _ = parse_pot([KeplerPotential(amp=GM)])
    R0, vT0, vz0 = 1.0, 0.9, 0.35  # inclined, eccentric
    E = 0.5 * (vT0**2 + vz0**2) - GM / R0
    a = -GM / (2.0 * E)
    T = 2.0 * numpy.pi * a**1.5 / numpy.sqrt(GM)
    yo = numpy.array([[R0, 0.0, vT0, 0.0, vz0, 0.0]])
    tol = numpy.log(1e-13)
    res, err, _, _ = oracle_integrate(pt, pa, yo, numpy.array([0.0, 0.5 * T, T]), 6, rtol=tol, atol=tol)
    assert err[0] == 0
    assert numpy.abs(res[0, -1, :5] - yo[0, :5]).max() < 1e-9
    assert abs((res[0, -1, 5] + numpy.pi) % (2 * numpy.pi) - numpy.pi) < 1e-9
    # the start (vR = 0, speed below circular) is the apocentre, so half a
    # period later the orbit is at pericentre: r = a (1 - e), vR = 0
    e = numpy.sqrt(1.0 - (R0 * numpy.hypot(vT0, vz0)) ** 2 / (GM * a))
    half = res[0, 1]
    assert abs(R0 - a * (1.0 + e)) < 1e-14
    assert abs(numpy.hypot(half[0], half[3]) - a * (1.0 - e)) < 1e-9


@pytest.mark.parametrize("code", [3, 0])
def test_spherical_orbits_conserve_angular_momentum(code):
    """In a spherical potential every kick is radial, so the symplectic
    steppers conserve the angular-momentum vector to rounding."""
    from galpy_amd.potential import parse_pot

    _, pt, pa, _ = parse_pot(spherical_b())
    yo = ics(16, seed=9)
    t = numpy.linspace(0.0, 20.0, 201)
    res, err, _, _ = oracle_integrate(pt, pa, yo, t, code, dt=0.01)
    assert numpy.all(err == 0)
    w = cyl_to_rect(res.reshape(-1, 6)).reshape(16, 201, 6)
    L = numpy.cross(w[..., :3], w[..., 3:])
    assert numpy.abs(L - L[:, :1]).max() < 1e-12 * numpy.abs(L).max()
