"""Spline-interpolated potentials: interpSphericalPotential (36) and
MovingObjectPotential (-6).  galpy's C interpolates both with GSL's natural
cubic spline; the oracle restates that spline and is pinned here against
scipy's CubicSpline(bc_type='natural') -- values, integrals and, through
the forces, the whole construction -- and against closed forms (a Plummer
sphere moving on a straight line; Kepler outside the grid)."""

import numpy
import pytest
from scipy import interpolate

from helpers import oracle_forces, oracle_planar_forces, oracle_potential


def test_interp_spherical_against_scipy_natural_spline():
    from galpy_amd import potential as P

    hern = P.HernquistPotential(amp=2.0, a=0.8)
    rgrid = numpy.geomspace(0.05, 10.0, 61)
    pot = P.interpSphericalPotential(rforce=hern, rgrid=rgrid)
    _, pt, pa, _ = P.parse_pot([pot])
    assert list(pt) == [36] and len(pa) == 1 + 2 * 61 + 6
    cs = interpolate.CubicSpline(rgrid, pot._rforce_grid, bc_type="natural")
    rng = numpy.random.default_rng(36)
    n = 200
    r = numpy.concatenate([rng.uniform(0.05, 10.0, n - 4), [0.01, 0.04, 10.0, 15.0]])
    th = rng.uniform(0.1, numpy.pi - 0.1, r.size)
    R, z = r * numpy.sin(th), r * numpy.cos(th)
    FR, Fz, Fp = oracle_forces(pt, pa, R, z)
    V = oracle_potential(pt, pa, R, z)
    Mmax, Phi0, Phimax = pa[-3], pa[-2], pa[-1]
    inside = (r >= 0.05) & (r < 10.0)
    Fr = numpy.where(r >= 10.0, -Mmax / r**2, numpy.where(r < 0.05, 0.0, cs(numpy.clip(r, 0.05, 10.0))))
    want_V = numpy.where(r >= 10.0, -Mmax / r + Phimax,
                         numpy.where(r < 0.05, 0.0, [-cs.integrate(0.05, x) + Phi0 for x in numpy.clip(r, 0.05, 10.0)]))
    assert numpy.allclose(FR, Fr * R / r, rtol=1e-13, atol=1e-15) and numpy.allclose(Fz, Fr * z / r, rtol=1e-13,
                                                                                     atol=1e-15)
    assert numpy.allclose(V, want_V, rtol=1e-13, atol=1e-15)
    assert numpy.all(Fp == 0.0)
    # the table reproduces the Hernquist force to interpolation accuracy
    assert numpy.allclose(FR[inside], hern.Rforce(R[inside], z[inside]), rtol=2e-3)
    # planar packing: the same spline in the plane
    _, ppt, ppa, _ = P.parse_pot_planar([pot])
    pFR, _ = oracle_planar_forces(ppt, ppa, r)
    assert numpy.allclose(pFR, Fr, rtol=1e-13, atol=1e-15)


class _LineOrbit:
    """A stand-in for an integrated orbit: uniform motion x = x0 + v t."""

    def __init__(self, t, x0, v):
        self.t = t
        self._x0, self._v = numpy.array(x0), numpy.array(v)

    def x(self, t):
        return self._x0[0] + self._v[0] * t

    def y(self, t):
        return self._x0[1] + self._v[1] * t

    def z(self, t):
        return self._x0[2] + self._v[2] * t


def test_moving_object_on_a_line_is_a_shifted_plummer():
    """A natural spline reproduces a straight line exactly, so the moving
    Plummer sphere's forces are the Plummer forces at x - x_obj(t)."""
    from galpy_amd import potential as P

    tgrid = numpy.linspace(0.0, 10.0, 41)
    orb = _LineOrbit(tgrid, [1.0, -0.5, 0.2], [-0.1, 0.15, 0.02])
    pl = P.PlummerPotential(amp=0.3, b=0.2)
    mo = P.MovingObjectPotential(orb, pot=pl, amp=1.3)
    _, pt, pa, _ = P.parse_pot([mo])
    assert list(pt) == [-6, 17] and pa[0] == 1
    rng = numpy.random.default_rng(6)
    n = 64
    R, z, phi = rng.uniform(0.2, 2.5, n), rng.uniform(-1.0, 1.0, n), rng.uniform(0.0, 2.0 * numpy.pi, n)
    t = numpy.concatenate([rng.uniform(0.0, 10.0, n - 4), [-1.0, 0.0, 10.0, 12.0]])  # clamped outside
    FR, Fz, Fp = oracle_forces(pt, pa, R, z, phi, t)
    tc = numpy.clip(t, 0.0, 10.0)
    X = numpy.stack([R * numpy.cos(phi), R * numpy.sin(phi), z], axis=1)
    O = numpy.stack([orb.x(tc), orb.y(tc), orb.z(tc)], axis=1)
    d = X - O
    F = -1.3 * 0.3 * d / ((d**2).sum(axis=1) + 0.04)[:, None] ** 1.5
    wR = F[:, 0] * numpy.cos(phi) + F[:, 1] * numpy.sin(phi)
    wp = R * (-F[:, 0] * numpy.sin(phi) + F[:, 1] * numpy.cos(phi))
    assert numpy.allclose(FR, wR, rtol=1e-12, atol=1e-15)
    assert numpy.allclose(Fz, F[:, 2], rtol=1e-12, atol=1e-15)
    assert numpy.allclose(Fp, wp, rtol=1e-12, atol=1e-15)
    V = oracle_potential(pt, pa, R, z, phi, t)
    assert numpy.allclose(V, -1.3 * 0.3 / numpy.sqrt((d**2).sum(axis=1) + 0.04), rtol=1e-12)
    # host class (scipy natural splines) equals the oracle
    assert numpy.allclose(mo.Rforce(R, z, phi, t), FR, rtol=1e-12, atol=1e-15)
    assert numpy.allclose(mo.zforce(R, z, phi, t), Fz, rtol=1e-12, atol=1e-15)


def test_moving_object_curved_orbit_against_scipy_splines():
    """On a curved orbit the object's position is the natural spline of the
    stored positions (scipy's CubicSpline, bc natural, over normalised
    time); planar packing included."""
    from galpy_amd import potential as P

    tgrid = numpy.linspace(-2.0, 6.0, 33)

    class Circ:
        t = tgrid

        def x(self, t):
            return 1.5 * numpy.cos(0.8 * t)

        def y(self, t):
            return 1.5 * numpy.sin(0.8 * t)

        def z(self, t):
            return 0.1 * numpy.sin(1.3 * t)

    mo = P.MovingObjectPotential(Circ(), pot=P.HernquistPotential(amp=0.2, a=0.3))
    _, pt, pa, _ = P.parse_pot([mo])
    _, ppt, ppa, _ = P.parse_pot_planar([mo])
    assert len(ppa) == len(pa) - 33
    rng = numpy.random.default_rng(7)
    n = 64
    R, z, phi, t = (rng.uniform(0.2, 2.5, n), rng.uniform(-0.5, 0.5, n), rng.uniform(0.0, 6.3, n),
                    rng.uniform(-2.0, 6.0, n))
    FR, Fz, Fp = oracle_forces(pt, pa, R, z, phi, t)
    assert numpy.allclose(mo.Rforce(R, z, phi, t), FR, rtol=1e-11, atol=1e-14)
    assert numpy.allclose(mo.zforce(R, z, phi, t), Fz, rtol=1e-11, atol=1e-14)
    assert numpy.allclose(mo.phitorque(R, z, phi, t), Fp, rtol=1e-11, atol=1e-14)
    assert numpy.allclose(mo(R, z, phi, t), oracle_potential(pt, pa, R, z, phi, t), rtol=1e-12)
    pFR, pFp = oracle_planar_forces(ppt, ppa, R, phi, t)
    # in the plane the planar kernel sees no height difference
    h = P.HernquistPotential(amp=0.2, a=0.3)
    ox, oy, _ = mo._pos(t)
    x, y = R * numpy.cos(phi), R * numpy.sin(phi)
    Rd = numpy.hypot(x - ox, y - oy)
    RF = h.Rforce(Rd, 0.0)
    assert numpy.allclose(pFR, -RF * (numpy.cos(phi) * (ox - x) + numpy.sin(phi) * (oy - y)) / Rd, rtol=1e-12)
