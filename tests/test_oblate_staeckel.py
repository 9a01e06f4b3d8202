"""OblateStaeckelWrapperPotential (-3): the wrapped potential put in Staeckel
form in prolate spheroidal coordinates (OblateStaeckelWrapperPotential.c).
Pinned by: around a KuzminKutuzovStaeckel potential of the same focal
length it is that potential (a Staeckel potential is its own Staeckel
form), forces and potential; around MWPotential2014 its forces are minus
the gradient of its potential; at the reference point it equals the wrapped
potential; the planar packing's force (no V term) is the full force in the
plane to the size of V(pi/2), which vanishes up to rounding."""

import numpy

from helpers import oracle_forces, oracle_planar_forces, oracle_potential


def _pts(seed, n=40):
    rng = numpy.random.default_rng(seed)
    return rng.uniform(0.1, 3.0, n), rng.uniform(-1.5, 1.5, n)


def test_staeckel_potential_is_its_own_staeckel_form():
    from galpy_amd import potential as P

    kks = P.KuzminKutuzovStaeckelPotential(amp=1.3, ac=5.0, Delta=0.7)
    w = P.OblateStaeckelWrapperPotential(pot=kks, delta=0.7, u0=(1.0, 0.2))
    _, pt, pa, _ = P.parse_pot([w])
    _, kt, ka, _ = P.parse_pot([kks])
    assert list(pt) == [-3, 16] and len(pa) == 1 + 3 + 5
    R, z = _pts(1)
    a, b = oracle_forces(pt, pa, R, z), oracle_forces(kt, ka, R, z)
    # the 1e-12 that galpy adds to tanh(u) in dU/du is the only difference
    assert numpy.allclose(a[0], b[0], rtol=1e-10, atol=0) and numpy.allclose(a[1], b[1], rtol=1e-10, atol=0)
    assert numpy.all(a[2] == 0.0)
    assert numpy.allclose(oracle_potential(pt, pa, R, z), oracle_potential(kt, ka, R, z), rtol=1e-13, atol=0)
    _, ppt, ppa, _ = P.parse_pot_planar([w])
    _, pkt, pka, _ = P.parse_pot_planar([kks])
    assert numpy.allclose(oracle_planar_forces(ppt, ppa, R)[0], oracle_planar_forces(pkt, pka, R)[0], rtol=1e-10)


def test_forces_are_minus_gradient_and_reference_point():
    from galpy_amd import potential as P

    mw = P.MWPotential2014_factory()
    w = P.OblateStaeckelWrapperPotential(pot=mw, delta=0.45, u0=(1.2, 0.1), amp=0.9)
    _, pt, pa, _ = P.parse_pot([w])
    R, z = _pts(2)
    FR, Fz, _ = oracle_forces(pt, pa, R, z)
    h = 1e-5
    V = lambda RR, zz: oracle_potential(pt, pa, RR, zz)
    dR = (-V(R + 2 * h, z) + 8 * V(R + h, z) - 8 * V(R - h, z) + V(R - 2 * h, z)) / (12 * h)
    dz = (-V(R, z + 2 * h) + 8 * V(R, z + h) - 8 * V(R, z - h) + V(R, z - 2 * h)) / (12 * h)
    assert numpy.allclose(FR, -dR, rtol=1e-7, atol=1e-9) and numpy.allclose(Fz, -dz, rtol=1e-7, atol=1e-9)
    # at the reference point (u0, pi/2) the Staeckel form is the wrapped potential
    R0 = 0.45 * numpy.sinh(w._u0)
    _, mt, ma, _ = P.parse_pot(mw)
    assert numpy.allclose(oracle_potential(pt, pa, [R0], [0.0]), 0.9 * oracle_potential(mt, ma, [R0], [0.0]),
                          rtol=1e-14)
    # the host class restates the same formulas
    assert numpy.allclose(w.Rforce(R, z), FR, rtol=1e-11, atol=1e-14)
    assert numpy.allclose(w.zforce(R, z), Fz, rtol=1e-11, atol=1e-14)
    assert numpy.allclose(w(R, z), oracle_potential(pt, pa, R, z), rtol=1e-12)
    # planar: the full force in the plane up to V(pi/2) ~ rounding
    _, ppt, ppa, _ = P.parse_pot_planar([w])
    assert numpy.allclose(oracle_planar_forces(ppt, ppa, R)[0], oracle_forces(pt, pa, R, 0 * R)[0], rtol=1e-9)
