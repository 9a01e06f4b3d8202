"""KuzminKutuzovStaeckelPotential (pot_type 16) pinned by what the
mathematics fixes: (l, n) are the two roots of the spheroidal-coordinate
quadratic (t + alpha)(t + gamma) - R^2 (t + gamma) - z^2 (t + alpha) = 0,
solved here independently in 40-digit mpmath, Phi = -amp / (sqrt(l) +
sqrt(n)) there, and the forces are mpmath derivatives of that Phi; far out
the potential is that of a point mass amp.  The oracle's full and planar
forces and potential and the Python class are held to these."""

import numpy
import pytest

from helpers import oracle_forces, oracle_planar_forces, oracle_potential


def _phi_mp(mpmath, amp, ac, Delta, R, z):
    gamma = mpmath.mpf(Delta) ** 2 / (1 - mpmath.mpf(ac) ** 2)
    alpha = gamma - mpmath.mpf(Delta) ** 2
    R, z = mpmath.mpf(R), mpmath.mpf(z)
    b = alpha + gamma - R**2 - z**2
    c = alpha * gamma - R**2 * gamma - z**2 * alpha
    disc = mpmath.sqrt(b**2 - 4 * c)
    lam, nu = (-b + disc) / 2, (-b - disc) / 2
    # both are coordinates: lam >= -alpha >= nu >= -gamma
    assert lam >= -alpha - mpmath.mpf(10) ** -30 and nu <= -alpha + mpmath.mpf(10) ** -30
    return -amp / (mpmath.sqrt(lam) + mpmath.sqrt(nu))


@pytest.mark.parametrize("ac,Delta", [(5.0, 1.0), (1.5, 0.3), (20.0, 2.0)])
def test_kks_against_mpmath(ac, Delta):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    from galpy_amd.potential import KuzminKutuzovStaeckelPotential, parse_pot

    pot = KuzminKutuzovStaeckelPotential(amp=1.3, ac=ac, Delta=Delta)
    _, pt, pa, _ = parse_pot([pot])
    rng = numpy.random.default_rng(5)
    R, z = rng.uniform(0.05, 5.0, 24), rng.normal(0.0, 1.0, 24)
    FR, Fz, _ = oracle_forces(pt, pa, R, z)
    P = oracle_potential(pt, pa, R, z)
    for i in range(len(R)):
        f = lambda a, b: _phi_mp(mpmath, 1.3, ac, Delta, a, b)
        want = (float(f(R[i], z[i])), -float(mpmath.diff(lambda a: f(a, z[i]), R[i])),
                -float(mpmath.diff(lambda b: f(R[i], b), z[i])))
        for got, w in zip((P[i], FR[i], Fz[i]), want):
            assert abs(got - w) <= 1e-12 * max(abs(w), 1e-3), (i, got, w)
    assert numpy.allclose(pot.Rforce(R, z), FR, rtol=1e-12, atol=0)
    assert numpy.allclose(pot.zforce(R, z), Fz, rtol=1e-12, atol=1e-15)
    assert numpy.allclose(pot(R, z), P, rtol=1e-13, atol=0)


def test_kks_point_mass_far_out_and_planar():
    from galpy_amd.potential import KuzminKutuzovStaeckelPotential, parse_pot, parse_pot_planar

    pot = KuzminKutuzovStaeckelPotential(amp=0.7, ac=5.0, Delta=1.0)
    _, pt, pa, _ = parse_pot([pot])
    r = 1e5
    for th in (0.1, 0.8, 1.5):
        val = oracle_potential(pt, pa, numpy.array([r * numpy.sin(th)]), numpy.array([r * numpy.cos(th)]))[0]
        assert abs(r * val + 0.7) < 1e-4
    _, ppt, ppa, _ = parse_pot_planar([pot])
    R = numpy.linspace(0.1, 4.0, 40)
    FRp = oracle_planar_forces(ppt, ppa, R)[0]
    FR, Fz, _ = oracle_forces(pt, pa, R, numpy.zeros_like(R))
    assert numpy.allclose(FRp, FR, rtol=1e-13, atol=0) and numpy.all(Fz == 0.0)
