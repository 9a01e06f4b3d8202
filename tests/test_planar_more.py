"""TransientLogSpiralPotential, CosmphiDiskPotential and HenonHeilesPotential
(planar only, integratePlanarOrbit's pot_type 2, 28 and 29).  The oracle's
planar force and torque are checked against mpmath derivatives of each
potential's closed form, written here from the class definitions
(TransientLogSpiralPotential.py, CosmphiDiskPotential.py,
HenonHeilesPotential.py); the host classes' forces against the oracle's;
CosmphiDisk's radial force is continuous at its break radius; the Henon-Heiles
energy is conserved; the full-orbit packing refuses them, as galpy has no 3D
version."""

import numpy
import pytest

from helpers import oracle_integrate, oracle_planar_forces


def _check(pot, fphi, R, phi, t, rtol, floor):
    """oracle F_R and F_phi against -d/dR and -d/dphi of fphi(R, phi) in mpmath."""
    import mpmath

    from galpy_amd.potential import parse_pot_planar

    _, pt, pa, _ = parse_pot_planar([pot])
    FR, Fp = oracle_planar_forces(pt, pa, R, phi, numpy.full(len(R), t))
    for i in range(len(R)):
        wR = -float(mpmath.diff(lambda r: fphi(r, mpmath.mpf(phi[i])), mpmath.mpf(R[i])))
        wp = -float(mpmath.diff(lambda ph: fphi(mpmath.mpf(R[i]), ph), mpmath.mpf(phi[i])))
        assert abs(FR[i] - wR) <= rtol * max(abs(wR), floor), (i, FR[i], wR)
        assert abs(Fp[i] - wp) <= rtol * max(abs(wp), floor), (i, Fp[i], wp)
    assert numpy.allclose(pot.Rforce(R, 0.0, phi, t), FR, rtol=rtol, atol=rtol * floor)
    assert numpy.allclose(pot.phitorque(R, 0.0, phi, t), Fp, rtol=rtol, atol=rtol * floor)
    return pt, pa


def test_transient_log_spiral_against_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    from galpy_amd.potential import TransientLogSpiralPotential, parse_pot

    amp, A, alpha, m, om, gam, sigma, to = 1.3, -0.035, -7.0, 2, 0.65, 0.7, 1.5, 2.0
    pot = TransientLogSpiralPotential(amp=amp, omegas=om, A=A, alpha=alpha, m=m, gamma=gam, sigma=sigma, to=to)
    with pytest.raises(NotImplementedError):
        parse_pot([pot])
    rng = numpy.random.default_rng(12)
    R, phi = rng.uniform(0.2, 4.0, 16), rng.uniform(0.0, 2.0 * numpy.pi, 16)
    for t in (to, to + sigma, 0.3):
        g = mpmath.exp(-(mpmath.mpf(t) - to) ** 2 / (2 * mpmath.mpf(sigma) ** 2))
        f = lambda r, ph: (amp * A * g / alpha
                           * mpmath.cos(alpha * mpmath.log(r) - m * (ph - mpmath.mpf(om) * t - mpmath.mpf(gam))))
        pt, pa = _check(pot, f, R, phi, t, 1e-12, 1e-4)
    assert list(pt) == [2] and len(pa) == 8 and pa[3] == sigma**2
    # the pitch-angle form gives alpha = m / tan(p)
    assert TransientLogSpiralPotential(p=0.3, m=2)._alpha == 2 / numpy.tan(0.3)


@pytest.mark.parametrize("rb", [None, 1.5])
def test_cosmphi_disk_against_mpmath(rb):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    from galpy_amd.potential import CosmphiDiskPotential, parse_pot, parse_pot_planar

    amp, phib, p, phio, m, r1 = 1.2, 0.4, 0.8, 0.02, 3, 1.3
    pot = CosmphiDiskPotential(amp=amp, phib=phib, p=p, phio=phio, m=m, r1=r1, rb=rb)
    with pytest.raises(NotImplementedError):
        parse_pot([pot])
    mphio, mphib = m * phio, m * phib
    r1p = mpmath.mpf(r1) ** p

    def f(r, ph):
        c = amp / r1p * mphio / m * mpmath.cos(m * ph - mphib)
        if rb is not None and r <= rb:
            rbp = mpmath.mpf(rb) ** p
            return c * rbp * (2 * r1p - rbp / r**p)
        return c * r**p

    rng = numpy.random.default_rng(13)
    R, phi = rng.uniform(0.2, 4.0, 24), rng.uniform(0.0, 2.0 * numpy.pi, 24)
    if rb is not None:  # stay off the kink, where the two one-sided derivatives differ in the second order
        R = R[numpy.abs(R - rb) > 1e-3]
        phi = phi[: len(R)]
    pt, pa = _check(pot, f, R, phi, 0.0, 1e-12, 1e-4)
    assert list(pt) == [28] and len(pa) == 9 and pa[4] == m
    if rb is not None:
        # at rb: F_R is continuous; the potential and F_phi are continuous
        # only for r1 = 1 (inside, rb^p (2 r1^p - 1) against rb^p outside),
        # as galpy's own definition has it
        Rs = numpy.array([rb * (1 - 1e-12), rb * (1 + 1e-12)])
        ph = numpy.array([0.9, 0.9])
        FR, Fp = oracle_planar_forces(pt, pa, Rs, ph)
        assert abs(FR[0] - FR[1]) < 1e-10 * abs(FR[0])
        assert abs(Fp[0] / Fp[1] - (2.0 * r1**p - 1.0)) < 1e-10
        one = CosmphiDiskPotential(amp=amp, phib=phib, p=p, phio=phio, m=m, r1=1.0, rb=rb)
        _, pt1, pa1, _ = parse_pot_planar([one])
        FR, Fp = oracle_planar_forces(pt1, pa1, Rs, ph)
        assert abs(Fp[0] - Fp[1]) < 1e-10 * abs(Fp[0])
        v = one(Rs, 0.0, ph)
        assert abs(v[0] - v[1]) < 1e-10 * abs(v[0])
    # (cp, sp) give the same potential as (phib, phio)
    alt = CosmphiDiskPotential(amp=amp, p=p, m=m, r1=r1, rb=rb, cp=mphio * numpy.cos(mphib),
                               sp=mphio * numpy.sin(mphib))
    assert numpy.allclose(alt(R, 0.0, phi), pot(R, 0.0, phi), rtol=1e-13, atol=0)


def test_henon_heiles_against_mpmath_and_energy():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    from galpy_amd.potential import HenonHeilesPotential, parse_pot, parse_pot_planar

    pot = HenonHeilesPotential(amp=0.8)
    with pytest.raises(NotImplementedError):
        parse_pot([pot])
    f = lambda r, ph: 0.8 * (r**2 / 2 + r**3 / 3 * mpmath.sin(3 * ph))
    rng = numpy.random.default_rng(14)
    R, phi = rng.uniform(0.05, 0.6, 16), rng.uniform(0.0, 2.0 * numpy.pi, 16)
    _check(pot, f, R, phi, 0.0, 1e-12, 1e-3)
    # Henon & Heiles' bound orbits at E = 1/12 (amp = 1): the oracle's DOP853
    # conserves E = (vx^2 + vy^2)/2 + (x^2 + y^2)/2 + x^2 y - y^3/3
    pot1 = HenonHeilesPotential()
    _, pt, pa, _ = parse_pot_planar([pot1])
    rows = []
    for y0, vy0 in ((0.1, 0.2), (-0.15, 0.05), (0.3, -0.1)):
        ep = 0.5 * y0**2 - y0**3 / 3.0
        vx0 = numpy.sqrt(2.0 * (1.0 / 12.0 - ep) - vy0**2)
        x, y, vx, vy = 0.0, y0, vx0, vy0
        Rr = numpy.hypot(x, y)
        ph = numpy.arctan2(y, x) % (2.0 * numpy.pi)
        rows.append([Rr, (x * vx + y * vy) / Rr, (x * vy - y * vx) / Rr, ph])
    yo = numpy.array(rows)
    t = numpy.linspace(0.0, 100.0, 501)
    res = oracle_integrate(pt, pa, yo, t, 6, planar=True)[0]
    Rr, vR, vT, ph = (res[..., k] for k in range(4))
    E = 0.5 * (vR**2 + vT**2) + pot1(Rr, 0.0, ph)
    assert numpy.abs(E * 12.0 - 1.0).max() < 1e-10  # DOP853 at rtol = atol = 1e-12 over 100 time units
