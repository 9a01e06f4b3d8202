"""EllipticalDiskPotential, LopsidedDiskPotential and SteadyLogSpiralPotential
(planar only, integratePlanarOrbit's pot_type 4, 6 and 3).  EllipticalDisk:
the oracle's planar force and torque are mpmath derivatives of
Phi = amp smooth(t) twophio/2 (R/r1)^p cos(2 (phi - phib)); Dehnen's growth
polynomial is 0 before tform, 1/2 half-way, 1 after tsteady, with zero slope
at both ends; the full-orbit packing refuses it, as galpy has no 3D
version."""

import numpy
import pytest

from helpers import oracle_planar_forces


def test_ellipdisk_against_mpmath_and_smooth():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    from galpy_amd.potential import EllipticalDiskPotential, parse_pot, parse_pot_planar

    pot = EllipticalDiskPotential(amp=1.3, phib=0.4, p=0.7, twophio=0.05, r1=1.2, tform=1.0, tsteady=3.0)
    _, pt, pa, _ = parse_pot_planar([pot])
    assert list(pt) == [4] and len(pa) == 6
    with pytest.raises(NotImplementedError):
        parse_pot([pot])
    rng = numpy.random.default_rng(4)
    R, phi = rng.uniform(0.1, 4.0, 20), rng.uniform(-3.0, 3.0, 20)
    for t, sm in ((0.5, 0.0), (2.5, 0.5), (5.0, 1.0)):
        FR, Fp = oracle_planar_forces(pt, pa, R, phi, numpy.full(20, t))
        for i in range(20):
            f = lambda r, ph: (1.3 / mpmath.mpf(1.2) ** 0.7 * sm * mpmath.mpf(0.05) / 2
                               * mpmath.mpf(r) ** 0.7 * mpmath.cos(2 * (ph - mpmath.mpf(0.4))))
            wR = -float(mpmath.diff(lambda r: f(r, phi[i]), R[i]))
            wp = -float(mpmath.diff(lambda ph: f(R[i], ph), phi[i]))
            assert abs(FR[i] - wR) <= 1e-13 * max(abs(wR), 1e-6) + 1e-18
            assert abs(Fp[i] - wp) <= 1e-13 * max(abs(wp), 1e-6) + 1e-18
        assert numpy.allclose(pot.Rforce(R, 0.0, phi, t), FR, rtol=1e-13, atol=1e-18)
        assert numpy.allclose(pot.phitorque(R, 0.0, phi, t), Fp, rtol=1e-13, atol=1e-18)
    # zero slope of the growth at tform = 1 and tform + tsteady = 4 (galpy's
    # tsteady is the duration of the growth)
    one = lambda t: oracle_planar_forces(pt, pa, numpy.array([1.0]), numpy.array([0.0]), numpy.array([t]))[0][0]
    full = one(10.0)
    for t0 in (1.0, 4.0):
        assert abs(one(t0 + 1e-4) - one(t0 - 1e-4)) < 1e-7 * abs(full)


def test_lopsided_against_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    from galpy_amd.potential import LopsidedDiskPotential, parse_pot, parse_pot_planar

    pot = LopsidedDiskPotential(amp=1.1, phib=0.6, p=-0.3, phio=0.02, r1=3.0)
    _, pt, pa, _ = parse_pot_planar([pot])
    assert list(pt) == [6] and list(pa) == [1.1, 0.02, -0.3, 0.6]  # r1 unused, as in galpy
    with pytest.raises(NotImplementedError):
        parse_pot([pot])
    rng = numpy.random.default_rng(6)
    R, phi = rng.uniform(0.1, 4.0, 20), rng.uniform(-3.0, 3.0, 20)
    FR, Fp = oracle_planar_forces(pt, pa, R, phi, numpy.zeros(20))
    f = lambda r, ph: mpmath.mpf(1.1) * mpmath.mpf(0.02) * mpmath.mpf(r) ** mpmath.mpf(-0.3) * mpmath.cos(ph - mpmath.mpf(0.6))
    for i in range(20):
        wR = -float(mpmath.diff(lambda r: f(r, phi[i]), R[i]))
        wp = -float(mpmath.diff(lambda ph: f(R[i], ph), phi[i]))
        assert abs(FR[i] - wR) <= 1e-13 * max(abs(wR), 1e-6)
        assert abs(Fp[i] - wp) <= 1e-13 * max(abs(wp), 1e-6)
    assert numpy.allclose(pot.Rforce(R, 0.0, phi), FR, rtol=1e-13, atol=0)
    assert numpy.allclose(pot.phitorque(R, 0.0, phi), Fp, rtol=1e-13, atol=0)


def test_steady_log_spiral_against_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    from galpy_amd.potential import SteadyLogSpiralPotential, parse_pot_planar

    pot = SteadyLogSpiralPotential(amp=1.2, omegas=0.65, A=-0.035, alpha=-7.0, m=2, gamma=0.7,
                                   tform=1.0, tsteady=1.5)
    _, pt, pa, _ = parse_pot_planar([pot])
    ts = 2.0 * numpy.pi / 0.65
    assert list(pt) == [3] and pa[1] == ts and pa[2] == ts + 1.5 * ts
    rng = numpy.random.default_rng(7)
    R, phi = rng.uniform(0.2, 4.0, 20), rng.uniform(-3.0, 3.0, 20)
    for t, sm in ((0.5 * ts, 0.0), (1.75 * ts, 0.5), (4.0 * ts, 1.0)):
        FR, Fp = oracle_planar_forces(pt, pa, R, phi, numpy.full(20, t))
        f = lambda r, ph: (mpmath.mpf(1.2) * sm * mpmath.mpf(-0.035) / -7
                           * mpmath.cos(-7 * mpmath.log(r) - 2 * (ph - mpmath.mpf(0.65) * t - mpmath.mpf(0.7))))
        for i in range(20):
            wR = -float(mpmath.diff(lambda r: f(r, phi[i]), R[i]))
            wp = -float(mpmath.diff(lambda ph: f(R[i], ph), phi[i]))
            assert abs(FR[i] - wR) <= 1e-12 * max(abs(wR), 1e-4) + 1e-18
            assert abs(Fp[i] - wp) <= 1e-12 * max(abs(wp), 1e-4) + 1e-18
        assert numpy.allclose(pot.Rforce(R, 0.0, phi, t), FR, rtol=1e-12, atol=1e-18)
        assert numpy.allclose(pot.phitorque(R, 0.0, phi, t), Fp, rtol=1e-12, atol=1e-18)
