"""TransientLogSpiralPotential (pot_type 2), CosmphiDiskPotential (28) and
HenonHeilesPotential (29), planar only, on the GPU: pointwise forces against
the oracle's, potential values against the host classes, planar orbits
(symplec4, DOP853) against the oracle's, and the Henon-Heiles energy along
device-integrated rows."""

import numpy
import pytest

from helpers import assert_parity, oracle_integrate, oracle_planar_forces, planar_ics, ulp_sensitivity

pytestmark = pytest.mark.gpu


def _pots():
    from galpy_amd import potential as P

    return [P.LogarithmicHaloPotential(normalize=1.0),
            P.TransientLogSpiralPotential(A=-0.02, alpha=-6.0, m=1.5, sigma=2.0, to=3.0),
            P.CosmphiDiskPotential(phio=0.01, p=0.7, m=3, r1=1.2, rb=0.8, phib=0.5)]


def _packed(pot):
    from galpy_amd.potential import parse_pot_planar

    _, pt, pa, _ = parse_pot_planar(pot)
    return pt, pa


def test_pointwise_forces_and_potential():
    """Device F_R, F_phi at angles all around the circle (phi from acos, in
    [0, 2 pi): the transient spiral has m = 1.5), radii on both sides of the
    break radius, times around the spiral's peak, against the oracle; the
    potential against the host classes."""
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, eval_device

    pot = _pots() + [P.HenonHeilesPotential(amp=0.1)]
    pt, pa = _packed(pot)
    tab = PotentialTable(pot, planar=True)
    rng = numpy.random.default_rng(51)
    n = 4096
    R = rng.uniform(0.1, 3.0, n)
    R[:64] = 0.8 * (1.0 + rng.uniform(-1e-9, 1e-9, 64))  # at the break radius
    phi = rng.uniform(0.0, 2.0 * numpy.pi, n)
    t = rng.uniform(-2.0, 8.0, n)
    dev = [torch.tensor(v, device="cuda") for v in (R, numpy.zeros_like(R), phi, t)]
    pot_d, FR, _, Fp = (o.cpu().numpy() for o in eval_device(tab, *dev))
    oFR, oFp = oracle_planar_forces(pt, pa, R, phi, t)
    assert numpy.abs(FR - oFR).max() <= 1e-13 * numpy.abs(oFR).max()
    assert numpy.abs(Fp - oFp).max() <= 1e-13 * numpy.abs(oFp).max()
    phw = numpy.arctan2(numpy.sin(phi), numpy.cos(phi))
    want = P.evaluatePotentials(pot, R, 0.0, phw, t)
    assert numpy.abs(pot_d - want).max() <= 1e-13 * numpy.abs(want).max()


@pytest.mark.parametrize("method,code,bar,factor", [("symplec4_hip", 3, 1e-10, 1e3),
                                                     ("dop853_hip", 6, 1e-8, 10.0)])
def test_planar_orbits(method, code, bar, factor):
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = _pots()
    pt, pa = _packed(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yp = planar_ics(256, seed=53)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, method)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code, planar=True)[0], yp)
    assert numpy.all(err == 0)
    assert_parity(res, ref, bar, sens, factor=factor, label=method)


def test_henon_heiles_orbits_and_energy():
    """Bound Henon-Heiles orbits at E = 1/12: device rows against the
    oracle's, and E conserved along them (DOP853, galpy's tolerances)."""
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, energy_drift_device, integratePlanarOrbit_hip

    pot = [P.HenonHeilesPotential()]
    pt, pa = _packed(pot)
    rng = numpy.random.default_rng(55)
    n = 512
    y0 = rng.uniform(-0.25, 0.35, n)
    vy0 = rng.uniform(-0.15, 0.15, n)
    vx0 = numpy.sqrt(2.0 * (1.0 / 12.0 - (0.5 * y0**2 - y0**3 / 3.0)) - vy0**2)
    x0 = numpy.full(n, 1e-3)
    R0 = numpy.hypot(x0, y0)
    yp = numpy.stack([R0, (x0 * vx0 + y0 * vy0) / R0, (x0 * vy0 - y0 * vx0) / R0,
                      numpy.arctan2(y0, x0) % (2.0 * numpy.pi)], axis=1)
    t = numpy.linspace(0.0, 50.0, 251)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "dop853_hip")
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 6, planar=True)[0], yp)
    assert_parity(res, ref, 1e-8, sens, factor=10.0, label="henon-heiles")
    tab = PotentialTable(pot, planar=True)
    e0, demax = energy_drift_device(tab, torch.tensor(res, device="cuda"), torch.tensor(t, device="cuda"))
    Rr, vR, vT, ph = (res[..., k] for k in range(4))
    E = P.evaluatePotentials(pot, Rr, 0.0, ph) + 0.5 * (vR**2 + vT**2)
    assert numpy.abs(e0.cpu().numpy() - E[:, 0]).max() <= 1e-14
    assert demax.cpu().numpy().max() < 1e-9
