"""DoubleExponentialDiskPotential (pot_type 11) on the GPU: pointwise forces
against the oracle's (in and off the plane, both packings) and the potential
against the oracle; full and planar orbits in disk + NFW halo against the
oracle's, short horizons (each force is ~2,500-term quadrature sums)."""

import numpy
import pytest

from helpers import (assert_parity, ics, oracle_forces, oracle_integrate, oracle_planar_forces,
                     oracle_potential, planar_ics, ulp_sensitivity)

pytestmark = pytest.mark.gpu


def _pot():
    from galpy_amd import potential as P

    return [P.DoubleExponentialDiskPotential(hr=0.3, hz=0.05, normalize=0.6), P.NFWPotential(a=2.0, normalize=0.4)]


def test_pointwise_forces_and_potential():
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, eval_device

    pot = _pot()
    _, pt, pa, _ = P.parse_pot(pot)
    _, ppt, ppa, _ = P.parse_pot_planar(pot)
    rng = numpy.random.default_rng(81)
    n = 1024
    R, z = rng.uniform(0.1, 3.0, n), rng.uniform(-0.5, 0.5, n)
    z[:64] = 0.0
    phi = rng.uniform(0.0, 2.0 * numpy.pi, n)
    dev = [torch.tensor(v, device="cuda") for v in (R, z, phi, numpy.zeros(n))]
    pot_d, FR, Fz, _ = (o.cpu().numpy() for o in eval_device(PotentialTable(pot), *dev))
    oFR, oFz, _ = oracle_forces(pt, pa, R, z, phi)
    assert numpy.abs(FR - oFR).max() <= 1e-13 * numpy.abs(oFR).max()
    assert numpy.abs(Fz - oFz).max() <= 1e-13 * numpy.abs(oFz).max()
    want = oracle_potential(pt, pa, R, z, phi)
    assert numpy.abs(pot_d - want).max() <= 1e-13 * numpy.abs(want).max()
    devp = [torch.tensor(v, device="cuda") for v in (R, numpy.zeros(n), phi, numpy.zeros(n))]
    _, pFR, _, _ = (o.cpu().numpy() for o in eval_device(PotentialTable(pot, planar=True), *devp))
    oFR, _ = oracle_planar_forces(ppt, ppa, R, phi)
    assert numpy.abs(pFR - oFR).max() <= 1e-13 * numpy.abs(oFR).max()


@pytest.mark.parametrize("method,code,bar,factor", [("symplec4_hip", 3, 1e-10, 1e3),
                                                     ("dop853_hip", 6, 1e-8, 10.0)])
def test_full_orbits(method, code, bar, factor):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = _pot()
    _, pt, pa, _ = P.parse_pot(pot)
    t = numpy.linspace(0.0, 2.0, 21)
    yo = ics(64, seed=83)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method, dt=0.01 if code == 3 else None)
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code, dt=0.01 if code == 3 else None)[0],
                                yo)
    assert_parity(res, ref, bar, sens, factor=factor, label=method)


def test_planar_orbits():
    from galpy_amd import potential as P
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = _pot()
    _, pt, pa, _ = P.parse_pot_planar(pot)
    t = numpy.linspace(0.0, 2.0, 21)
    yp = planar_ics(64, seed=85)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "dop853_hip")
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 6, planar=True)[0], yp)
    assert_parity(res, ref, 1e-8, sens, factor=10.0, label="planar dop853")
