"""interpSphericalPotential (36) and MovingObjectPotential (-6) on the GPU:
natural cubic splines built on the host, evaluated on the device; pointwise
forces and potential against the oracle's, full and planar orbits against
the oracle's.  The moving object follows an orbit integrated on the GPU."""

import numpy
import pytest

from helpers import (assert_parity, ics, oracle_forces, oracle_integrate, oracle_potential, planar_ics,
                     ulp_sensitivity)

pytestmark = pytest.mark.gpu


def _interp():
    from galpy_amd import potential as P

    halo = P.interpSphericalPotential(rforce=P.HernquistPotential(amp=2.0, a=0.8), rgrid=numpy.geomspace(0.02, 30.0, 201))
    return [halo, P.MiyamotoNagaiPotential(a=0.5, b=0.1, normalize=0.3)]


def _moving():
    from galpy_amd import potential as P
    from galpy_amd.orbit import Orbit

    sat = Orbit([1.5, 0.1, 0.9, 0.2, 0.05, 0.3])
    sat.integrate(numpy.linspace(0.0, 12.0, 241), P.MWPotential2014_factory(), method="dop853_hip")
    return list(P.MWPotential2014_factory()) + [P.MovingObjectPotential(sat, pot=P.PlummerPotential(amp=0.02, b=0.1))]


@pytest.mark.parametrize("which", ["interp", "moving"])
def test_pointwise_forces_and_potential(which):
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, eval_device

    pot = _interp() if which == "interp" else _moving()
    _, pt, pa, _ = P.parse_pot(pot)
    rng = numpy.random.default_rng(111)
    n = 4096
    R, z = rng.uniform(0.01, 3.0, n), rng.uniform(-1.0, 1.0, n)
    phi, t = rng.uniform(0.0, 2.0 * numpy.pi, n), rng.uniform(-1.0, 13.0, n)
    dev = [torch.tensor(v, device="cuda") for v in (R, z, phi, t)]
    pot_d, FR, Fz, Fp = (o.cpu().numpy() for o in eval_device(PotentialTable(pot), *dev))
    oFR, oFz, oFp = oracle_forces(pt, pa, R, z, phi, t)
    assert numpy.abs(FR - oFR).max() <= 1e-13 * numpy.abs(oFR).max()
    assert numpy.abs(Fz - oFz).max() <= 1e-13 * numpy.abs(oFz).max()
    assert numpy.abs(Fp - oFp).max() <= 1e-13 * max(numpy.abs(oFp).max(), 1e-300)
    want = oracle_potential(pt, pa, R, z, phi, t)
    assert numpy.abs(pot_d - want).max() <= 1e-13 * numpy.abs(want).max()


@pytest.mark.parametrize("which", ["interp", "moving"])
@pytest.mark.parametrize("method,code,bar,factor", [("symplec4_hip", 3, 1e-10, 1e3),
                                                     ("dop853_hip", 6, 1e-8, 10.0)])
def test_full_orbits(which, method, code, bar, factor):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = _interp() if which == "interp" else _moving()
    _, pt, pa, _ = P.parse_pot(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yo = ics(256, seed=113)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code)[0], yo)
    assert_parity(res, ref, bar, sens, factor=factor, label=f"{which} {method}")


@pytest.mark.parametrize("which", ["interp", "moving"])
def test_planar_orbits(which):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = _interp() if which == "interp" else _moving()
    _, pt, pa, _ = P.parse_pot_planar(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yp = planar_ics(256, seed=115)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "dop853_hip")
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 6, planar=True)[0], yp)
    assert_parity(res, ref, 1e-8, sens, factor=10.0, label=f"{which} planar")
