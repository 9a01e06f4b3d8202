"""OblateStaeckelWrapperPotential (-3) on the GPU: the Staeckel form of a
KuzminKutuzovStaeckel potential and of MWPotential2014 (three wrapped
components, so the wrapper's constant is added once), pointwise forces and
potential against the oracle, full and planar orbits against the oracle;
and a CylindricallySeparable wrapper around three components, whose
potential subtracts its Phi(Rp, 0) once."""

import numpy
import pytest

from helpers import (assert_parity, ics, oracle_forces, oracle_integrate, oracle_potential, planar_ics,
                     ulp_sensitivity)

pytestmark = pytest.mark.gpu


def _pot(which):
    from galpy_amd import potential as P

    if which == "kks":
        return [P.OblateStaeckelWrapperPotential(pot=P.KuzminKutuzovStaeckelPotential(amp=1.3, ac=5.0, Delta=0.7),
                                                 delta=0.7, u0=(1.0, 0.2)),
                P.NFWPotential(a=2.0, normalize=0.3)]
    if which == "mw":
        return [P.OblateStaeckelWrapperPotential(pot=P.MWPotential2014_factory(), delta=0.45, u0=(1.2, 0.1), amp=0.9),
                P.PlummerPotential(amp=0.05, b=0.3)]
    return [P.CylindricallySeparablePotentialWrapper(pot=P.MWPotential2014_factory(), Rp=1.1, amp=0.8),
            P.HernquistPotential(amp=0.3, a=0.8)]


@pytest.mark.parametrize("which", ["kks", "mw", "cylsep"])
def test_pointwise_forces_and_potential(which):
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, eval_device

    pot = _pot(which)
    _, pt, pa, _ = P.parse_pot(pot)
    rng = numpy.random.default_rng(31)
    n = 4096
    R, z = rng.uniform(0.05, 4.0, n), rng.uniform(-2.0, 2.0, n)
    phi, t = rng.uniform(0.0, 2.0 * numpy.pi, n), rng.uniform(0.0, 10.0, n)
    dev = [torch.tensor(v, device="cuda") for v in (R, z, phi, t)]
    pot_d, FR, Fz, Fp = (o.cpu().numpy() for o in eval_device(PotentialTable(pot), *dev))
    oFR, oFz, oFp = oracle_forces(pt, pa, R, z, phi, t)
    assert numpy.abs(FR - oFR).max() <= 1e-12 * numpy.abs(oFR).max()
    assert numpy.abs(Fz - oFz).max() <= 1e-12 * numpy.abs(oFz).max()
    assert numpy.abs(Fp - oFp).max() <= 1e-12 * max(numpy.abs(oFR).max(), 1e-300)
    want = oracle_potential(pt, pa, R, z, phi, t)
    assert numpy.abs(pot_d - want).max() <= 1e-13 * numpy.abs(want).max()


@pytest.mark.parametrize("which", ["kks", "mw"])
@pytest.mark.parametrize("method,code,bar,factor", [("symplec4_hip", 3, 1e-10, 1e3),
                                                     ("dop853_hip", 6, 1e-8, 10.0)])
def test_full_orbits(which, method, code, bar, factor):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = _pot(which)
    _, pt, pa, _ = P.parse_pot(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yo = ics(128, seed=33)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code)[0], yo)
    assert_parity(res, ref, bar, sens, factor=factor, label=f"{which} {method}")


@pytest.mark.parametrize("which", ["kks", "mw"])
def test_planar_orbits(which):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = _pot(which)
    _, pt, pa, _ = P.parse_pot_planar(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yp = planar_ics(128, seed=35)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "dop853_hip")
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 6, planar=True)[0], yp)
    assert_parity(res, ref, 1e-8, sens, factor=10.0, label=f"{which} planar")
