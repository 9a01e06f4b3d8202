"""RotateAndTilt (-8), KuzminLike (-10) and CylindricallySeparable (-12)
wrappers on the GPU, nested with the rotating wrappers: pointwise forces and
potential values against the oracle's, full and planar orbits against the
oracle's, and energy conservation along device rows for a static
composition (through the energy kernel's potential path)."""

import numpy
import pytest

from helpers import (assert_parity, ics, oracle_forces, oracle_integrate, oracle_potential, planar_ics,
                     ulp_sensitivity)

pytestmark = pytest.mark.gpu


def _nested():
    from galpy_amd import potential as P

    halo = P.LogarithmicHaloPotential(normalize=0.6, q=0.8, b=0.7)
    inner = [P.RotateAndTiltWrapperPotential(pot=halo, zvec=[0.2, 0.1, 1.0], galaxy_pa=0.3, offset=[0.01, 0.0, -0.02]),
             P.KuzminLikeWrapperPotential(pot=P.HernquistPotential(normalize=0.3), a=0.6, b=0.2)]
    return [P.SolidBodyRotationWrapperPotential(pot=inner, omega=0.5, pa=0.2),
            P.CylindricallySeparablePotentialWrapper(pot=P.NFWPotential(normalize=0.2), Rp=1.1)]


def _static():
    from galpy_amd import potential as P

    halo = P.LogarithmicHaloPotential(normalize=0.6, q=0.8, b=0.7)
    return [P.RotateAndTiltWrapperPotential(pot=halo, inclination=0.7, sky_pa=0.2, galaxy_pa=0.4),
            P.KuzminLikeWrapperPotential(pot=P.HernquistPotential(normalize=0.2), a=0.6, b=0.1),
            P.CylindricallySeparablePotentialWrapper(pot=P.MiyamotoNagaiPotential(a=0.5, b=0.1, normalize=0.2),
                                                     Rp=1.0)]


@pytest.mark.parametrize("which", ["nested", "static"])
def test_pointwise_forces_and_potential(which):
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, eval_device

    pot = _nested() if which == "nested" else _static()
    _, pt, pa, _ = P.parse_pot(pot)
    rng = numpy.random.default_rng(91)
    n = 4096
    R, z = rng.uniform(0.1, 3.0, n), rng.uniform(-1.0, 1.0, n)
    phi, t = rng.uniform(0.0, 2.0 * numpy.pi, n), rng.uniform(-3.0, 3.0, n)
    dev = [torch.tensor(v, device="cuda") for v in (R, z, phi, t)]
    pot_d, FR, Fz, Fp = (o.cpu().numpy() for o in eval_device(PotentialTable(pot), *dev))
    oFR, oFz, oFp = oracle_forces(pt, pa, R, z, phi, t)
    assert numpy.abs(FR - oFR).max() <= 1e-13 * numpy.abs(oFR).max()
    assert numpy.abs(Fz - oFz).max() <= 1e-13 * numpy.abs(oFz).max()
    assert numpy.abs(Fp - oFp).max() <= 1e-13 * numpy.abs(oFp).max()
    want = oracle_potential(pt, pa, R, z, phi, t)
    assert numpy.abs(pot_d - want).max() <= 1e-13 * numpy.abs(want).max()


@pytest.mark.parametrize("method,code,bar,factor", [("symplec4_hip", 3, 1e-10, 1e3),
                                                     ("dop853_hip", 6, 1e-8, 10.0)])
def test_full_orbits(method, code, bar, factor):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = _nested()
    _, pt, pa, _ = P.parse_pot(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yo = ics(256, seed=93)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code)[0], yo)
    assert_parity(res, ref, bar, sens, factor=factor, label=method)


def test_planar_orbits():
    from galpy_amd import potential as P
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = _static()[1:]  # no planar RotateAndTilt in galpy
    _, pt, pa, _ = P.parse_pot_planar(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yp = planar_ics(256, seed=95)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "dop853_hip")
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 6, planar=True)[0], yp)
    assert_parity(res, ref, 1e-8, sens, factor=10.0, label="planar dop853")


def test_energy_conserved_static():
    """E along device rows from the energy kernel (the mapped wrappers'
    potential path) equals E from the host classes on the same rows, and is
    conserved (median drift; orbits that graze the cored halo's centre
    drift more under DOP853's error control, for the oracle too)."""
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, energy_drift_device, integrateFullOrbit_hip

    pot = _static()
    t = numpy.linspace(0.0, 50.0, 501)
    res, err = integrateFullOrbit_hip(pot, ics(256, seed=97).copy(), t, "dop853_hip")
    assert numpy.all(err == 0)
    e0, de = (v.cpu().numpy() for v in energy_drift_device(PotentialTable(pot), torch.tensor(res, device="cuda"),
                                                           torch.tensor(t, device="cuda")))
    R, vR, vT, z, vz, phi = (res[..., k] for k in range(6))
    E = P.evaluatePotentials(pot, R, z, phi) + 0.5 * (vR**2 + vT**2 + vz**2)
    d_np = numpy.abs((E - E[:, :1]) / E[:, :1]).max(axis=1)
    assert numpy.abs(e0 - E[:, 0]).max() <= 1e-13 * numpy.abs(E[:, 0]).max()
    # drifts are relative to E(0), which comes close to 0 for some orbits
    # here: compare them as energies
    assert (numpy.abs(de - d_np) * numpy.abs(E[:, 0])).max() <= 1e-13
    assert numpy.median(d_np) < 1e-10
