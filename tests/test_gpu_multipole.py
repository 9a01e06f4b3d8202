"""MultipoleExpansionPotential (pot_type 44) on the GPU, static (axisymmetric
and non-axisymmetric) and time-dependent, and a DiskMultipoleExpansion-
Potential (44 plus its pot_type 26 pairs): pointwise forces and potential
against the oracle (inside, below and above the radial grid), full and
planar orbits against the oracle."""

import numpy
import pytest

from helpers import (assert_parity, ics, orbit_diffs, oracle_forces, oracle_integrate, oracle_potential,
                     planar_ics, ulp_sensitivity)

pytestmark = pytest.mark.gpu


def _bar_density(R, z, phi):
    """A flattened halo with an m = 2 bar-like and an m = 1 lopsided term."""
    r = numpy.sqrt(R * R + z * z)
    return (numpy.exp(-r) / (0.2 + r) * (1.0 + 0.5 * (R / (r + 1e-300)) ** 2 * numpy.cos(2.0 * phi) * numpy.exp(-r))
            * numpy.exp(-0.5 * (z / (0.6 + r)) ** 2) + 0.05 * numpy.exp(-r) * R * numpy.sin(phi))


def _pot(which):
    from galpy_amd import potential as P

    rg = numpy.geomspace(0.01, 40.0, 241)
    if which == "axi":
        halo = P.MultipoleExpansionPotential.from_density(
            lambda R, z: numpy.exp(-numpy.hypot(R, z / 0.7)) / (0.1 + numpy.hypot(R, z / 0.7)), L=6, rgrid=rg,
            amp=3.0)
        return [halo, P.MiyamotoNagaiPotential(a=0.5, b=0.1, normalize=0.3)]
    if which == "nonaxi":
        return [P.MultipoleExpansionPotential.from_density(_bar_density, L=5, rgrid=rg, amp=4.0),
                P.NFWPotential(a=2.0, normalize=0.4)]
    if which == "tdep":
        dens = lambda R, z, phi, t=0.0: _bar_density(R, z, phi - 0.5 * t) * (1.0 + 0.1 * numpy.sin(t))
        return [P.MultipoleExpansionPotential.from_density(dens, L=3, rgrid=numpy.geomspace(0.01, 40.0, 121),
                                                           tgrid=numpy.linspace(-1.0, 12.0, 14), amp=4.0),
                P.NFWPotential(a=2.0, normalize=0.4)]
    dens = lambda R, z: 13.5 * numpy.exp(-3.0 * R) * numpy.exp(-27.0 * numpy.fabs(z))
    return [P.DiskMultipoleExpansionPotential(dens=dens, Sigma={"type": "exp", "h": 1.0 / 3.0, "amp": 1.0},
                                              hz={"type": "exp", "h": 1.0 / 27.0}, L=6,
                                              rgrid=numpy.geomspace(0.01, 20.0, 201)),
            P.NFWPotential(a=2.0, normalize=0.4)]


@pytest.mark.parametrize("which", ["axi", "nonaxi", "tdep", "disk"])
def test_pointwise_forces_and_potential(which):
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, eval_device

    pot = _pot(which)
    _, pt, pa, _ = P.parse_pot(pot)
    rng = numpy.random.default_rng(441)
    n = 4096
    r = numpy.concatenate([rng.uniform(0.02, 5.0, n - 64), rng.uniform(0.001, 0.01, 32), rng.uniform(40.0, 90.0, 32)])
    th = rng.uniform(0.05, numpy.pi - 0.05, n)
    R, z = r * numpy.sin(th), r * numpy.cos(th)
    phi, t = rng.uniform(0.0, 2.0 * numpy.pi, n), rng.uniform(-2.0, 13.0, n)
    dev = [torch.tensor(v, device="cuda") for v in (R, z, phi, t)]
    pot_d, FR, Fz, Fp = (o.cpu().numpy() for o in eval_device(PotentialTable(pot), *dev))
    oFR, oFz, oFp = oracle_forces(pt, pa, R, z, phi, t)
    assert numpy.abs(FR - oFR).max() <= 1e-13 * numpy.abs(oFR).max()
    assert numpy.abs(Fz - oFz).max() <= 1e-13 * numpy.abs(oFz).max()
    assert numpy.abs(Fp - oFp).max() <= 1e-13 * max(numpy.abs(oFp).max(), 1e-300)
    want = oracle_potential(pt, pa, R, z, phi, t)
    assert numpy.abs(pot_d - want).max() <= 1e-13 * numpy.abs(want).max()


@pytest.mark.parametrize("which", ["axi", "nonaxi", "tdep", "disk"])
@pytest.mark.parametrize("method,code,bar,factor", [("symplec4_hip", 3, 1e-10, 1e3),
                                                     ("dop853_hip", 6, 1e-8, 10.0)])
def test_full_orbits(which, method, code, bar, factor):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = _pot(which)
    _, pt, pa, _ = P.parse_pot(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yo = ics(128, seed=443)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code)[0], yo)
    if code == 6:
        # the radial quintics are twice continuously differentiable only:
        # DOP853's error estimate does not bound a step across a node, so the
        # oracle's fused-multiply-add twin sets the yardstick (as for the
        # B-spline interpolated potential), with 1e-7 as the bar
        fma = oracle_integrate(pt, pa, yo, t, code, fma=True)[0]
        sens = numpy.maximum(sens, orbit_diffs(fma, ref))
        bar = 1e-7
    assert_parity(res, ref, bar, sens, factor=factor, label=f"{which} {method}")


@pytest.mark.parametrize("which", ["nonaxi", "disk"])
def test_planar_orbits(which):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = _pot(which)
    _, pt, pa, _ = P.parse_pot_planar(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yp = planar_ics(128, seed=445)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "symplec4_hip")
    assert numpy.all(err == 0)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 3, planar=True)[0], yp)
    assert_parity(res, ref, 1e-10, sens, factor=1e3, label="planar")
