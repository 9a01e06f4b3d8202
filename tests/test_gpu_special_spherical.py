"""GPU parity of EinastoPotential (41), TwoPowerSphericalPotential (42) and
ExpTruncNFWPotential (46) against the oracle, whose forces, potentials and
special functions (incomplete gamma, E1, 2F1(A, B; A+1; -x)) are pinned to
mpmath and to the reference's own hyp2f1 build.

Bars: device forces and potentials 1e-13 relative to the largest value of a
case (the device sums the incomplete-gamma series and the 2F1 tables in its
own order); orbits fixed-step symplec4 1e-10 and DOP853 1e-8, held to the
oracle's own one-ulp sensitivity times 1e3 / 10 where that exceeds the bar."""

import numpy
import pytest

from helpers import (
    assert_parity,
    ics,
    oracle_forces,
    oracle_integrate,
    oracle_potential,
    planar_ics,
    ulp_sensitivity,
)

pytestmark = pytest.mark.gpu


def _singles():
    from galpy_amd import potential as P

    return {
        "Einasto n=2.5": P.EinastoPotential(amp=1.3, h=1.7, n=2.5),
        "Einasto n=0.5": P.EinastoPotential(amp=1.1, h=0.9, n=0.5),
        "Einasto n=6": P.EinastoPotential(amp=0.8, h=0.05, n=6.0),
        "ExpTruncNFW": P.ExpTruncNFWPotential(amp=1.3, a=1.2, rc=3.0),
        "TwoPowerSpherical": P.TwoPowerSphericalPotential(amp=1.3, a=1.2, alpha=1.3, beta=3.6),
        "TwoPowerSpherical beta=3": P.TwoPowerSphericalPotential(amp=1.3, a=2.0, alpha=0.5, beta=3.0),
        "TwoPowerSpherical beta=7.5": P.TwoPowerSphericalPotential(amp=1.3, a=0.7, alpha=1.2, beta=7.5),
    }


def _packed(pot, planar=False):
    from galpy_amd.potential import parse_pot, parse_pot_planar

    _, pt, pa, _ = (parse_pot_planar if planar else parse_pot)(pot)
    return pt, pa


def test_device_forces_and_potential_match_oracle():
    import torch

    from galpy_amd.adjoint import _RawTable
    from galpy_amd.integrate import eval_device

    rng = numpy.random.default_rng(31)
    R = numpy.concatenate([rng.uniform(0.001, 8.0, 2000), [1e-4, 2.99, 3.0, 3.01, 50.0, 400.0]])
    z = numpy.concatenate([rng.normal(0.0, 1.0, 2000), [0.0, 0.1, 0.0, -0.2, 3.0, 10.0]])
    phi = numpy.zeros_like(R)
    for name, p in _singles().items():
        pt, pa = _packed([p])
        tab = _RawTable(pt, pa)
        dev = [torch.tensor(v, device="cuda") for v in (R, z, phi, phi)]
        pot, FR, Fz, Fphi = (o.cpu().numpy() for o in eval_device(tab, *dev))
        oFR, oFz, _ = oracle_forces(pt, pa, R, z)
        rel = numpy.abs(FR - oFR) / numpy.abs(oFR)
        assert rel.max() <= 1e-13, (name, rel.max(), R[rel.argmax()])
        assert numpy.abs(Fz - oFz).max() <= 1e-13 * numpy.abs(oFz).max(), name
        assert numpy.all(Fphi == 0.0)
        if name.startswith("TwoPowerSpherical"):
            assert numpy.all(numpy.isnan(pot))
        else:
            opot = oracle_potential(pt, pa, R, z)
            assert numpy.abs(pot - opot).max() <= 1e-13 * numpy.abs(opot).max(), name
        del tab


def _with_disk(p):
    from galpy_amd import potential as P

    return [P.MiyamotoNagaiPotential(a=0.5, b=0.05, normalize=0.4), p]


@pytest.mark.parametrize("method,code,dt", [("symplec4_hip", 3, 0.01), ("dop853_hip", 6, None)])
def test_full_orbits_each_with_a_disk(method, code, dt):
    from galpy_amd.integrate import integrateFullOrbit_hip
    from galpy_amd.potential import EinastoPotential, ExpTruncNFWPotential, TwoPowerSphericalPotential

    yo = ics(128, seed=32)
    t = numpy.linspace(0.0, 8.0, 81)
    for halo in (EinastoPotential(h=1.5, n=3.0, normalize=0.6), ExpTruncNFWPotential(a=2.0, rc=5.0, normalize=0.6),
                 TwoPowerSphericalPotential(a=3.0, alpha=1.2, beta=3.4, normalize=0.6)):
        pot = _with_disk(halo)
        pt, pa = _packed(pot)
        res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method, dt=dt)
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code, dt=dt)[0], yo)
        assert numpy.all(err == 0)
        assert_parity(res, ref, 1e-10 if code == 3 else 1e-8, sens, 1e3 if code == 3 else 10.0,
                      label=f"{type(halo).__name__} {method}")


def test_planar_orbits():
    from galpy_amd.integrate import integratePlanarOrbit_hip
    from galpy_amd.potential import EinastoPotential, TwoPowerSphericalPotential

    pot = [EinastoPotential(h=1.5, n=3.0, normalize=0.5),
           TwoPowerSphericalPotential(a=3.0, alpha=0.8, beta=3.0, normalize=0.5)]
    pt, pa = _packed(pot, True)
    yo = planar_ics(128, seed=33)
    t = numpy.linspace(0.0, 8.0, 81)
    for method, code, dt in (("symplec4_hip", 3, 0.01), ("dop853_hip", 6, None)):
        res, err = integratePlanarOrbit_hip(pot, yo.copy(), t, method, dt=dt)
        ref, sens = ulp_sensitivity(
            lambda y: oracle_integrate(pt, pa, y, t, code, dt=dt, planar=True)[0], yo)
        assert numpy.all(err == 0)
        assert_parity(res, ref, 1e-10 if code == 3 else 1e-8, sens, 1e3 if code == 3 else 10.0,
                      label=f"planar {method}")


def test_energy_is_conserved():
    import torch

    from galpy_amd.integrate import PotentialTable, energy_drift_device, integrateFullOrbit_hip
    from galpy_amd.potential import EinastoPotential, ExpTruncNFWPotential

    for halo in (EinastoPotential(h=1.5, n=3.0, normalize=0.6), ExpTruncNFWPotential(a=2.0, rc=5.0, normalize=0.6)):
        pot = _with_disk(halo)
        yo = ics(64, seed=34)
        t = numpy.linspace(0.0, 8.0, 81)
        res, err = integrateFullOrbit_hip(pot, yo.copy(), t, "dop853_hip", rtol=1e-12, atol=1e-12)
        assert numpy.all(err == 0)
        rows = torch.tensor(res, device="cuda")
        e0, de = energy_drift_device(PotentialTable(pot), rows, torch.tensor(t, device="cuda"))
        assert float(de.abs().max()) < 1e-9, type(halo).__name__
