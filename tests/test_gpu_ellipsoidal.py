"""GPU parity of the EllipsoidalPotential family (TriaxialHernquist,
TriaxialNFW, TriaxialJaffe, TwoPowerTriaxial, PerfectEllipsoid,
TriaxialGaussian, PowerTriaxial; pot_types 21, 22, 23, 43, 30, 37, 38)
against the oracle, whose forces and potentials are pinned to mpmath
(spherical limit) and to the gradient of the potential (triaxial and
rotated frames).

Bars: device forces 1e-13 relative to the largest force of a case (the
device multiplies by stored reciprocals where galpy divides, so single
terms differ in the last bits); potential values the same.  Orbits as for
the other potentials: fixed-step symplec4 1e-10, DOP853 1e-8 at galpy's
default tolerances, held to the oracle's own one-ulp sensitivity times 1e3
(fixed steps) or 10 (DOP853) where that exceeds the bar -- triaxial and
rotating frames make some orbits chaotic."""

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


def _members(**kw):
    from galpy_amd import potential as P

    return {
        "TriaxialHernquist": P.TriaxialHernquistPotential(amp=1.7, a=1.3, **kw),
        "TriaxialNFW": P.TriaxialNFWPotential(amp=2.1, a=1.8, **kw),
        "TriaxialJaffe": P.TriaxialJaffePotential(amp=1.4, a=0.9, **kw),
        "TwoPowerTriaxial": P.TwoPowerTriaxialPotential(amp=1.9, a=1.6, alpha=1.2, beta=3.7, **kw),
        "PerfectEllipsoid": P.PerfectEllipsoidPotential(amp=1.3, a=1.5, **kw),
        "TriaxialGaussian": P.TriaxialGaussianPotential(amp=2.3, sigma=1.4, **kw),
        "PowerTriaxial": P.PowerTriaxialPotential(amp=1.1, alpha=1.3, r1=1.2, **kw),
    }


FRAMES = {"aligned": {"b": 0.8, "c": 0.55}, "pa": {"b": 0.8, "c": 0.55, "pa": 0.4},
          "zvec+pa": {"b": 0.9, "c": 0.6, "pa": -0.7, "zvec": [0.3, -0.2, 1.0]},
          "spherical": {}}


def _packed(pot, planar=False):
    from galpy_amd.potential import parse_pot, parse_pot_planar

    _, pt, pa, _ = (parse_pot_planar if planar else parse_pot)(pot)
    return pt, pa


@pytest.mark.parametrize("frame", list(FRAMES))
def test_device_forces_and_potential_match_oracle(frame):
    import torch

    from galpy_amd.adjoint import _RawTable
    from galpy_amd.integrate import eval_device

    rng = numpy.random.default_rng(21)
    R = rng.uniform(0.05, 6.0, 2000)
    z = rng.normal(0.0, 1.0, 2000)
    phi = rng.uniform(-numpy.pi, numpy.pi, 2000)
    t = numpy.zeros_like(R)
    for name, p in _members(**FRAMES[frame]).items():
        pt, pa = _packed([p])
        tab = _RawTable(pt, pa)
        dev = [torch.tensor(v, device="cuda") for v in (R, z, phi, t)]
        pot, FR, Fz, Fphi = (o.cpu().numpy() for o in eval_device(tab, *dev))
        oFR, oFz, oFphi = oracle_forces(pt, pa, R, z, phi)
        scale = max(numpy.abs(oFR).max(), numpy.abs(oFz).max())
        for got, want, key in ((FR, oFR, "FR"), (Fz, oFz, "Fz"), (Fphi, oFphi, "Fphi")):
            assert numpy.abs(got - want).max() <= 1e-13 * scale, (name, frame, key)
        if name != "TwoPowerTriaxial":  # psi needs 2F1: no device potential
            opot = oracle_potential(pt, pa, R, z, phi)
            assert numpy.abs(pot - opot).max() <= 1e-13 * numpy.abs(opot).max(), (name, frame)
        else:
            assert numpy.all(numpy.isnan(pot))
        del tab


def _halo_disk(frame):
    from galpy_amd import potential as P

    kw = FRAMES[frame]
    return [P.MiyamotoNagaiPotential(a=0.5, b=0.05, normalize=0.4),
            P.TriaxialNFWPotential(a=2.0, normalize=0.6, **kw)]


@pytest.mark.parametrize("frame", ["aligned", "zvec+pa"])
@pytest.mark.parametrize("method,code,dt", [("symplec4_hip", 3, 0.01), ("dop853_hip", 6, None)])
def test_full_orbits_halo_and_disk(frame, method, code, dt):
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = _halo_disk(frame)
    pt, pa = _packed(pot)
    yo = ics(128, seed=22)
    t = numpy.linspace(0.0, 8.0, 81)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method, dt=dt)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code, dt=dt)[0], yo)
    assert numpy.all(err == 0)
    assert_parity(res, ref, 1e-10 if code == 3 else 1e-8, sens, 1e3 if code == 3 else 10.0,
                  label=f"{frame} {method}")


def test_each_member_alone_symplec4():
    from galpy_amd.integrate import integrateFullOrbit_hip

    t = numpy.linspace(0.0, 4.0, 41)
    yo = ics(64, seed=23)
    for name, p in _members(**FRAMES["pa"]).items():
        pt, pa = _packed([p])
        res, err = integrateFullOrbit_hip([p], yo.copy(), t, "symplec4_hip", dt=0.01)
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 3, dt=0.01)[0], yo)
        assert numpy.all(err == 0), name
        assert_parity(res, ref, 1e-10, sens, 1e3, label=name)


@pytest.mark.parametrize("method,code", [("symplec4_hip", 3), ("dop853_hip", 6)])
def test_planar_orbits(method, code):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = [P.PerfectEllipsoidPotential(a=1.2, b=0.7, c=0.5, pa=0.3, normalize=0.5),
           P.TriaxialHernquistPotential(a=1.5, b=0.85, c=0.6, normalize=0.5)]
    pt, pa = _packed(pot, True)
    yo = planar_ics(128, seed=24)
    t = numpy.linspace(0.0, 8.0, 81)
    dt = 0.01 if code == 3 else None
    res, err = integratePlanarOrbit_hip(pot, yo.copy(), t, method, dt=dt)
    ref, sens = ulp_sensitivity(
        lambda y: oracle_integrate(pt, pa, y, t, code, dt=dt, planar=True)[0], yo)
    assert numpy.all(err == 0)
    assert_parity(res, ref, 1e-10 if code == 3 else 1e-8, sens, 1e3 if code == 3 else 10.0,
                  label=f"planar {method}")


def test_energy_is_conserved_in_the_aligned_halo():
    """A static triaxial halo conserves E: the device energy kernel (its own
    potential quadrature) sees DOP853 orbits keep E to ~1e-9."""
    import torch

    from galpy_amd.integrate import PotentialTable, energy_drift_device, integrateFullOrbit_hip

    pot = _halo_disk("aligned")
    yo = ics(64, seed=25)
    t = numpy.linspace(0.0, 8.0, 81)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, "dop853_hip", rtol=1e-12, atol=1e-12)
    assert numpy.all(err == 0)
    rows = torch.tensor(res, device="cuda")
    e0, de = energy_drift_device(PotentialTable(pot), rows, torch.tensor(t, device="cuda"))
    assert float(de.abs().max()) < 1e-9 * float(e0.abs().max())


def test_refused_for_variational_equations():
    from galpy_amd.integrate import integrateFullOrbit_dxdv_hip

    pot = _halo_disk("aligned")
    w0 = numpy.array([[1.0, 0.1, 0.2, 0.1, 1.0, 0.05]])
    with pytest.raises(Exception):
        integrateFullOrbit_dxdv_hip(pot, w0, numpy.ones_like(w0), numpy.linspace(0, 1, 3), "dop853_hip")
