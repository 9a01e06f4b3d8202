"""GPU parity of IAS15 (wez_ias15.c, odeint 7) against the oracle's
restatement, pinned by tests/test_oracle_ias15.py.

IAS15 is adaptive, so the DOP853 bar applies: 1e-8 (relative, unit floor,
phi mod 2 pi) at every output.  Step counts are compared in total only
(within 1%): the controller accepts a step only when the next one would not
be shorter, a ratio that sits at 1 to within rounding once the step size has
settled, so which attempts are accepted differs between any two builds
while every accepted step meets the tolerance.  Forces of a step are all
taken at the step's end time (as in the reference), so in a rotating bar the
method is first order in time and the oracle's own result moves by up to
1.7e-3 under a one-ulp change of its initial state: there each orbit is held
to 10x that sensitivity.  Full and planar orbits, MWPotential2014's
specialised kernel and the run-time dispatch kernels, the reversed grid and
a caller's dt shorter than the grid spacing (rows past the run are NaN on
both sides)."""

import numpy
import pytest

from helpers import assert_parity, bar_potential, ics, oracle_integrate, phase_diff, planar_ics, ulp_sensitivity

pytestmark = pytest.mark.gpu


def _packed(pot, planar=False):
    from galpy_amd.potential import parse_pot, parse_pot_planar

    _, pt, pa, _ = (parse_pot_planar if planar else parse_pot)(pot)
    return pt, pa


def test_ias15_mw2014_matches_oracle():
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device
    from galpy_amd.potential import MWPotential2014

    yo = ics(512, seed=11)
    t = numpy.linspace(0.0, 10.0, 101)
    stats = torch.zeros((len(yo), 3), dtype=torch.float64, device="cuda")
    res, err, _ = integrate_device(PotentialTable(MWPotential2014), torch.tensor(yo, device="cuda"),
                                   torch.tensor(t, device="cuda"), "ias15_hip", stats=stats)
    ref, rerr, st, _ = oracle_integrate(*_packed(MWPotential2014), yo, t, 7)
    assert int((err != 0).sum()) == 0 and numpy.all(rerr == 0)
    s = stats.cpu().numpy()
    assert abs(s[:, 0].sum() / st[:, 0].sum() - 1.0) < 0.01
    assert abs((s[:, 0] + s[:, 1]).sum() / (st[:, 0] + st[:, 1]).sum() - 1.0) < 0.01
    assert phase_diff(res.cpu().numpy(), ref) < 1e-8


def test_ias15_generic_planar_and_reversed():
    from galpy_amd.integrate import integrateFullOrbit_hip, integratePlanarOrbit_hip
    from galpy_amd.potential import HernquistPotential, NFWPotential, PlummerPotential

    t = numpy.linspace(0.0, 10.0, 101)
    pot = [HernquistPotential(a=1.3), NFWPotential(a=3.0, amp=2.0), PlummerPotential(b=0.5, amp=0.2)]
    yo = ics(128, seed=12)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, "ias15_hip")
    ref, rerr, _, _ = oracle_integrate(*_packed(pot), yo, t, 7)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    assert phase_diff(res, ref) < 1e-8
    back = numpy.ascontiguousarray(t[::-1])
    res, err = integrateFullOrbit_hip(pot, ref[:, -1].copy(), back, "ias15_hip")
    rb, rerr, _, _ = oracle_integrate(*_packed(pot), ref[:, -1], back, 7)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    assert phase_diff(res, rb) < 1e-8
    yp = planar_ics(128, seed=13)
    res, err = integratePlanarOrbit_hip(bar_potential(), yp.copy(), t, "ias15_hip")
    ppt, ppa = _packed(bar_potential(), True)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(ppt, ppa, y, t, 7, planar=True)[0], yp)
    assert numpy.all(err == 0)
    assert_parity(res, ref, 1e-8, sens, factor=10.0, label="ias15 rotating bar")


def test_ias15_caller_dt_and_frontend():
    from galpy_amd.integrate import integrateFullOrbit_hip
    from galpy_amd.orbit import Orbit
    from galpy_amd.potential import MWPotential2014

    yo = ics(16, seed=14)
    t = numpy.linspace(0.0, 10.0, 101)
    res, err = integrateFullOrbit_hip(MWPotential2014, yo.copy(), t, "ias15_hip", dt=0.05)
    ref, rerr, _, _ = oracle_integrate(*_packed(MWPotential2014), yo, t, 7, dt=0.05)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    assert numpy.array_equal(numpy.isnan(res), numpy.isnan(ref)) and numpy.isnan(ref[:, -1]).all()
    ok = numpy.isfinite(ref[..., 0])
    assert phase_diff(res[ok], ref[ok]) < 1e-8
    o = Orbit(ics(4, seed=15))
    o.integrate(numpy.linspace(0.0, 5.0, 51), MWPotential2014, method="ias15_hip")
    assert o.getOrbit().shape == (4, 51, 6) and numpy.isfinite(o.getOrbit()).all()
