"""GPU parity of RK4, RK6 (bovy_rk.c) and symplec6 (bovy_symplecticode.c:403-1000)
against the oracle's restatements, pinned by tests/test_oracle_rk_symplec6.py.

Bars as for the other fixed-step methods: 1e-10 (relative, unit floor, phi
mod 2 pi) at every output with the oracle's per-orbit step, the estimated
step identical (the RK estimator's ceil(err^(1/5)) / ceil(err^(1/7)) and the
symplectic halving alike); full and planar orbits, MWPotential2014's
specialised kernels and the run-time dispatch kernels."""

import numpy
import pytest

from helpers import bar_potential, ics, oracle_integrate, phase_diff, planar_ics

pytestmark = pytest.mark.gpu

METHODS = [("rk4_hip", 1), ("rk6_hip", 2), ("symplec6_hip", 4)]


def _packed(pot, planar=False):
    from galpy_amd.potential import parse_pot, parse_pot_planar

    _, pt, pa, _ = (parse_pot_planar if planar else parse_pot)(pot)
    return pt, pa


def _mw():
    from galpy_amd.potential import MWPotential2014

    return MWPotential2014


@pytest.mark.parametrize("method,code", METHODS)
def test_estimated_step_mw2014(method, code):
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device

    pot = _mw()
    yo = ics(512, seed=1)
    t = numpy.linspace(0.0, 10.0, 101)
    tab = PotentialTable(pot)
    stats = torch.zeros((len(yo), 3), dtype=torch.float64, device="cuda")
    res, err, _ = integrate_device(tab, torch.tensor(yo, device="cuda"), torch.tensor(t, device="cuda"),
                                   method, stats=stats)
    ref, rerr, st, _ = oracle_integrate(*_packed(pot), yo, t, code)
    assert int((err != 0).sum()) == 0 and numpy.all(rerr == 0)
    s = stats.cpu().numpy()
    assert numpy.array_equal(s[:, 2], st[:, 3]) and numpy.array_equal(s[:, 0], st[:, 0])
    assert phase_diff(res.cpu().numpy(), ref) < 1e-10


@pytest.mark.parametrize("method,code", METHODS)
def test_fixed_step_generic_and_planar(method, code):
    from galpy_amd.integrate import integrateFullOrbit_hip, integratePlanarOrbit_hip
    from galpy_amd.potential import HernquistPotential, NFWPotential, PlummerPotential

    t = numpy.linspace(0.0, 10.0, 101)
    for pot in (bar_potential(), [HernquistPotential(a=1.3), NFWPotential(a=3.0, amp=2.0),
                                  PlummerPotential(b=0.5, amp=0.2)]):
        yo = ics(128, seed=2)
        res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method, dt=0.01)
        ref, rerr, _, _ = oracle_integrate(*_packed(pot), yo, t, code, dt=0.01)
        assert numpy.all(err == 0) and numpy.all(rerr == 0)
        assert phase_diff(res, ref) < 1e-10
        yp = planar_ics(128, seed=3)
        res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, method)  # estimated step
        ref, rerr, _, _ = oracle_integrate(*_packed(pot, True), yp, t, code, planar=True)
        assert numpy.all(err == 0) and numpy.all(rerr == 0)
        assert phase_diff(res, ref) < 1e-10


def test_orbit_frontend_methods():
    from galpy_amd.orbit import Orbit

    o = Orbit(ics(4, seed=4))
    t = numpy.linspace(0.0, 5.0, 51)
    for method in ("rk4_hip", "rk6_hip", "symplec6_hip"):
        o.integrate(t, _mw(), method=method, dt=0.01)
        assert o.getOrbit().shape == (4, 51, 6)


def test_no_integrate_dxdv_for_ias15():
    # galpy has no ias15 integrate_dxdv (Orbit.integrate_dxdv refuses it);
    # every other integrator has one (tests/test_gpu_dxdv.py)
    from galpy_amd import _lib
    from galpy_amd.integrate import integrateFullOrbit_dxdv_hip
    from helpers import cyl_to_rect

    y0 = cyl_to_rect(ics(2, seed=5))
    with pytest.raises(RuntimeError):
        integrateFullOrbit_dxdv_hip(_mw(), y0, numpy.ones_like(y0), numpy.linspace(0, 1, 3), "ias15_hip")
    assert "ias15" in _lib.last_error()


def test_dopr54_matches_oracle():
    """DOPR54 (odeint 5): adaptive, so the DOP853 bar, 1e-8 at every output,
    with the oracle's accepted and rejected step counts."""
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device, integratePlanarOrbit_hip

    pot = _mw()
    yo = ics(512, seed=8)
    t = numpy.linspace(0.0, 10.0, 101)
    stats = torch.zeros((len(yo), 3), dtype=torch.float64, device="cuda")
    res, err, _ = integrate_device(PotentialTable(pot), torch.tensor(yo, device="cuda"),
                                   torch.tensor(t, device="cuda"), "dopr54_hip", stats=stats)
    ref, rerr, st, _ = oracle_integrate(*_packed(pot), yo, t, 5)
    assert int((err != 0).sum()) == 0 and numpy.all(rerr == 0)
    s = stats.cpu().numpy()
    same = (s[:, 0] == st[:, 0]) & (s[:, 1] == st[:, 1])
    assert same.mean() > 0.99
    assert phase_diff(res.cpu().numpy(), ref) < 1e-8
    yp = planar_ics(128, seed=9)
    res, err = integratePlanarOrbit_hip(bar_potential(), yp.copy(), t, "dopr54_hip")
    ref, rerr, _, _ = oracle_integrate(*_packed(bar_potential(), True), yp, t, 5, planar=True)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    assert phase_diff(res, ref) < 1e-8
