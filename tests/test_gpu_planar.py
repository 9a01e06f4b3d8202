"""GPU parity of planar orbits (integratePlanarOrbit / integratePlanarOrbit_hip)
against the oracle's integratePlanarOrbit on the same seeded inputs.

Bars, as for full orbits: fixed-step methods match to 1e-10 (relative, unit
floor, phi mod 2 pi) with the identical per-orbit step; DOP853 to 1e-8 at
galpy's default tolerances (rtol = atol = 1e-12).  At sizes the oracle does
not reach in seconds, the fixed-step planar kernels are checked against the
GPU's own full-orbit kernels in the plane z = 0 (the same orbit, to
rounding), and DOP853 against the oracle on a sample of the orbits plus the
energy drift of all of them."""

import numpy
import pytest

from helpers import bar_potential, ics, in_plane, oracle_integrate, phase_diff, planar_ics

pytestmark = pytest.mark.gpu


def _mw():
    from galpy_amd.potential import MWPotential2014

    return MWPotential2014


def _packed(pot):
    from galpy_amd.potential import parse_pot_planar

    _, pt, pa, _ = parse_pot_planar(pot)
    return pt, pa


def _hip(pot, yo, t, method, dt=None):
    from galpy_amd.integrate import integratePlanarOrbit_hip

    return integratePlanarOrbit_hip(pot, yo.copy(), t, method, dt=dt)


def _oracle(pot, yo, t, code, dt=None):
    return oracle_integrate(*_packed(pot), yo, t, code, dt=dt, planar=True)


@pytest.mark.parametrize("method,code", [("symplec4_hip", 3), ("leapfrog_hip", 0)])
def test_planar_symplectic_estimated_dt(method, code):
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device

    pot = _mw()
    yo = planar_ics(256, seed=1)
    t = numpy.linspace(0.0, 10.0, 1001)
    res, err = _hip(pot, yo, t, method)
    ref, rerr, rstats, _ = _oracle(pot, yo, t, code)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    assert res.shape == (256, 1001, 4)
    assert phase_diff(res, ref) < 1e-10
    # the device twin reports the step it chose: the oracle's, orbit by orbit
    table = PotentialTable(pot, planar=True)
    stats = torch.zeros((len(yo), 3), dtype=torch.float64, device="cuda")
    dres, derr, stats = integrate_device(table, torch.tensor(yo, device="cuda"),
                                         torch.tensor(t, device="cuda"), method, stats=stats)
    torch.cuda.synchronize()
    s = stats.cpu().numpy()
    assert numpy.array_equal(s[:, 2], rstats[:, 3]) and numpy.array_equal(s[:, 0], rstats[:, 0])
    assert torch.equal(dres.cpu(), torch.tensor(res))


@pytest.mark.parametrize("method,code", [("symplec4_hip", 3), ("leapfrog_hip", 0)])
def test_planar_symplectic_fixed_dt_bar(method, code):
    pot = bar_potential()
    yo = planar_ics(128, seed=5)
    t = numpy.linspace(0.0, 10.0, 101)
    res, err = _hip(pot, yo, t, method, dt=0.01)
    ref, *_ = _oracle(pot, yo, t, code, dt=0.01)
    assert numpy.all(err == 0)
    assert phase_diff(res, ref) < 1e-10


@pytest.mark.parametrize("which", ["mw", "bar"])
def test_planar_dop853(which):
    pot = _mw() if which == "mw" else bar_potential()
    yo = planar_ics(128, seed=3)
    t = numpy.linspace(0.0, 20.0, 201)
    res, err = _hip(pot, yo, t, "dop853_hip")
    ref, rerr, _, _ = _oracle(pot, yo, t, 6)
    assert numpy.all(err == 0)
    assert phase_diff(res, ref) < 1e-8


def test_planar_other_potentials():
    from galpy_amd.potential import HernquistPotential, LogarithmicHaloPotential, NFWPotential

    for pot in ([HernquistPotential(a=1.3)], [LogarithmicHaloPotential(q=0.8)],
                [LogarithmicHaloPotential(q=0.9, b=0.8)],
                [HernquistPotential(a=1.3), NFWPotential(a=3.0, amp=2.0)]):  # generic kernel
        yo = planar_ics(64, seed=6)
        t = numpy.linspace(0.0, 10.0, 101)
        res, err = _hip(pot, yo, t, "symplec4_hip", dt=0.01)
        ref, *_ = _oracle(pot, yo, t, 3, dt=0.01)
        assert phase_diff(res, ref) < 1e-10
        res, err = _hip(pot, yo, t, "dop853_hip")
        ref, *_ = _oracle(pot, yo, t, 6)
        assert phase_diff(res, ref) < 1e-8


def test_planar_indiv_t_backward_and_edge_sizes():
    pot = _mw()
    yo = planar_ics(64, seed=7)
    rng = numpy.random.default_rng(0)
    t = numpy.stack([numpy.linspace(0.0, L, 51) for L in rng.uniform(1.0, 5.0, len(yo))])
    tb = numpy.linspace(0.0, -5.0, 51)
    for tt in (t, tb):
        for method, code in (("symplec4_hip", 3), ("dop853_hip", 6)):
            res, err = _hip(pot, yo, tt, method)
            ref, *_ = _oracle(pot, yo, tt, code)
            assert phase_diff(res, ref) < (1e-10 if code == 3 else 1e-8)
    for n, tt in ((1, numpy.array([0.0, 0.5])), (65, numpy.linspace(0, 1, 3))):
        y = planar_ics(n, seed=9)
        for method, code in (("symplec4_hip", 3), ("leapfrog_hip", 0), ("dop853_hip", 6)):
            res, err = _hip(pot, y, tt, method)
            ref, *_ = _oracle(pot, y, tt, code)
            assert res.shape == (n, len(tt), 4)
            assert phase_diff(res, ref) < (1e-8 if code == 6 else 1e-10)
    from galpy_amd.integrate import integratePlanarOrbit_hip

    res, err = integratePlanarOrbit_hip(pot, numpy.zeros((0, 4)), numpy.linspace(0, 1, 5), "symplec4_hip")
    assert res.shape == (0, 5, 4)


def test_planar_dop853_lockstep_and_order():
    """Every lane passing an output time in the same step (overflowing the
    planar dense-output queue), and a call large enough for the pilot and
    longest-first order: results identical to the unordered, unqueued ones."""
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device

    pot = _mw()
    yo = numpy.repeat(planar_ics(1, seed=12), 130, axis=0)
    t = numpy.linspace(0.0, 20.0, 201)
    res, err = _hip(pot, yo, t, "dop853_hip")
    assert numpy.all(err == 0)
    assert numpy.array_equal(res, numpy.broadcast_to(res[:1], res.shape))
    ref, *_ = _oracle(pot, yo[:1], t, 6)
    assert phase_diff(res[:1], ref) < 1e-8
    n, split = 70000, 60000
    table = PotentialTable(pot, planar=True)
    td = torch.linspace(0.0, 3.0, 4, dtype=torch.float64, device="cuda")
    y = torch.tensor(planar_ics(n, seed=21), device="cuda")
    whole, err, _ = integrate_device(table, y.clone(), td, "dop853_hip")
    parts = [integrate_device(table, y[a:b].clone(), td, "dop853_hip")[0] for a, b in ((0, split), (split, n))]
    torch.cuda.synchronize()
    assert int((err != 0).sum()) == 0
    assert torch.equal(whole, torch.cat(parts))


def test_planar_kernels_at_scale():
    """262,144 orbits.  symplec4 at a fixed step: the planar kernels against
    the full-orbit kernels run on the same orbits in the plane (z = vz = 0),
    both on the GPU -- the same orbit, to rounding.  DOP853 controls the
    error with dim 4 (planar) or 6 (full), so the two take different steps
    and agree only to the tolerance; there the planar kernel is checked
    against the oracle on every 512th orbit and by the energy (Jacobi) drift
    of all orbits."""
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device
    from galpy_amd.potential import evaluatePotentials

    n = 262144
    for pot, omega in ((_mw(), 0.0), (bar_potential(), 1.85)):
        yo = planar_ics(n, seed=31)
        tp, tf = PotentialTable(pot, planar=True), PotentialTable(pot)
        t = torch.linspace(0.0, 10.0, 101, dtype=torch.float64, device="cuda")
        rp, ep, _ = integrate_device(tp, torch.tensor(yo, device="cuda"), t, "symplec4_hip", dt=0.01)
        rf, ef, _ = integrate_device(tf, torch.tensor(in_plane(yo), device="cuda"), t, "symplec4_hip",
                                     dt=0.01)
        torch.cuda.synchronize()
        assert int((ep != 0).sum()) == 0 and int((ef != 0).sum()) == 0
        rf = rf.cpu().numpy()
        assert numpy.all(rf[..., 3:5] == 0.0)
        assert phase_diff(in_plane(rp.cpu().numpy()), rf) < 1e-11
        del rf
        tn = numpy.linspace(0.0, 10.0, 21)
        rp, ep, _ = integrate_device(tp, torch.tensor(yo, device="cuda"), torch.tensor(tn, device="cuda"),
                                     "dop853_hip")
        torch.cuda.synchronize()
        assert int((ep != 0).sum()) == 0
        rp = rp.cpu().numpy()
        ref, *_ = _oracle(pot, yo[::512], tn, 6)
        assert phase_diff(rp[::512], ref) < 1e-8
        R, vR, vT, phi = numpy.moveaxis(rp, -1, 0)
        E = evaluatePotentials(pot, R, 0.0, phi, tn[None, :]) + 0.5 * (vR**2 + vT**2) - omega * R * vT
        dE = numpy.abs(E / E[:, :1] - 1.0).max(axis=1)
        assert numpy.median(dE) < 1e-10 and numpy.mean(dE <= 1e-8) > 0.999, numpy.quantile(dE, [0.5, 0.999])


def test_planar_orbit_frontend():
    from galpy_amd.orbit import Orbit

    pot = _mw()
    o = Orbit(planar_ics(8, seed=4))
    t = numpy.linspace(0.0, 10.0, 201)
    o.integrate(t, pot, method="dop853_hip")
    assert o.getOrbit().shape == (8, 201, 4)
    E = o.E()
    assert numpy.abs(E / E[:, :1] - 1.0).max() < 1e-10
    with pytest.raises(AttributeError):
        o.z()
    ref, *_ = _oracle(pot, planar_ics(8, seed=4), t, 6)
    assert phase_diff(o.getOrbit(), ref) < 1e-8
    # a full table for planar rows is refused, not silently reinterpreted
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device

    with pytest.raises(ValueError):
        integrate_device(PotentialTable(pot), torch.zeros((4, 4), dtype=torch.float64, device="cuda"),
                         torch.linspace(0, 1, 3, dtype=torch.float64, device="cuda"), "symplec4_hip")


def test_planar_full_rows_rejected_on_host():
    from galpy_amd.integrate import integratePlanarOrbit_hip

    with pytest.raises(ValueError):
        integratePlanarOrbit_hip(_mw(), ics(2, seed=1), numpy.linspace(0, 1, 3), "symplec4_hip")
