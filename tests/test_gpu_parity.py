"""GPU parity: the HIP integrators against the CPU oracle on the same seeded
inputs, through the C-ABI (integrateFullOrbit / integrateFullOrbit_hip).

Bars (stated per test): fixed-step methods match the oracle's final phase
space to 1e-10 (relative, unit floor, phi mod 2pi) and pick the identical
per-orbit step; DOP853 matches at every output time to 1e-8 relative at
galpy's default tolerances (rtol=atol=1e-12)."""

import numpy
import pytest

from helpers import bar_potential, ics, oracle_integrate, phase_diff

pytestmark = pytest.mark.gpu


def _mw():
    from galpy_amd.potential import MWPotential2014

    return MWPotential2014


def _packed(pot):
    from galpy_amd.potential import parse_pot

    _, pt, pa, _ = parse_pot(pot)
    return pt, pa


def _hip(pot, yo, t, method, dt=None):
    from galpy_amd.integrate import integrateFullOrbit_hip

    return integrateFullOrbit_hip(pot, yo.copy(), t, method, dt=dt)


@pytest.mark.parametrize("method,code", [("symplec4_hip", 3), ("leapfrog_hip", 0)])
def test_symplectic_estimated_dt_mw2014(method, code):
    # config-0 path: no dt, per-orbit step estimation, MWPotential2014
    pot = _mw()
    yo = ics(256, seed=1)
    t = numpy.linspace(0.0, 10.0, 1001)
    res, err = _hip(pot, yo, t, method)
    ref, rerr, stats, _ = oracle_integrate(*_packed(pot), yo, t, code)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    assert phase_diff(res[:, -1], ref[:, -1]) < 1e-10
    assert phase_diff(res, ref) < 1e-10


def test_symplec4_dt_choice_identical():
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device

    pot = _mw()
    yo = ics(512, seed=2)
    t = numpy.linspace(0.0, 100.0, 10001)[:101]
    table = PotentialTable(pot)
    dyo = torch.tensor(yo, device="cuda")
    dt_ = torch.tensor(t, device="cuda")
    stats = torch.zeros((len(yo), 3), dtype=torch.float64, device="cuda")
    res, err, stats = integrate_device(table, dyo, dt_, "symplec4_hip", stats=stats)
    torch.cuda.synchronize()
    ref, rerr, rstats, _ = oracle_integrate(*_packed(pot), yo, t, 3)
    s = stats.cpu().numpy()
    assert numpy.array_equal(s[:, 2], rstats[:, 3])  # identical step per orbit
    assert numpy.array_equal(s[:, 0], rstats[:, 0])  # identical orbit-step counts
    assert phase_diff(res.cpu().numpy(), ref) < 1e-10


def test_dop853_mw2014():
    pot = _mw()
    yo = ics(128, seed=3)
    t = numpy.linspace(0.0, 50.0, 101)
    res, err = _hip(pot, yo, t, "dop853_hip")
    ref, rerr, stats, _ = oracle_integrate(*_packed(pot), yo, t, 6)
    assert numpy.all(err == 0)
    assert phase_diff(res, ref) < 1e-8


def test_dop853_longest_first_order_leaves_results_unchanged():
    """More orbits than resident lanes turns on the pilot run and the
    longest-first claim order; each orbit must still come out bit-identical
    to the same orbit integrated in a call too small to be reordered."""
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device
    from galpy_amd.potential import HernquistPotential, NFWPotential

    n, split = 70000, 60000
    t = torch.linspace(0.0, 3.0, 4, dtype=torch.float64, device="cuda")
    for pot in (_mw(), [HernquistPotential(a=1.3), NFWPotential(a=3.0, amp=2.0)]):
        table = PotentialTable(pot)
        yo = torch.tensor(ics(n, seed=21), device="cuda")
        whole, err, _ = integrate_device(table, yo.clone(), t, "dop853_hip")
        parts = [integrate_device(table, yo[a:b].clone(), t, "dop853_hip")[0] for a, b in ((0, split), (split, n))]
        torch.cuda.synchronize()
        assert int((err != 0).sum()) == 0
        assert torch.equal(whole, torch.cat(parts))


def test_dop853_lockstep_lanes_all_need_dense_output():
    """130 copies of one orbit step in lockstep, so every lane of a wavefront
    passes an output time in the same step (more dense-output jobs at once
    than the per-wavefront queue holds): every copy must come out identical
    and match the oracle."""
    for pot in (_mw(), bar_potential()):
        yo = numpy.repeat(ics(1, seed=12), 130, axis=0)
        t = numpy.linspace(0.0, 20.0, 201)
        res, err = _hip(pot, yo, t, "dop853_hip")
        assert numpy.all(err == 0)
        assert numpy.array_equal(res, numpy.broadcast_to(res[:1], res.shape))
        pt, pa = _packed(pot)
        ref, rerr, _, _ = oracle_integrate(pt, pa, yo[:1], t, 6)
        assert phase_diff(res[:1], ref) < 1e-8


def test_dop853_energy_outliers_are_the_oracles():
    """The largest |dE/E| of the DOP853 benchmark orbits (first 65,536 of the
    1M-orbit MWPotential2014 workload, t = linspace(0, 500, 1001), rtol = atol
    = 1e-12) belong to the method, not to this implementation: the oracle
    integrates the same orbits to within 1e-8 of the GPU trajectories and its
    own energy drift equals the GPU's to within 1% (and 1e-12)."""
    import torch

    from galpy_amd import workloads
    from galpy_amd.integrate import PotentialTable, energy_drift_device, integrate_device

    from helpers import oracle_potential

    w = workloads.make("mw2014_dop853_1m", n=65536)
    table = PotentialTable(w["pot"])
    t = torch.tensor(w["t"], device="cuda")
    res, err, _ = integrate_device(table, torch.tensor(w["yo"], device="cuda"), t, "dop853_hip")
    _, demax = energy_drift_device(table, res, t)
    worst = torch.argsort(demax, descending=True)[:6].cpu().numpy()
    rows = res[torch.as_tensor(worst, device="cuda")].cpu().numpy()
    assert int((err != 0).sum()) == 0
    pt, pa = _packed(w["pot"])
    ref, rerr, _, _ = oracle_integrate(pt, pa, w["yo"][worst], w["t"], 6)
    assert numpy.all(rerr == 0)
    assert phase_diff(rows, ref) < 1e-8

    def drift(r):
        R, vR, vT, z, vz, phi = (r[..., k].ravel() for k in range(6))
        E = 0.5 * (vR**2 + vT**2 + vz**2) + oracle_potential(pt, pa, R, z, phi)
        E = E.reshape(r.shape[:2])
        return numpy.abs(E - E[:, :1]).max(axis=1) / numpy.abs(E[:, 0])

    dg, dr = drift(rows), drift(ref)
    assert dg.max() > 1e-8  # the outliers are outliers
    assert numpy.all(numpy.abs(dg - dr) <= 0.01 * dr + 1e-12), (dg, dr)


def test_dop853_bar_jacobi():
    pot = bar_potential()
    yo = ics(128, seed=4)
    t = numpy.linspace(0.0, 20.0, 201)
    res, err = _hip(pot, yo, t, "dop853_hip")
    ref, rerr, stats, _ = oracle_integrate(*_packed(pot), yo, t, 6)
    assert numpy.all(err == 0)
    assert phase_diff(res, ref) < 1e-8


@pytest.mark.parametrize("method,code", [("symplec4_hip", 3), ("leapfrog_hip", 0)])
def test_symplectic_fixed_dt_bar(method, code):
    pot = bar_potential()
    yo = ics(128, seed=5)
    t = numpy.linspace(0.0, 10.0, 101)
    res, err = _hip(pot, yo, t, method, dt=0.01)
    ref, rerr, stats, _ = oracle_integrate(*_packed(pot), yo, t, code, dt=0.01)
    assert phase_diff(res, ref) < 1e-10


def test_hernquist_loghalo_flattened():
    from galpy_amd.potential import HernquistPotential, LogarithmicHaloPotential

    for pot in ([HernquistPotential(a=1.3)], [LogarithmicHaloPotential(q=0.8)],
                [LogarithmicHaloPotential(q=0.9, b=0.8)]):
        yo = ics(64, seed=6)
        t = numpy.linspace(0.0, 10.0, 101)
        res, err = _hip(pot, yo, t, "symplec4_hip", dt=0.01)
        ref, rerr, stats, _ = oracle_integrate(*_packed(pot), yo, t, 3, dt=0.01)
        assert phase_diff(res, ref) < 1e-10
        res, err = _hip(pot, yo, t, "dop853_hip")
        ref, rerr, stats, _ = oracle_integrate(*_packed(pot), yo, t, 6)
        assert phase_diff(res, ref) < 1e-8


def test_indiv_t_and_backward():
    pot = _mw()
    yo = ics(64, seed=7)
    rng = numpy.random.default_rng(0)
    t = numpy.stack([numpy.linspace(0.0, L, 51) for L in rng.uniform(1.0, 5.0, len(yo))])
    for method, code in (("symplec4_hip", 3), ("dop853_hip", 6)):
        res, err = _hip(pot, yo, t, method)
        ref, *_ = oracle_integrate(*_packed(pot), yo, t, code)
        assert phase_diff(res, ref) < (1e-10 if code == 3 else 1e-8)
    tb = numpy.linspace(0.0, -5.0, 51)
    for method, code in (("symplec4_hip", 3), ("dop853_hip", 6)):
        res, err = _hip(pot, yo, tb, method)
        ref, *_ = oracle_integrate(*_packed(pot), yo, tb, code)
        assert phase_diff(res, ref) < (1e-10 if code == 3 else 1e-8)


def test_edge_sizes():
    pot = _mw()
    yo = ics(3, seed=8)
    # nt = 2, a single orbit, and an orbit count that is not a multiple of 64
    for n, t in ((1, numpy.array([0.0, 0.5])), (3, numpy.linspace(0, 1, 11)), (65, numpy.linspace(0, 1, 3))):
        y = ics(n, seed=9)
        for method, code in (("symplec4_hip", 3), ("leapfrog_hip", 0), ("dop853_hip", 6)):
            res, err = _hip(pot, y, t, method)
            ref, *_ = oracle_integrate(*_packed(pot), y, t, code)
            assert res.shape == (n, len(t), 6)
            assert phase_diff(res, ref) < (1e-8 if code == 6 else 1e-10)
    # zero orbits is a no-op
    from galpy_amd.integrate import integrateFullOrbit_hip

    res, err = integrateFullOrbit_hip(pot, numpy.zeros((0, 6)), numpy.linspace(0, 1, 5), "symplec4_hip")
    assert res.shape == (0, 5, 6)


def test_unsupported_method_and_potential_raise():
    from galpy_amd import _lib

    lib = _lib.load()
    pot = _mw()
    pt, pa = _packed(pot)
    yo = ics(4, seed=10)
    t = numpy.linspace(0, 1, 5)
    res = numpy.empty((4, 5, 6))
    err = numpy.zeros(4, dtype=numpy.int32)
    for code in (8, -1, 100):  # galpy has no odeint codes outside 0-7
        err[:] = 0
        lib.integrateFullOrbit(4, yo.copy(), 5, t, 0, len(pt), pt, pa, None, -9999.99, 0.0, 0.0, res, err,
                               code, None)
        assert numpy.all(err == -1)
        assert "odeint_type" in _lib.last_error()
    bad = numpy.array([999], dtype=numpy.int32)  # no galpy potential has this code
    err[:] = 0
    lib.integrateFullOrbit(4, yo.copy(), 5, t, 0, 1, bad, pa, None, -9999.99, 0.0, 0.0, res, err, 3, None)
    assert numpy.all(err == -1)
    assert "pot_type 999" in _lib.last_error()


def test_galpypaper_golden_numbers():
    """tests/test_galpypaper.py:293-360 of the reference, on the GPU path."""
    from galpy_amd.potential import evaluatePotentials

    pot = _mw()
    ts = numpy.linspace(0.0, 10.0, 1001)
    o, err = _hip(pot, numpy.array([0.8, 0.3, 0.75, 0.0, 0.2, 0.0]), ts, "symplec4_hip")
    R, vR, vT, z, vz, phi = o.T
    E = evaluatePotentials(pot, R[0], z[0]) + 0.5 * (vR[0] ** 2 + vT[0] ** 2 + vz[0] ** 2)
    assert abs(E + 1.2547650648697966) < 1e-5
    r = numpy.sqrt(R**2 + z**2)
    assert abs(r.min() - 0.44231993168097) < 1e-5
    assert abs(r.max() - 0.87769030382105) < 1e-5
    assert abs(numpy.abs(z).max() - 0.077452357289016) < 1e-5
    e = (r.max() - r.min()) / (r.max() + r.min())
    assert abs(e - 0.32982348199330563) < 1e-5
    ER = 0.5 * vR[-1] ** 2 + 0.5 * vT[-1] ** 2 + evaluatePotentials(pot, R[-1], 0.0)
    Ez = 0.5 * vz[-1] ** 2 + evaluatePotentials(pot, R[-1], z[-1]) - evaluatePotentials(pot, R[-1], 0.0)
    assert abs(ER + 1.27601734263047) < 1e-5
    assert abs(Ez - 0.021252201847851909) < 1e-5
    assert abs(8.0 * R[200] - 3.5470772876920007) < 1e-3
    assert abs(220.0 * vR[500] - 45.202530965094553) < 1e-3
