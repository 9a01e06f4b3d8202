"""GPU parity of the spherical potentials (PowerSpherical, Kepler, Jaffe,
Isochrone, Plummer, PseudoIsothermal, Burkert, DehnenCoreSpherical,
DehnenSpherical, HomogeneousSphere) and the Kuzmin disk against the oracle,
whose forces, potentials and Hessians are pinned to mpmath.

Bars as for MWPotential2014: fixed-step methods 1e-10 (relative, unit floor,
phi mod 2 pi) with the oracle's per-orbit step; DOP853 1e-8 at galpy's
default tolerances; integrate_dxdv the same on orbit and deviation;
device forces and potential values 1e-13 against the oracle's; adjoint
gradients w.r.t. every amplitude 1e-6 against central differences.

Two of these potentials are not smooth: the homogeneous sphere's force has
a kink at its surface and the razor-thin Kuzmin disk's vertical force flips
sign at z = 0.  An adaptive step that straddles either is accepted or
rejected on an error estimate that no longer measures the error, so the
oracle's own DOP853 result moves by up to ~3e-8 when its inputs change by
one ulp (by ~1e-13 with those two components removed), and a Kepler orbit
near escape amplifies per-step rounding far more than MWPotential2014 does.
Where a test runs such a case, an orbit is held to the bar or, if the
oracle's own one-ulp sensitivity of that orbit times a stated factor
exceeds the bar, to that product (10 for DOP853, where a flipped step
decision is the mechanism; 1e3 for fixed steps, where ~2,000 force
evaluations each differ by a few ulp)."""

import json
import os

import numpy
import pytest

from helpers import (
    assert_parity,
    cyl_to_rect,
    ics,
    oracle_forces,
    oracle_integrate,
    oracle_integrate_dxdv,
    oracle_potential,
    phase_diff,
    planar_ics,
    spherical_a,
    spherical_b,
    ulp_sensitivity,
)

pytestmark = pytest.mark.gpu

GROUPS = {"spherical A": spherical_a, "spherical B": spherical_b}
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _single_types():
    from galpy_amd import potential as P

    return {
        "PowerSpherical": [P.PowerSphericalPotential(alpha=1.7, normalize=1.0)],
        "PowerSpherical alpha=2": [P.PowerSphericalPotential(alpha=2.0, normalize=1.0)],
        "Kepler": [P.KeplerPotential(normalize=1.0)],
        "Jaffe": [P.JaffePotential(a=0.8, normalize=1.0)],
        "Isochrone": [P.IsochronePotential(b=0.7, normalize=1.0)],
        "Plummer": [P.PlummerPotential(normalize=1.0)],
        "PseudoIsothermal": [P.PseudoIsothermalPotential(a=1.3, normalize=1.0)],
        "KuzminDisk": [P.KuzminDiskPotential(a=0.6, normalize=1.0)],
        "Burkert": [P.BurkertPotential(normalize=1.0)],
        "DehnenCoreSpherical": [P.DehnenCoreSphericalPotential(a=1.2, normalize=1.0)],
        "DehnenSpherical": [P.DehnenSphericalPotential(a=0.9, alpha=1.5, normalize=1.0)],
        "DehnenSpherical alpha=0.5": [P.DehnenSphericalPotential(a=1.4, alpha=0.5, normalize=1.0)],
        "HomogeneousSphere": [P.HomogeneousSpherePotential(R=1.5, normalize=1.0)],
        "FlattenedPower": [P.FlattenedPowerPotential(alpha=0.5, q=0.9, core=0.1, normalize=1.0)],
        "FlattenedPower alpha=0": [P.FlattenedPowerPotential(alpha=0.0, q=0.8, core=0.2, normalize=1.0)],
    }


def _packed(pot, planar=False):
    from galpy_amd.potential import parse_pot, parse_pot_planar

    _, pt, pa, _ = (parse_pot_planar if planar else parse_pot)(pot)
    return pt, pa


def _rel(a, b):
    return (numpy.abs(a - b) / numpy.maximum(numpy.abs(b), 1.0)).max()


def test_device_forces_and_potential_match_oracle_and_mpmath():
    """gh_eval_hip (the kernels' own force and potential code) at the golden
    points and at random points, one component at a time."""
    import torch

    from galpy_amd.adjoint import _RawTable
    from galpy_amd.integrate import eval_device

    golden = json.load(open(os.path.join(GOLDEN, "potentials.json")))
    names = ["powersph", "powersph_a2", "kepler", "jaffe", "isochrone", "plummer", "pseudoiso", "kuzmin",
             "burkert", "dehnencore", "dehnen", "dehnen_a05", "homogsphere", "flatpower", "flatpower_a0"]
    for name in names:
        case = golden[name]
        pts = case["points"]
        pt = numpy.array(case["pot_type"], dtype=numpy.int32)
        pa = numpy.array(case["pot_args"], dtype=numpy.float64)
        tab = _RawTable(pt, pa)
        rng = numpy.random.default_rng(3)
        R = numpy.concatenate([[p["R"] for p in pts], rng.uniform(0.02, 6.0, 1000)])
        z = numpy.concatenate([[p["z"] for p in pts], rng.normal(0.0, 0.6, 1000)])
        dev = [torch.tensor(v, device="cuda") for v in (R, z, numpy.zeros_like(R), numpy.zeros_like(R))]
        pot, FR, Fz, Fphi = (o.cpu().numpy() for o in eval_device(tab, *dev))
        oFR, oFz, _ = oracle_forces(pt, pa, R, z)
        opot = oracle_potential(pt, pa, R, z)
        for got, want, key in ((FR, oFR, "FR"), (Fz, oFz, "Fz"), (pot, opot, "pot")):
            assert numpy.abs(got - want).max() <= 1e-13 * numpy.abs(want).max(), (name, key)
            gold = numpy.array([p[key] for p in pts])
            n = len(pts)
            assert numpy.all(numpy.abs(got[:n] - gold) <= 1e-12 * numpy.maximum(numpy.abs(gold), 1e-6)), (name, key)
        assert numpy.all(Fphi == 0.0)
        del tab


@pytest.mark.parametrize("group", list(GROUPS))
@pytest.mark.parametrize("method,code", [("symplec4_hip", 3), ("leapfrog_hip", 0), ("dop853_hip", 6)])
def test_full_orbits(group, method, code):
    """Symplectic steppers with galpy's per-orbit step estimation (no dt)."""
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = GROUPS[group]()
    yo = ics(256, seed=11)
    t = numpy.linspace(0.0, 10.0, 101)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
    ref, rerr, _, _ = oracle_integrate(*_packed(pot), yo, t, code)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    if code != 6:
        assert phase_diff(res, ref) < 1e-10
    else:
        _, sens = ulp_sensitivity(lambda y: oracle_integrate(*_packed(pot), y, t, code)[0], yo)
        assert_parity(res, ref, 1e-8, sens, 10.0, label=group)


@pytest.mark.parametrize("group", list(GROUPS))
@pytest.mark.parametrize("method,code", [("symplec4_hip", 3), ("dop853_hip", 6)])
def test_planar_orbits(group, method, code):
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = GROUPS[group]()
    yo = planar_ics(256, seed=12)
    t = numpy.linspace(0.0, 10.0, 101)
    res, err = integratePlanarOrbit_hip(pot, yo.copy(), t, method)
    ref, rerr, _, _ = oracle_integrate(*_packed(pot, True), yo, t, code, planar=True)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    if code != 6:
        assert phase_diff(res, ref) < 1e-10
    else:
        _, sens = ulp_sensitivity(
            lambda y: oracle_integrate(*_packed(pot, True), y, t, code, planar=True)[0], yo)
        assert_parity(res, ref, 1e-8, sens, 10.0, label=group)


def test_each_type_alone_orbits_and_tangent_maps():
    """Each component on its own through the run-time dispatch kernels:
    symplec4 at dt = 0.01 (orbit), and the full tangent map (orbit plus
    deviation, Hessian-vector products of that component)."""
    from galpy_amd.integrate import integrateFullOrbit_dxdv_hip, integrateFullOrbit_hip

    t = numpy.linspace(0.0, 5.0, 51)
    yo = ics(64, seed=13)
    w0 = cyl_to_rect(yo)
    d0 = numpy.random.default_rng(14).normal(size=w0.shape)
    for name, pot in _single_types().items():
        pt, pa = _packed(pot)
        res, err = integrateFullOrbit_hip(pot, yo.copy(), t, "symplec4_hip", dt=0.01)
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 3, dt=0.01)[0], yo)
        assert numpy.all(err == 0), name
        assert_parity(res, ref, 1e-10, sens, 1e3, label=name)
        res, err = integrateFullOrbit_dxdv_hip(pot, w0, d0, t, "symplec4_hip", dt=0.01)
        ref, sens = ulp_sensitivity(
            lambda y: oracle_integrate_dxdv(pt, pa, y, t, 3, dim=3, dt=0.01)[0], numpy.hstack([w0, d0]),
            phi_last=False)
        assert numpy.all(err == 0), name
        assert_parity(res, ref, 1e-10, sens, 1e3, phi_last=False, label=name + " dxdv")


@pytest.mark.parametrize("group", list(GROUPS))
@pytest.mark.parametrize("dim", [3, 2])
def test_dxdv(group, dim):
    from galpy_amd.integrate import integrateFullOrbit_dxdv_hip, integratePlanarOrbit_dxdv_hip

    pot = GROUPS[group]()
    if dim == 3:
        y0 = cyl_to_rect(ics(128, seed=15))
    else:
        R, vR, vT, phi = planar_ics(128, seed=15).T
        c, s = numpy.cos(phi), numpy.sin(phi)
        y0 = numpy.stack([R * c, R * s, vR * c - vT * s, vR * s + vT * c], axis=1)
    d0 = numpy.random.default_rng(16).normal(size=y0.shape)
    f = integrateFullOrbit_dxdv_hip if dim == 3 else integratePlanarOrbit_dxdv_hip
    pt, pa = _packed(pot, dim == 2)
    t = numpy.linspace(0.0, 10.0, 101)
    w = numpy.hstack([y0, d0])
    res, err = f(pot, y0, d0, t, "symplec4_hip")  # estimated step
    ref, rerr = oracle_integrate_dxdv(pt, pa, w, t, 3, dim=dim)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    assert _rel(res, ref) < 1e-10
    res, err = f(pot, y0, d0, t, "dop853_hip")
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate_dxdv(pt, pa, y, t, 6, dim=dim)[0], w, phi_last=False)
    assert numpy.all(err == 0)
    assert_parity(res, ref, 1e-8, sens, 10.0, phi_last=False, label=f"{group} dim {dim}")


def test_energy_kernel_matches_host_potential():
    """The energy-drift kernel's E(0) (its own potential evaluation) equals
    the host classes' energy of the initial states."""
    import torch

    from galpy_amd.integrate import PotentialTable, energy_drift_device
    from galpy_amd.potential import evaluatePotentials

    yo = ics(512, seed=17)
    for pot in list(GROUPS.values()) + [lambda: _single_types()["DehnenSpherical alpha=0.5"]]:
        pot = pot()
        rows = torch.tensor(yo[:, None, :].copy(), device="cuda")
        t = torch.zeros(1, dtype=torch.float64, device="cuda")
        e0, de = energy_drift_device(PotentialTable(pot), rows, t)
        R, vR, vT, z, vz, _ = yo.T
        want = evaluatePotentials(pot, R, z) + 0.5 * (vR**2 + vT**2 + vz**2)
        assert numpy.abs(e0.cpu().numpy() - want).max() <= 1e-13 * numpy.abs(want).max()
        assert float(de.abs().max()) == 0.0


def test_kepler_period_on_device():
    """DOP853 on the GPU returns a Kepler orbit to its initial state after
    T = 2 pi a^1.5 / sqrt(GM)."""
    from galpy_amd.integrate import integrateFullOrbit_hip
    from galpy_amd.potential import KeplerPotential

    GM = 1.3
    yo = numpy.array([[1.0, 0.0, 0.9, 0.0, 0.35, 0.0], [1.2, 0.1, 0.8, 0.05, -0.3, 1.0]])
    E = 0.5 * (yo[:, 1] ** 2 + yo[:, 2] ** 2 + yo[:, 4] ** 2) - GM / numpy.hypot(yo[:, 0], yo[:, 3])
    a = -GM / (2.0 * E)
    T = 2.0 * numpy.pi * a**1.5 / numpy.sqrt(GM)
    t = numpy.stack([numpy.zeros(2), T], axis=1)
    res, err = integrateFullOrbit_hip([KeplerPotential(amp=GM)], yo.copy(), t, "dop853_hip", rtol=1e-13, atol=1e-13)
    assert numpy.all(err == 0)
    assert phase_diff(res[:, -1], yo) < 1e-9


def _rect_to_cyl(w):
    x, y, z, vx, vy, vz = numpy.asarray(w).T
    R = numpy.hypot(x, y)
    c, s = x / R, y / R
    return numpy.stack([R, vx * c + vy * s, -vx * s + vy * c, z, vz, numpy.arctan2(y, x)], axis=1)


def _plane_flips(spec, w0, h, nsteps, dt, code):
    """Orbits for which some +-h stencil of a central difference puts a step
    on opposite sides of z = 0 (the oracle's trajectories, one output per
    step)."""
    t = numpy.arange(nsteps + 1) * dt
    flip = numpy.zeros(len(w0), dtype=bool)
    for j in range(6):
        e = numpy.zeros_like(w0)
        e[:, j] = h
        zp = oracle_integrate(spec.pot_type, spec.pot_args, _rect_to_cyl(w0 + e), t, code, dt=dt)[0][..., 3]
        zm = oracle_integrate(spec.pot_type, spec.pot_args, _rect_to_cyl(w0 - e), t, code, dt=dt)[0][..., 3]
        flip |= numpy.any(numpy.sign(zp) != numpy.sign(zm), axis=1)
    return flip


@pytest.mark.parametrize("method,code", [("symplec4_hip", 3), ("leapfrog_hip", 0)])
def test_adjoint_amplitude_gradients(method, code):
    """dL/d amp of every component equals central differences (h = 1e-6)
    within 1e-6 relative; the gradient w.r.t. every initial condition as
    well, except for an orbit whose +h and -h runs put some step on opposite
    sides of the Kuzmin disk's plane: the disk's vertical kick flips sign
    there, so that stencil straddles a jump and is no derivative (on the
    oracle alone, the one such orbit of these 128 under symplec4 differs
    from its tangent map by 100%)."""
    import torch

    from galpy_amd.adjoint import ParamSpec, _RawTable, flow_raw, orbit_flow

    for pot in (spherical_a(), spherical_b()):
        spec = ParamSpec(pot, [(i, 0) for i in range(len(pot))])
        w0 = cyl_to_rect(ics(128, seed=18))
        w0t = torch.tensor(w0, device="cuda")
        nsteps, dt, h = 400, 0.01, 1e-6
        rng = numpy.random.default_rng(19)
        target = torch.tensor(w0 + rng.normal(0, 0.05, w0.shape), device="cuda")
        theta0 = spec.theta0()
        theta = torch.tensor(theta0, device="cuda", requires_grad=True)
        w = w0t.clone().requires_grad_(True)
        ((orbit_flow(w, theta, spec, dt, nsteps, method) - target) ** 2).sum().backward()

        def L(th, ww):
            return ((flow_raw(_RawTable(*spec.packed(th)), ww, dt, nsteps, method) - target) ** 2).sum().item()

        for k in range(len(theta0)):
            e = numpy.zeros_like(theta0)
            e[k] = h
            fd = (L(theta0 + e, w0t) - L(theta0 - e, w0t)) / (2 * h)
            assert abs(theta.grad[k].item() - fd) <= 1e-6 * abs(fd), (k, theta.grad[k].item(), fd)
        tab = _RawTable(spec.pot_type, spec.pot_args)
        g = w.grad.cpu().numpy()
        keep = numpy.ones(len(w0), dtype=bool)
        if any(int(c) == 19 for c in spec.pot_type):
            keep = ~_plane_flips(spec, w0, h, nsteps, dt, code)
            assert keep.sum() >= 0.95 * len(keep)
        per_orbit = lambda ww: ((flow_raw(tab, ww, dt, nsteps, method) - target) ** 2).sum(dim=1)
        for j in range(6):
            e = torch.zeros_like(w0t)
            e[:, j] = h
            fd = ((per_orbit(w0t + e) - per_orbit(w0t - e)) / (2 * h)).cpu().numpy()
            assert numpy.abs(g[keep, j] - fd[keep]).max() <= 1e-6 * max(1e-3, numpy.abs(fd[keep]).max()), j


def test_flattened_power_planar_and_null_potential():
    """FlattenedPowerPotential's planar packing (amp, alpha, core^2) held in
    the full layout at q^2 = 1; a NullPotential leaves orbits on straight
    lines."""
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip, integratePlanarOrbit_hip

    t = numpy.linspace(0.0, 10.0, 101)
    for pot in ([P.FlattenedPowerPotential(alpha=0.5, q=0.9, core=0.1, normalize=1.0)],
                [P.FlattenedPowerPotential(alpha=0.0, q=0.8, core=0.2, normalize=1.0)]):
        yo = planar_ics(128, seed=21)
        pt, pa = _packed(pot, True)
        assert len(pa) == 3
        for method, code, tol in (("symplec4_hip", 3, 1e-10), ("dop853_hip", 6, 1e-8)):
            res, err = integratePlanarOrbit_hip(pot, yo.copy(), t, method)
            ref, rerr, _, _ = oracle_integrate(pt, pa, yo, t, code, planar=True)
            assert numpy.all(err == 0) and numpy.all(rerr == 0)
            assert phase_diff(res, ref) < tol, method
    yo = ics(64, seed=22)
    res, err = integrateFullOrbit_hip([P.NullPotential()], yo.copy(), t[:11], "dop853_hip")
    assert numpy.all(err == 0)
    x = cyl_to_rect(res.reshape(-1, 6)).reshape(64, 11, 6)
    x0 = cyl_to_rect(yo)
    assert numpy.abs(x[..., :3] - (x0[:, None, :3] + x0[:, None, 3:] * t[None, :11, None])).max() < 1e-12


def test_mn3_exponential_disk_orbits():
    """MN3ExponentialDiskPotential reaches the device as three Miyamoto-Nagai
    components (galpy's packing): full and planar orbits with a halo, the
    fixed-step and DOP853 bars."""
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip, integratePlanarOrbit_hip

    t = numpy.linspace(0.0, 10.0, 101)
    pot = [P.MN3ExponentialDiskPotential(hr=1.0 / 3.0, hz=1.0 / 16.0, normalize=0.6),
           P.NFWPotential(a=2.0, normalize=0.4)]
    yo = ics(128, seed=23)
    pt, pa = _packed(pot)
    assert list(pt) == [5, 5, 5, 9]
    for method, code, tol in (("symplec4_hip", 3, 1e-10), ("dop853_hip", 6, 1e-8)):
        res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
        ref, rerr, _, _ = oracle_integrate(pt, pa, yo, t, code)
        assert numpy.all(err == 0) and numpy.all(rerr == 0)
        assert phase_diff(res, ref) < tol, method
    yp = planar_ics(128, seed=24)
    ppt, ppa = _packed(pot, True)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "dop853_hip")
    ref, rerr, _, _ = oracle_integrate(ppt, ppa, yp, t, 6, planar=True)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    assert phase_diff(res, ref) < 1e-8


def test_kuzmin_kutuzov_staeckel():
    """KuzminKutuzovStaeckelPotential (pot_type 16): the device's forces and
    potential against the oracle's, orbits (full, planar) at the fixed-step
    and DOP853 bars; integrate_dxdv refuses it (no device Hessian)."""
    import torch

    from galpy_amd import _lib
    from galpy_amd import potential as P
    from galpy_amd.adjoint import _RawTable
    from galpy_amd.integrate import eval_device, integrateFullOrbit_dxdv_hip, integrateFullOrbit_hip
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = [P.KuzminKutuzovStaeckelPotential(ac=5.0, Delta=1.0, normalize=0.7),
           P.NFWPotential(a=2.0, normalize=0.3)]
    pt, pa = _packed(pot)
    rng = numpy.random.default_rng(9)
    R, z = rng.uniform(0.05, 6.0, 2000), rng.normal(0.0, 0.8, 2000)
    tab = _RawTable(pt, pa)
    dev = [torch.tensor(v, device="cuda") for v in (R, z, numpy.zeros_like(R), numpy.zeros_like(R))]
    pot_d, FR, Fz, _ = (o.cpu().numpy() for o in eval_device(tab, *dev))
    oFR, oFz, _ = oracle_forces(pt, pa, R, z)
    for got, want in ((FR, oFR), (Fz, oFz), (pot_d, oracle_potential(pt, pa, R, z))):
        assert numpy.abs(got - want).max() <= 1e-13 * numpy.abs(want).max()
    del tab
    t = numpy.linspace(0.0, 10.0, 101)
    yo = ics(128, seed=25)
    for method, code, tol in (("symplec4_hip", 3, 1e-10), ("dop853_hip", 6, 1e-8)):
        res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
        ref, rerr, _, _ = oracle_integrate(pt, pa, yo, t, code)
        assert numpy.all(err == 0) and numpy.all(rerr == 0)
        assert phase_diff(res, ref) < tol, method
    yp = planar_ics(128, seed=26)
    ppt, ppa = _packed(pot, True)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "dop853_hip")
    ref, rerr, _, _ = oracle_integrate(ppt, ppa, yp, t, 6, planar=True)
    assert numpy.all(err == 0) and numpy.all(rerr == 0)
    assert phase_diff(res, ref) < 1e-8
    y0 = cyl_to_rect(ics(2, seed=27))
    with pytest.raises(RuntimeError):
        integrateFullOrbit_dxdv_hip(pot, y0, numpy.ones_like(y0), numpy.linspace(0, 1, 3), "symplec4_hip")
    assert "Hessian" in _lib.last_error()


def test_softened_needle_bar():
    """SoftenedNeedleBarPotential (pot_type 25, rotating): the device's
    forces, torque and potential at t != 0 against the oracle's; orbits with
    a halo (full: symplec4 and DOP853; planar: DOP853), each orbit held to
    its bar or to 10x the oracle's own one-ulp sensitivity; integrate_dxdv
    refuses it (non-axisymmetric)."""
    import torch

    from galpy_amd import _lib
    from galpy_amd import potential as P
    from galpy_amd.adjoint import _RawTable
    from galpy_amd.integrate import eval_device, integrateFullOrbit_dxdv_hip, integrateFullOrbit_hip
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = [P.LogarithmicHaloPotential(normalize=1.0),
           P.SoftenedNeedleBarPotential(amp=0.05, a=1.5, b=0.1, c=0.5, pa=0.3, omegab=1.6)]
    pt, pa = _packed(pot)
    rng = numpy.random.default_rng(10)
    n = 2000
    R, z = rng.uniform(0.05, 6.0, n), rng.normal(0.0, 0.8, n)
    phi, t = rng.uniform(-numpy.pi, numpy.pi, n), rng.uniform(0.0, 5.0, n)
    tab = _RawTable(pt, pa)
    dev = [torch.tensor(v, device="cuda") for v in (R, z, phi, t)]
    pot_d, FR, Fz, Fp = (o.cpu().numpy() for o in eval_device(tab, *dev))
    oFR, oFz, oFp = oracle_forces(pt, pa, R, z, phi, t)
    for got, want in ((FR, oFR), (Fz, oFz), (Fp, oFp), (pot_d, oracle_potential(pt, pa, R, z, phi, t))):
        assert numpy.abs(got - want).max() <= 1e-13 * numpy.abs(want).max()
    del tab
    tt = numpy.linspace(0.0, 10.0, 101)
    yo = ics(128, seed=28)
    for method, code, bar in (("symplec4_hip", 3, 1e-10), ("dop853_hip", 6, 1e-8)):
        res, err = integrateFullOrbit_hip(pot, yo.copy(), tt, method)
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, tt, code)[0], yo)
        assert numpy.all(err == 0)
        assert_parity(res, ref, bar, sens, factor=10.0 if code == 6 else 1e3, label=method)
    yp = planar_ics(128, seed=29)
    ppt, ppa = _packed(pot, True)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), tt, "dop853_hip")
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(ppt, ppa, y, tt, 6, planar=True)[0], yp)
    assert numpy.all(err == 0)
    assert_parity(res, ref, 1e-8, sens, factor=10.0, label="planar")
    y0 = cyl_to_rect(ics(2, seed=30))
    with pytest.raises(RuntimeError):
        integrateFullOrbit_dxdv_hip(pot, y0, numpy.ones_like(y0), numpy.linspace(0, 1, 3), "dop853_hip")
    assert "Hessian" in _lib.last_error()


def test_elliptical_disk_planar_orbits():
    """EllipticalDiskPotential (planar only, code 4, growing between tform
    and tform + tsteady), LopsidedDiskPotential (code 6) and
    SteadyLogSpiralPotential (code 3, growing) with a halo:
    planar orbits, symplec4 and DOP853, each
    orbit held to its bar or to 10x the oracle's own one-ulp sensitivity."""
    from galpy_amd import potential as P
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = [P.LogarithmicHaloPotential(normalize=1.0),
           P.EllipticalDiskPotential(twophio=0.05, p=0.5, phib=0.3, tform=1.0, tsteady=2.0),
           P.LopsidedDiskPotential(phio=0.02, p=-0.2, phib=1.1),
           P.SteadyLogSpiralPotential(A=-0.02, tform=0.2, tsteady=0.5)]
    ppt, ppa = _packed(pot, True)
    t = numpy.linspace(0.0, 10.0, 101)
    yp = planar_ics(128, seed=31)
    for method, code, bar, factor in (("symplec4_hip", 3, 1e-10, 1e3), ("dop853_hip", 6, 1e-8, 10.0)):
        res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, method)
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(ppt, ppa, y, t, code, planar=True)[0], yp)
        assert numpy.all(err == 0)
        assert_parity(res, ref, bar, sens, factor=factor, label=method)


def test_planar_only_pointwise_forces_potential_and_energy():
    """EllipticalDisk (code 4), LopsidedDisk (6) and SteadyLogSpiral (3, at a
    non-integer multiplicity m = 1.5) on the device, point by point: planar
    F_R and F_phi against the oracle's at angles all around the circle (y < 0
    included: the spiral's phase takes phi in [0, 2 pi) as galpy's
    evalPlanarRectForce forms it) and at times just before, at and after the
    ends of both growth windows; potential values against the host classes'
    _evaluate (phi in (-pi, pi], as stored rows carry it); and the planar
    energy drift kernel against numpy on device-integrated rows."""
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, energy_drift_device, eval_device, integratePlanarOrbit_hip

    from helpers import oracle_planar_forces

    spiral = P.SteadyLogSpiralPotential(A=-0.02, alpha=-7.0, m=1.5, tform=0.2, tsteady=0.5)
    pot = [P.LogarithmicHaloPotential(normalize=1.0),
           P.EllipticalDiskPotential(twophio=0.05, p=0.5, phib=0.3, tform=1.0, tsteady=2.0),
           P.LopsidedDiskPotential(phio=0.02, p=-0.2, phib=1.1), spiral]
    ppt, ppa = _packed(pot, True)
    tab = PotentialTable(pot, planar=True)
    ts = 2.0 * numpy.pi / 0.65
    edges = [1.0, 3.0, 0.2 * ts, 0.7 * ts]  # EllipticalDisk tform, tform + tsteady; spiral's
    times = numpy.concatenate([[e - 1e-6, e, e + 1e-6, e - 1e-12, e + 1e-12] for e in edges] + [[0.0, 2.0, 10.0]])
    rng = numpy.random.default_rng(41)
    npt = 64
    R = numpy.tile(rng.uniform(0.2, 3.0, npt), len(times))
    phi = numpy.tile(rng.uniform(0.0, 2.0 * numpy.pi, npt), len(times))
    t = numpy.repeat(times, npt)
    assert (numpy.sin(phi) < 0).sum() > npt // 4
    dev = [torch.tensor(v, device="cuda") for v in (R, numpy.zeros_like(R), phi, t)]
    pot_d, FR, _, Fp = (o.cpu().numpy() for o in eval_device(tab, *dev))
    oFR, oFp = oracle_planar_forces(ppt, ppa, R, phi, t)
    assert numpy.abs(FR - oFR).max() <= 1e-13 * numpy.abs(oFR).max()
    assert numpy.abs(Fp - oFp).max() <= 1e-13 * numpy.abs(oFp).max()
    phw = numpy.arctan2(numpy.sin(phi), numpy.cos(phi))
    want = P.evaluatePotentials(pot, R, 0.0, phw, t)
    assert numpy.abs(pot_d - want).max() <= 1e-13 * numpy.abs(want).max()
    # energy drift along device rows, against numpy on the same rows
    yp = planar_ics(64, seed=43)
    tt = numpy.linspace(0.0, 5.0, 51)
    rows, err = integratePlanarOrbit_hip(pot, yp.copy(), tt, "dop853_hip")
    assert numpy.all(err == 0)
    e0, demax = energy_drift_device(tab, torch.tensor(rows, device="cuda"), torch.tensor(tt, device="cuda"))
    Rr, vR, vT, ph = (rows[..., k] for k in range(4))
    E = P.evaluatePotentials(pot, Rr, 0.0, ph, numpy.broadcast_to(tt, Rr.shape)) + 0.5 * (vR**2 + vT**2)
    d_np = numpy.abs((E - E[:, :1]) / E[:, :1]).max(axis=1)
    assert numpy.abs(e0.cpu().numpy() - E[:, 0]).max() <= 1e-13 * numpy.abs(E[:, 0]).max()
    assert numpy.abs(demax.cpu().numpy() - d_np).max() <= 1e-12 * max(d_np.max(), 1e-6)
