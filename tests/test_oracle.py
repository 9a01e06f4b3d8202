"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the
checker of the HIP path: mpmath fixtures, galpy's published numbers,
closed-form orbits, convergence orders and invariants.  CPU only."""

import json
import os

import numpy
import pytest

from helpers import (
    LN1EM12,
    oracle,
    oracle_forces,
    oracle_integrate,
    oracle_potential,
)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _mw_packed():
    from galpy_amd.potential import MWPotential2014, parse_pot

    _, pt, pa, _ = parse_pot(MWPotential2014)
    return pt, pa


# ---------------------------------------------------------------- special fns
def test_incgamma_vs_mpmath():
    import ctypes

    lib = oracle()
    rows = _load("incgamma.json")
    for a in sorted({r["a"] for r in rows}):
        sel = [r for r in rows if r["a"] == a]
        x = numpy.array([r["x"] for r in sel])
        lo = numpy.array([r["lower"] for r in sel])
        up = numpy.array([r["upper"] for r in sel])
        out_lo = numpy.empty_like(x)
        out_up = numpy.empty_like(x)
        P = ctypes.c_void_p
        lib.oracle_lower_incgamma(ctypes.c_double(a), len(x), x.ctypes.data_as(P), out_lo.ctypes.data_as(P))
        lib.oracle_upper_incgamma(ctypes.c_double(a), len(x), x.ctypes.data_as(P), out_up.ctypes.data_as(P))
        assert numpy.all(numpy.abs(out_lo / lo - 1) < 5e-14), a
        # the upper function underflows relative accuracy only where it is
        # negligible against Gamma(a); compare relative to Gamma(a) there
        ga = lo[-1] + up[-1]
        assert numpy.all(numpy.abs(out_up - up) <= 5e-14 * numpy.maximum(numpy.abs(up), 1e-3 * ga)), a


def test_host_table_incgamma_vs_mpmath():
    """The HIP library's host-fitted Gamma(a,x) table, evaluated on the host
    exactly as the device evaluates it, against the mpmath fixture."""
    import ctypes

    from galpy_amd import _lib
    from galpy_amd.potential import PowerSphericalPotentialwCutoff, parse_pot

    lib = _lib.load()
    fn = lib.gh_host_bulge_lower_gamma
    V = ctypes.c_void_p
    fn.argtypes = [ctypes.c_int, V, V, ctypes.c_int, ctypes.c_int, V, V]
    rows = _load("incgamma.json")
    for a in (0.1, 0.6, 1.0, 1.5, 0.35):
        npot, pt, pa, _ = parse_pot([PowerSphericalPotentialwCutoff(alpha=3.0 - 2.0 * a, rc=0.3)])
        sel = [r for r in rows if r["a"] == a]
        x = numpy.array([r["x"] for r in sel])
        lo = numpy.array([r["lower"] for r in sel])
        out = numpy.empty_like(x)
        assert fn(npot, pt.ctypes.data, pa.ctypes.data, 0, len(x), x.ctypes.data, out.ctypes.data) == 0
        assert numpy.all(numpy.abs(out / lo - 1) < 5e-14), a


def test_host_log_ge1_vs_mpmath():
    """The force kernels' logarithm (NFW's log(1 + r/a)), evaluated on the
    host by the same code the device runs, within 1.5 ulp of mpmath's 40-digit
    log over [1, 2^60], including the reduction boundaries sqrt(2) 2^k and
    exact powers of two."""
    import mpmath

    from galpy_amd import _lib

    lib = _lib.load()
    rng = numpy.random.default_rng(11)
    y = numpy.concatenate([
        1.0 + rng.uniform(0.0, 3.0, 3000),
        numpy.exp2(rng.uniform(0.0, 60.0, 3000)),
        1.0 + numpy.array([0.0, 1e-300, 1e-16, 2.2e-16, 1e-12, 1e-8, 1e-4]),
        numpy.sqrt(2.0) * numpy.exp2(numpy.arange(0, 20)),
        numpy.nextafter(numpy.sqrt(2.0) * numpy.exp2(numpy.arange(0, 20)), 0.0),
        numpy.exp2(numpy.arange(0, 60)).astype(float),
    ])
    out = numpy.empty_like(y)
    assert lib.gh_host_log_ge1(len(y), y.ctypes.data, out.ctypes.data) == 0
    mpmath.mp.dps = 40
    ref = numpy.array([float(mpmath.log(mpmath.mpf(float(v)))) for v in y])
    ulp = numpy.spacing(numpy.maximum(numpy.abs(ref), numpy.finfo(float).tiny))
    assert numpy.all(numpy.abs(out - ref) <= 1.5 * ulp)
    assert out[numpy.flatnonzero(y == 1.0)[0]] == 0.0


def test_device_sincos_vs_mpmath():
    """The force kernels' sine and cosine (one Cody-Waite reduction by pi/2,
    Taylor polynomials to degree 19), run on the host from the same source:
    within 2 ulp of the magnitude scale (max(|value|, 2^-53)) of the 40-digit
    values over the angles the rotating components see (|x| up to 1e5,
    multiples of pi/2 and their neighbours, every quadrant and sign), and the
    library's values beyond the reduction's range."""
    import mpmath

    from galpy_amd import _lib

    lib = _lib.load()
    rng = numpy.random.default_rng(17)
    k = numpy.arange(-40, 41)
    x = numpy.concatenate([
        rng.uniform(-10.0, 10.0, 4000), rng.uniform(-1e5, 1e5, 2000), numpy.exp2(rng.uniform(-60, 0, 500)),
        k * (numpy.pi / 2), numpy.nextafter(k * (numpy.pi / 2), numpy.inf), numpy.nextafter(k * (numpy.pi / 2), -numpy.inf),
        numpy.array([0.0, -0.0, numpy.pi / 4, -numpy.pi / 4, 3e9, -5e12]),
    ])
    s, c = numpy.empty_like(x), numpy.empty_like(x)
    assert lib.gh_host_sincos(len(x), x.ctypes.data, s.ctypes.data, c.ctypes.data) == 0
    mpmath.mp.dps = 40
    rs = numpy.array([float(mpmath.sin(mpmath.mpf(float(v)))) for v in x])
    rc = numpy.array([float(mpmath.cos(mpmath.mpf(float(v)))) for v in x])
    for got, ref in ((s, rs), (c, rc)):
        ulp = numpy.spacing(numpy.maximum(numpy.abs(ref), 2.0**-53))
        bad = numpy.abs(got - ref) > 2.0 * ulp
        assert not bad.any(), (x[bad][:5], got[bad][:5], ref[bad][:5])
    big = numpy.abs(x) > 1e9
    assert numpy.array_equal(s[big], numpy.sin(x[big])) and numpy.array_equal(c[big], numpy.cos(x[big]))


# ---------------------------------------------------------------- potentials
NEW_SPHERICAL = ["powersph", "powersph_a2", "kepler", "jaffe", "isochrone", "plummer", "pseudoiso", "kuzmin",
                 "burkert", "dehnencore", "dehnen", "dehnen_a05", "homogsphere", "flatpower", "flatpower_a0"]


@pytest.mark.parametrize("name", ["psc", "mn", "nfw", "hernquist", "loghalo_q09", "loghalo_triax", "dehnenbar"]
                         + NEW_SPHERICAL)
def test_oracle_potentials_vs_mpmath(name):
    case = _load("potentials.json")[name]
    pts = case["points"]
    R = numpy.array([p["R"] for p in pts])
    z = numpy.array([p["z"] for p in pts])
    phi = numpy.array([p["phi"] for p in pts])
    t = numpy.array([p["t"] for p in pts])
    FR, Fz, Fp = oracle_forces(case["pot_type"], case["pot_args"], R, z, phi, t)
    pot = oracle_potential(case["pot_type"], case["pot_args"], R, z, phi, t)
    for got, key in ((FR, "FR"), (Fz, "Fz"), (Fp, "Fphi"), (pot, "pot")):
        want = numpy.array([p[key] for p in pts])
        scale = numpy.max(numpy.abs(want)) + 1e-300
        assert numpy.all(numpy.abs(got - want) <= 1e-12 * numpy.maximum(numpy.abs(want), 1e-6 * scale)), (name, key)


def test_mwpotential2014_parameters():
    """galpy tests/test_potential.py:4784-4812."""
    from galpy_amd.potential import MWPotential2014

    pot = MWPotential2014
    assert pot[0].rc == 1.9 / 8.0 and pot[0].alpha == 1.8
    assert abs(pot[0].Rforce(1.0, 0.0) + 0.05) < 1e-14
    assert abs(pot[1]._a - 3.0 / 8.0) < 1e-14 and abs(pot[1]._b - 0.28 / 8.0) < 1e-14
    assert abs(pot[1].Rforce(1.0, 0.0) + 0.60) < 1e-14
    assert abs(pot[2].a - 16.0 / 8.0) < 1e-14
    assert abs(pot[2].Rforce(1.0, 0.0) + 0.35) < 1e-14
    # the packed amplitudes reproduce the same forces through the oracle
    pt, pa = _mw_packed()
    for i, want in enumerate((-0.05, -0.60, -0.35)):
        npack = {15: 5, 5: 3, 9: 2}
        off = sum(npack[c] for c in pt[:i])
        FR, _, _ = oracle_forces(pt[i:i + 1], pa[off:off + npack[pt[i]]], numpy.array([1.0]), 0.0)
        assert abs(FR[0] - want) < 1e-14


def test_packing_layout():
    """_parse_pot codes and argument order (integrateFullOrbit.py:65-181)."""
    from galpy_amd.potential import (
        DehnenBarPotential,
        HernquistPotential,
        LogarithmicHaloPotential,
        MWPotential2014,
        parse_pot,
    )

    npot, pt, pa, tf = parse_pot(MWPotential2014)
    assert npot == 3 and list(pt) == [15, 5, 9] and tf == []
    assert pa.shape == (10,) and pa[1] == 1.8 and pa[2] == 1.9 / 8 and pa[3] == 0.0 and pa[4] == 0.0
    assert pa[6] == 3.0 / 8.0 and pa[7] == 0.28 / 8.0 and pa[9] == 2.0
    bar = DehnenBarPotential(omegab=1.85, rb=0.8, Af=0.01)
    npot, pt, pa, _ = parse_pot([LogarithmicHaloPotential(q=0.9), bar, HernquistPotential(a=2.0)])
    assert list(pt) == [0, 1, 8]
    assert list(pa[:4]) == [1.0, 0.9, (10**-8) ** 2.0, 2.0]  # core = 1e-8, squared as galpy does
    tb = 2 * numpy.pi / 1.85
    numpy.testing.assert_allclose(pa[4:10], [0.01, -4 * tb, -2 * tb, 0.8, 1.85, 25 * numpy.pi / 180], rtol=1e-15)
    assert list(pa[10:]) == [1.0, 2.0]
    pa_tri = parse_pot([LogarithmicHaloPotential(q=0.9, b=0.8)])[2]
    assert pa_tri[3] == 1.0 - 1.0 / 0.8**2


# ---------------------------------------------------------------- integrators
def _galpypaper(res):
    from galpy_amd.potential import MWPotential2014, evaluatePotentials

    R, vR, vT, z, vz, phi = res.T
    pot = MWPotential2014
    r = numpy.sqrt(R**2 + z**2)
    E = evaluatePotentials(pot, R[0], z[0]) + 0.5 * (vR[0] ** 2 + vT[0] ** 2 + vz[0] ** 2)
    ER = 0.5 * vR[-1] ** 2 + 0.5 * vT[-1] ** 2 + evaluatePotentials(pot, R[-1], 0.0)
    Ez = 0.5 * vz[-1] ** 2 + evaluatePotentials(pot, R[-1], z[-1]) - evaluatePotentials(pot, R[-1], 0.0)
    return dict(E=E, rperi=r.min(), rap=r.max(), zmax=numpy.abs(z).max(),
                e=(r.max() - r.min()) / (r.max() + r.min()), ER=ER, Ez=Ez, R2=8 * R[200], vR5=220 * vR[500])


@pytest.mark.parametrize("method", [3, 0, 6])
def test_galpypaper_golden_numbers_oracle(method):
    """galpy tests/test_galpypaper.py:293-360 (MWPotential2014, symplec4_c default)."""
    pt, pa = _mw_packed()
    ts = numpy.linspace(0.0, 10.0, 1001)
    res, err, _, _ = oracle_integrate(pt, pa, [0.8, 0.3, 0.75, 0.0, 0.2, 0.0], ts, method)
    g = _galpypaper(res[0])
    assert abs(g["E"] + 1.2547650648697966) < 1e-5
    assert abs(g["rperi"] - 0.44231993168097) < 1e-5
    assert abs(g["rap"] - 0.87769030382105) < 1e-5
    assert abs(g["zmax"] - 0.077452357289016) < 1e-5
    assert abs(g["e"] - 0.32982348199330563) < 1e-5
    assert abs(g["ER"] + 1.27601734263047) < 1e-5
    assert abs(g["Ez"] - 0.021252201847851909) < 1e-5
    assert abs(g["R2"] - 3.5470772876920007) < 1e-3
    assert abs(g["vR5"] - 45.202530965094553) < 1e-3


def _circular(pot_type, pot_args, R0, vc, t, method, dt=None, rtol=LN1EM12, atol=LN1EM12):
    res, err, stats, _ = oracle_integrate(pot_type, pot_args, [R0, 0.0, vc, 0.0, 0.0, 0.0], t, method,
                                          dt=dt, rtol=rtol, atol=atol)
    phi_exact = vc / R0 * t
    dphi = numpy.abs((res[0, :, 5] - phi_exact + numpy.pi) % (2 * numpy.pi) - numpy.pi)
    return res[0], dphi, stats


def test_circular_orbit_closed_form():
    """A circular orbit in a spherical potential has phi(t) = vc t / R exactly;
    DOP853 reproduces it at its tolerance, the symplectic steppers at the
    order of their truncation error."""
    pt, pa = [8], [1.0, 1.3]  # Hernquist
    R0 = 1.1
    FR, _, _ = oracle_forces(pt, pa, numpy.array([R0]), 0.0)
    vc = numpy.sqrt(-R0 * FR[0])
    t = numpy.linspace(0.0, 40.0, 81)
    res, dphi, _ = _circular(pt, pa, R0, vc, t, 6)
    assert dphi.max() < 1e-9
    assert numpy.abs(res[:, 0] - R0).max() < 1e-10 and numpy.abs(res[:, 2] - vc).max() < 1e-10
    # symplec4 order 4, leapfrog order 2: halving dt divides the phase error by 16 / 4
    for method, order in ((3, 4), (0, 2)):
        errs = [_circular(pt, pa, R0, vc, t, method, dt=h)[1].max() for h in (0.1, 0.05)]
        ratio = errs[0] / errs[1]
        assert abs(numpy.log2(ratio) - order) < 0.2, (method, ratio)


def test_dop853_tolerance_convergence():
    pt, pa = _mw_packed()
    yo = [1.0, 0.1, 1.05, 0.05, 0.02, 0.3]
    t = numpy.linspace(0.0, 20.0, 41)
    ref, *_ = oracle_integrate(pt, pa, yo, t, 6, rtol=numpy.log(1e-14), atol=numpy.log(1e-14))
    e = []
    for tol in (1e-8, 1e-10):
        r, *_ = oracle_integrate(pt, pa, yo, t, 6, rtol=numpy.log(tol), atol=numpy.log(tol))
        e.append(numpy.abs(r - ref)[..., :5].max())
    assert e[0] < 2e-5 and e[1] < e[0] / 20


def test_time_reversibility_symplectic():
    """Symmetric DKD steppers retrace their steps: forward then backward with
    the same step returns the initial condition to rounding."""
    pt, pa = _mw_packed()
    yo = numpy.array([1.2, 0.1, 0.9, 0.05, -0.03, 0.4])
    t = numpy.linspace(0.0, 5.0, 51)
    for method in (3, 0):
        fwd, *_ = oracle_integrate(pt, pa, yo, t, method, dt=0.01)
        back, *_ = oracle_integrate(pt, pa, fwd[0, -1] * [1, -1, -1, 1, -1, 1], t, method, dt=0.01)
        end = back[0, -1] * [1, -1, -1, 1, -1, 1]
        assert numpy.abs(end[:5] - yo[:5]).max() < 1e-11, method


def test_energy_and_jacobi_conservation():
    from galpy_amd.potential import MWPotential2014, evaluatePotentials

    pt, pa = _mw_packed()
    yo = numpy.array([[1.0, 0.1, 1.1, 0.0, 0.1, 0.0], [0.7, -0.2, 0.8, 0.05, 0.0, 1.0]])
    t = numpy.linspace(0.0, 50.0, 501)
    for method in (3, 6):
        res, *_ = oracle_integrate(pt, pa, yo, t, method)
        for o in res:
            R, vR, vT, z, vz, phi = o.T
            E = evaluatePotentials(MWPotential2014, R, z) + 0.5 * (vR**2 + vT**2 + vz**2)
            assert numpy.abs(E / E[0] - 1).max() < 1e-8
    # Jacobi integral in the rotating bar (test_orbit.py:180-195)
    from helpers import bar_potential
    from galpy_amd.potential import parse_pot

    pot = bar_potential()
    _, bpt, bpa, _ = parse_pot(pot)
    res, *_ = oracle_integrate(bpt, bpa, yo, t, 6)
    for o in res:
        R, vR, vT, z, vz, phi = o.T
        EJ = evaluatePotentials(pot, R, z, phi, t) + 0.5 * (vR**2 + vT**2 + vz**2) - 1.85 * R * vT
        assert numpy.abs(EJ / EJ[0] - 1).max() < 1e-8


def test_orbits_equal_single_orbit():
    """tests/test_orbits.py:931-1020: N orbits at once == N single orbits."""
    from helpers import ics

    pt, pa = _mw_packed()
    yo = ics(6, seed=3)
    t = numpy.linspace(0.0, 10.0, 101)
    for method in (3, 0, 6):
        many, *_ = oracle_integrate(pt, pa, yo, t, method, nthreads=4)
        for i in range(len(yo)):
            one, *_ = oracle_integrate(pt, pa, yo[i], t, method, nthreads=1)
            assert numpy.array_equal(one[0], many[i])


def test_estimated_step_is_power_of_two_fraction():
    """symplec4_estimate_step returns init_dt / 2^k, k <= 14
    (bovy_symplecticode.c:665-808, _MAX_DT_REDUCE)."""
    from helpers import ics

    pt, pa = _mw_packed()
    t = numpy.linspace(0.0, 1.0, 101)
    _, _, stats, _ = oracle_integrate(pt, pa, ics(50, seed=5), t, 3)
    k = numpy.log2((t[1] - t[0]) / stats[:, 3])
    assert numpy.allclose(k, numpy.round(k)) and k.min() >= 0 and k.max() <= 14
    assert numpy.array_equal(stats[:, 0], numpy.round((t[1] - t[0]) / stats[:, 3]) * 100)
