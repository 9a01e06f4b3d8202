"""The oracle's IAS15 (wez_ias15.c) pinned by what the mathematics fixes.

* Its tables (and the device's, read from the kernel source): the spacings
  are the 8-point Gauss-Radau nodes on [0, 1] (the fixed end 0 and the roots
  of (P_7 + P_8)(2h - 1) / h), and the other three tables follow from them by
  the recurrences of Everhart's method, all computed here in 40-digit mpmath:
  every entry within 4 ulps (the reference's own literals lie up to 3 ulps
  from the exact values), and the device's tables equal the oracle's digit
  for digit.  A transposed entry or a wrong digit before the last one or two
  fails this.
* Accuracy against DOP853 at rtol = atol = 1e-14 and energy conservation over
  100 time units in MWPotential2014; a Kepler orbit back at its start after
  one period and at pericentre a (1 - e) half a period in.
* Every force of a step is taken at the step's end time, as in the
  reference: in a static bar IAS15 agrees with DOP853 at 1e-14 to 1e-6 (2e-8
  for orbits crossing the bar's force kink at rb, 1e-12 otherwise), in the
  same bar rotating (omegab = 1.85) it is off by more than 1e-2 (forces at
  the substep times would keep it accurate there too).
* The reference's run-time rules: planar orbits equal full ones at z = 0,
  integrating back over the reversed grid recovers the initial state, and a
  caller's dt shorter than the grid spacing stops the run after nt steps'
  worth of time (the rows past it are NaN)."""

import os
import re

import numpy
import pytest

from helpers import ics, oracle_integrate, phase_diff, planar_ics

TIGHT = numpy.log(1e-14)
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mw():
    from galpy_amd.potential import MWPotential2014, parse_pot

    _, pt, pa, _ = parse_pot(MWPotential2014)
    return pt, pa


def _radau_tables():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    # (P_7 + P_8)(x) vanishes at x = -1 and at the 7 interior Radau nodes
    c7 = mpmath.taylor(lambda x: mpmath.legendre(7, x) + mpmath.legendre(8, x), 0, 8)
    roots = sorted(mpmath.polyroots(c7[::-1], maxsteps=200, extraprec=200), key=lambda r: mpmath.re(r))
    h = [mpmath.mpf(0)] + [(mpmath.re(r) + 1) / 2 for r in roots[1:]]
    assert abs(roots[0] + 1) < mpmath.mpf(10) ** -30 and len(h) == 8
    rr = [[(1 / (h[k] - h[j]) if j < k else 0) for j in range(7)] for k in range(8)]
    cc = [[mpmath.mpf(0)] * 7 for _ in range(7)]
    dd = [[mpmath.mpf(0)] * 7 for _ in range(7)]
    cc[0][0] = dd[0][0] = mpmath.mpf(1)
    for k in range(1, 7):
        cc[k][k] = dd[k][k] = mpmath.mpf(1)
        cc[k][0] = -h[k] * cc[k - 1][0]
        dd[k][0] = h[1] * dd[k - 1][0]
        for j in range(1, k):
            cc[k][j] = cc[k - 1][j - 1] - h[k] * cc[k - 1][j]
            dd[k][j] = dd[k - 1][j - 1] + h[j + 1] * dd[k - 1][j]
    return h, rr, cc, dd


def _read_table(text, name):
    m = re.search(name + r"(\[\d+\])+\s*=\s*\{(.*?)\};", text, re.S)
    assert m, name
    return [float(v) for v in re.findall(r"[-+]?\d+\.?\d*(?:e[-+]?\d+)?", m.group(2))]


@pytest.mark.parametrize("path", ["oracle/integrators.hpp", "galpy_amd/csrc/ias15_dev.h"])
def test_radau_tables_follow_from_the_nodes(path):
    h, rr, cc, dd = _radau_tables()
    text = open(os.path.join(ROOT, path)).read()
    for name, exact in (("h", [h]), ("rr", rr), ("cc", cc), ("dd", dd)):
        want = numpy.array([float(v) for row in exact for v in row])
        got = numpy.array(_read_table(text, r"\b" + name))
        assert got.shape == want.shape, name
        # the reference's values were computed in double precision and lie
        # within 3 ulps of the exact ones: at most 4 ulps here, exact zeros
        assert numpy.all((want == 0) == (got == 0)), name
        ulps = numpy.abs(got - want) / numpy.spacing(numpy.abs(want))
        assert ulps.max() <= 4.0, (name, ulps.max())


def test_device_and_oracle_tables_are_identical():
    dev = open(os.path.join(ROOT, "galpy_amd/csrc/ias15_dev.h")).read()
    orc = open(os.path.join(ROOT, "oracle/integrators.hpp")).read()
    for name in ("h", "rr", "cc", "dd"):
        assert _read_table(dev, r"\b" + name) == _read_table(orc, r"\b" + name), name


def test_ias15_accuracy_and_energy():
    from galpy_amd.potential import MWPotential2014, evaluatePotentials

    pt, pa = _mw()
    yo = ics(8, seed=0)
    t = numpy.linspace(0.0, 10.0, 101)
    ref, *_ = oracle_integrate(pt, pa, yo, t, 6, rtol=TIGHT, atol=TIGHT)
    res, err, st, _ = oracle_integrate(pt, pa, yo, t, 7)
    assert numpy.all(err == 0)
    assert phase_diff(res, ref) < 1e-10
    # the controller accepts a step only if the next one is not shorter, so
    # a large share of attempts is rejected (as in the reference)
    assert numpy.all(st[:, 1] > 0) and numpy.all(st[:, 0] > 10)
    res, *_ = oracle_integrate(pt, pa, yo, numpy.linspace(0.0, 100.0, 1001), 7)
    R, vR, vT, z, vz, _ = numpy.moveaxis(res, -1, 0)
    E = evaluatePotentials(MWPotential2014, R, z) + 0.5 * (vR**2 + vT**2 + vz**2)
    assert numpy.abs(E / E[:, :1] - 1.0).max() < 1e-12


def test_ias15_kepler_period():
    from galpy_amd.potential import KeplerPotential, parse_pot

    GM = 1.3
    _, pt, pa, _ = parse_pot([KeplerPotential(amp=GM)])
    R0, vT0, vz0 = 1.0, 0.9, 0.35
    E = 0.5 * (vT0**2 + vz0**2) - GM / R0
    a = -GM / (2.0 * E)
    T = 2.0 * numpy.pi * a**1.5 / numpy.sqrt(GM)
    yo = numpy.array([[R0, 0.0, vT0, 0.0, vz0, 0.0]])
    res, err, _, _ = oracle_integrate(pt, pa, yo, numpy.linspace(0.0, T, 21), 7)
    assert err[0] == 0
    assert numpy.abs(res[0, -1, :5] - yo[0, :5]).max() < 1e-10
    assert abs((res[0, -1, 5] + numpy.pi) % (2 * numpy.pi) - numpy.pi) < 1e-10
    e = numpy.sqrt(1.0 - (R0 * numpy.hypot(vT0, vz0)) ** 2 / (GM * a))
    half = res[0, 10]
    assert abs(numpy.hypot(half[0], half[3]) - a * (1.0 - e)) < 1e-10


def test_ias15_planar_equals_full_in_the_plane():
    from galpy_amd.potential import MWPotential2014, parse_pot_planar

    pt, pa = _mw()
    _, ppt, ppa, _ = parse_pot_planar(MWPotential2014)
    yp = planar_ics(16, seed=3)
    full = numpy.zeros((16, 6))
    full[:, [0, 1, 2, 5]] = yp
    t = numpy.linspace(0.0, 10.0, 51)
    resp, errp, _, _ = oracle_integrate(ppt, ppa, yp, t, 7, planar=True)
    resf, errf, _, _ = oracle_integrate(pt, pa, full, t, 7)
    assert numpy.all(errp == 0) and numpy.all(errf == 0)
    assert numpy.abs(resf[..., [3, 4]]).max() == 0.0
    assert phase_diff(resp, resf[..., [0, 1, 2, 5]]) < 1e-12


def test_ias15_backward_over_the_reversed_grid():
    pt, pa = _mw()
    yo = ics(8, seed=5)
    t = numpy.linspace(0.0, 10.0, 101)
    fwd, err, _, _ = oracle_integrate(pt, pa, yo, t, 7)
    back, err2, _, _ = oracle_integrate(pt, pa, fwd[:, -1], t[::-1].copy(), 7)
    assert numpy.all(err == 0) and numpy.all(err2 == 0)
    assert phase_diff(back[:, ::-1], fwd) < 1e-9


def test_ias15_caller_dt_bounds_the_run():
    pt, pa = _mw()
    yo = ics(4, seed=6)
    t = numpy.linspace(0.0, 10.0, 101)
    res, err, _, _ = oracle_integrate(pt, pa, yo, t, 7, dt=0.05)
    assert numpy.all(err == 0)
    # the accepted steps add up to |nt dt| = 5.05 (> 5.05 after the last
    # one): rows up to t = 5 are written, none past the step that ends the run
    ref, *_ = oracle_integrate(pt, pa, yo, t, 7)
    assert numpy.all(numpy.isfinite(res[:, :51]))
    assert numpy.isnan(res[:, -1]).all()
    assert phase_diff(res[:, :51], ref[:, :51]) < 1e-9


def test_ias15_forces_at_the_step_end_time():
    from galpy_amd.potential import DehnenBarPotential, LogarithmicHaloPotential, parse_pot_planar

    t = numpy.linspace(0.0, 10.0, 101)
    yp = planar_ics(32, seed=13)
    errs = []
    for omegab in (0.0, 1.85):
        pot = [LogarithmicHaloPotential(q=0.9), DehnenBarPotential(omegab=omegab, rb=0.8, Af=0.01)]
        _, pt, pa, _ = parse_pot_planar(pot)
        res, err, _, _ = oracle_integrate(pt, pa, yp, t, 7, planar=True)
        ref, *_ = oracle_integrate(pt, pa, yp, t, 6, rtol=TIGHT, atol=TIGHT, planar=True)
        assert numpy.all(err == 0)
        errs.append(phase_diff(res, ref))
    assert errs[0] < 1e-6 and errs[1] > 1e-2, errs
