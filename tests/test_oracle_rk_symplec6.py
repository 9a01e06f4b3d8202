"""The oracle's RK4 / RK6 (bovy_rk.c) and symplec6 (bovy_symplecticode.c:403-1000)
pinned by what the mathematics fixes: the order of convergence at fixed
steps against a DOP853 solution at rtol = atol = 1e-14 (error ratios near
2^4, 2^6, 2^6 per halving), conservation of Lz to rounding by symplec6 in an
axisymmetric potential and its time reversibility, the estimated symplectic
step init_dt / 2^k, and the reference's RK step rule (dt divided by
ceil(err^(1/order+1)), ndt = (long)(init_dt / dt) steps per output)."""

import numpy
import pytest

from helpers import LN1EM12, cyl_to_rect, ics, oracle_integrate, phase_diff

TIGHT = numpy.log(1e-14)


def _mw():
    from galpy_amd.potential import MWPotential2014, parse_pot

    _, pt, pa, _ = parse_pot(MWPotential2014)
    return pt, pa


@pytest.mark.parametrize("code,order", [(1, 4), (2, 6), (4, 6)])
def test_convergence_order(code, order):
    pt, pa = _mw()
    yo = ics(8, seed=3)
    t = numpy.linspace(0.0, 5.0, 51)
    ref, *_ = oracle_integrate(pt, pa, yo, t, 6, rtol=TIGHT, atol=TIGHT)
    dts = (0.05, 0.025) if order == 6 else (0.02, 0.01)
    errs = []
    for dt in dts:
        res, err, _, _ = oracle_integrate(pt, pa, yo, t, code, dt=dt)
        assert numpy.all(err == 0)
        errs.append(phase_diff(res, ref))
    ratio = errs[0] / errs[1]
    assert 2 ** (order - 1.2) < ratio < 2 ** (order + 1.5), (errs, ratio)


def test_symplec6_conserves_lz_and_is_reversible():
    pt, pa = _mw()
    yo = ics(8, seed=4)
    t = numpy.linspace(0.0, 20.0, 201)
    res, err, st, _ = oracle_integrate(pt, pa, yo, t, 4, dt=0.02)
    Lz = res[..., 0] * res[..., 2]
    assert numpy.abs(Lz - Lz[:, :1]).max() < 1e-13 * numpy.abs(Lz).max()
    # with the velocities reversed at the end, the symmetric splitting retraces the orbit
    flip = numpy.array([1, -1, -1, 1, -1, 1])
    back, *_ = oracle_integrate(pt, pa, res[:, -1] * flip, t, 4, dt=0.02)
    assert phase_diff(back[:, -1] * flip, yo) < 1e-10


def test_symplec6_estimated_step_is_power_of_two_fraction():
    pt, pa = _mw()
    yo = ics(64, seed=5)
    t = numpy.linspace(0.0, 10.0, 101)
    _, err, st, _ = oracle_integrate(pt, pa, yo, t, 4)
    k = numpy.log2(0.1 / st[:, 3])
    assert numpy.all(err == 0) and numpy.allclose(k, numpy.round(k), atol=1e-12)
    # a looser tolerance never asks for a smaller step
    _, _, st2, _ = oracle_integrate(pt, pa, yo, t, 4, rtol=numpy.log(1e-8), atol=numpy.log(1e-8))
    assert numpy.all(st2[:, 3] >= st[:, 3])


@pytest.mark.parametrize("code", [1, 2])
def test_rk_step_rule(code):
    """dt = init_dt / m for an integer m and (long)(init_dt / dt) steps per
    output interval (bovy_rk.c:86-92, 367-372)."""
    pt, pa = _mw()
    yo = ics(64, seed=6)
    t = numpy.linspace(0.0, 10.0, 101)
    _, err, st, _ = oracle_integrate(pt, pa, yo, t, code)
    assert numpy.all(err == 0)
    dt = st[:, 3]
    m = 0.1 / dt
    assert numpy.allclose(m, numpy.round(m), rtol=1e-12)
    ndt = (0.1 / dt).astype(numpy.int64)
    assert numpy.array_equal(st[:, 0].astype(numpy.int64), ndt * 100)


def test_dopr54_converges_with_tolerance_and_conserves_energy():
    """bovy_dopr54 (bovy_rk.c:501-760): the error against DOP853 at 1e-14
    falls with the tolerance (a factor ~100 per factor 100), no orbit is
    flagged, and at galpy's default tolerances the energy is conserved."""
    from galpy_amd.potential import MWPotential2014, evaluatePotentials

    pt, pa = _mw()
    yo = ics(8, seed=7)
    t = numpy.linspace(0.0, 10.0, 101)
    ref, *_ = oracle_integrate(pt, pa, yo, t, 6, rtol=TIGHT, atol=TIGHT)
    errs = []
    for tol in (1e-8, 1e-10, 1e-12):
        res, err, st, _ = oracle_integrate(pt, pa, yo, t, 5, rtol=numpy.log(tol), atol=numpy.log(tol))
        assert numpy.all(err == 0)
        errs.append(phase_diff(res, ref))
    assert errs[0] / errs[1] > 30 and errs[1] / errs[2] > 30, errs
    res, *_ = oracle_integrate(pt, pa, yo, numpy.linspace(0.0, 50.0, 501), 5)
    R, vR, vT, z, vz, _ = numpy.moveaxis(res, -1, 0)
    E = evaluatePotentials(MWPotential2014, R, z) + 0.5 * (vR**2 + vT**2 + vz**2)
    assert numpy.abs(E / E[:, :1] - 1.0).max() < 1e-9
