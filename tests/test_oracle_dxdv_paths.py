"""Pins the oracle's integrate_dxdv paths (integrateFullOrbit_dxdv,
integratePlanarOrbit_dxdv: the variational steppers with full trajectories,
DOP853 on the 12- and 8-dim variational systems) before they check the GPU.
CPU only.

The deviation of a variational run is the derivative of the final state
with respect to the initial one along the deviation, so it is pinned by
central finite differences of the pinned plain integrators; the planar
Hessian path is pinned against the full one in the plane; the state
transition matrix is symplectic; the orbit of a variational run is the plain
run's, bit for bit, for the symplectic steppers."""

import numpy
import pytest

from helpers import (
    bar_potential,
    cyl_to_rect,
    ics,
    oracle_integrate,
    oracle_integrate_dxdv,
    planar_ics,
)


def _mw(planar=False):
    from galpy_amd.potential import MWPotential2014, parse_pot, parse_pot_planar

    _, pt, pa, _ = (parse_pot_planar if planar else parse_pot)(MWPotential2014)
    return pt, pa


def _polar_to_rect(v):
    R, vR, vT, phi = numpy.atleast_2d(v).T
    c, s = numpy.cos(phi), numpy.sin(phi)
    return numpy.stack([R * c, R * s, vR * c - vT * s, vR * s + vT * c], axis=1)


def _rect_rows_full(res):
    """Rows (R,vR,vT,z,vz,phi) -> rectangular (x,y,z,vx,vy,vz)."""
    R, vR, vT, z, vz, phi = numpy.moveaxis(res, -1, 0)
    c, s = numpy.cos(phi), numpy.sin(phi)
    return numpy.stack([R * c, R * s, z, vR * c - vT * s, vR * s + vT * c, vz], axis=-1)


def _rect_rows_planar(res):
    R, vR, vT, phi = numpy.moveaxis(res, -1, 0)
    c, s = numpy.cos(phi), numpy.sin(phi)
    return numpy.stack([R * c, R * s, vR * c - vT * s, vR * s + vT * c], axis=-1)


def _rect_to_cyl_rows(w):
    """Full rectangular (x,y,z,vx,vy,vz) -> cylindrical rows for the oracle."""
    x, y, z, vx, vy, vz = numpy.moveaxis(w, -1, 0)
    R = numpy.hypot(x, y)
    phi = numpy.arctan2(y, x)
    c, s = numpy.cos(phi), numpy.sin(phi)
    return numpy.stack([R, vx * c + vy * s, vy * c - vx * s, z, vz, phi], axis=-1)


def _rect_to_polar_rows(w):
    x, y, vx, vy = numpy.moveaxis(w, -1, 0)
    R = numpy.hypot(x, y)
    phi = numpy.arctan2(y, x)
    c, s = numpy.cos(phi), numpy.sin(phi)
    return numpy.stack([R, vx * c + vy * s, vy * c - vx * s, phi], axis=-1)


@pytest.mark.parametrize("code", [3, 0])
def test_full_symplectic_dxdv_orbit_and_fd(code):
    pt, pa = _mw()
    y0 = cyl_to_rect(ics(8, seed=1))
    rng = numpy.random.default_rng(2)
    d0 = rng.normal(size=(8, 6))
    t = numpy.linspace(0.0, 5.0, 51)
    res, err = oracle_integrate_dxdv(pt, pa, numpy.hstack([y0, d0]), t, code)
    assert numpy.all(err == 0)
    # the orbit is the plain run's, bit for bit (same steps, same forces)
    plain, *_ = oracle_integrate(pt, pa, _rect_to_cyl_rows(y0), t, code)
    assert numpy.abs(res[..., :6] - _rect_rows_full(plain)).max() < 1e-13
    # the deviation is the directional derivative of the plain integrator
    eps = 1e-6
    # fixed step: the estimated step would jump between the +/- runs
    resf, _ = oracle_integrate_dxdv(pt, pa, numpy.hstack([y0, d0]), t, code, dt=0.01)
    hi, *_ = oracle_integrate(pt, pa, _rect_to_cyl_rows(y0 + eps * d0), t, code, dt=0.01)
    lo, *_ = oracle_integrate(pt, pa, _rect_to_cyl_rows(y0 - eps * d0), t, code, dt=0.01)
    fd = (_rect_rows_full(hi) - _rect_rows_full(lo)) / (2 * eps)
    dev = resf[..., 6:]
    assert numpy.abs(dev - fd).max() < 1e-6 * max(1.0, numpy.abs(fd).max())


@pytest.mark.parametrize("code", [3, 0])
def test_planar_symplectic_dxdv_is_full_in_plane_and_symplectic(code):
    ptp, pap = _mw(planar=True)
    pt, pa = _mw()
    y0 = _polar_to_rect(planar_ics(6, seed=3))
    t = numpy.linspace(0.0, 5.0, 51)
    M = numpy.empty((6, 4, 4))
    for j in range(4):
        d0 = numpy.zeros((6, 4))
        d0[:, j] = 1.0
        rp, ep = oracle_integrate_dxdv(ptp, pap, numpy.hstack([y0, d0]), t, code, dim=2, dt=0.01)
        # the same orbit and deviation in the full tangent map (z = vz = 0)
        wf = numpy.zeros((6, 12))
        wf[:, [0, 1, 3, 4]] = y0
        wf[:, [6, 7, 9, 10]] = d0
        rf, ef = oracle_integrate_dxdv(pt, pa, wf, t, code, dim=3, dt=0.01)
        assert numpy.all(ep == 0) and numpy.all(ef == 0)
        assert numpy.all(rf[..., [2, 5, 8, 11]] == 0.0)
        assert numpy.abs(rp - rf[..., [0, 1, 3, 4, 6, 7, 9, 10]]).max() < 1e-11
        M[:, :, j] = rp[:, -1, 4:]
    # (q, p) ordering of the deviation block: (dx, dy, dvx, dvy)
    Om = numpy.block([[numpy.zeros((2, 2)), numpy.eye(2)], [-numpy.eye(2), numpy.zeros((2, 2))]])
    for m in M:
        assert numpy.abs(m.T @ Om @ m - Om).max() < 1e-11


def test_planar_dxdv_orbit_is_plain_planar_orbit():
    ptp, pap = _mw(planar=True)
    yo = planar_ics(6, seed=4)
    y0 = _polar_to_rect(yo)
    t = numpy.linspace(0.0, 5.0, 51)
    res, err = oracle_integrate_dxdv(ptp, pap, numpy.hstack([y0, numpy.ones((6, 4))]), t, 3, dim=2)
    plain, *_ = oracle_integrate(ptp, pap, yo, t, 3, planar=True)
    assert numpy.abs(res[..., :4] - _rect_rows_planar(plain)).max() < 1e-13


@pytest.mark.parametrize("dim", [3, 2])
def test_dop853_dxdv_matches_finite_differences(dim):
    pt, pa = _mw(planar=dim == 2)
    if dim == 3:
        y0 = cyl_to_rect(ics(4, seed=5))
        to_rows, rows_to_rect = _rect_to_cyl_rows, _rect_rows_full
    else:
        y0 = _polar_to_rect(planar_ics(4, seed=5))
        to_rows, rows_to_rect = _rect_to_polar_rows, _rect_rows_planar
    rng = numpy.random.default_rng(6)
    d0 = rng.normal(size=y0.shape)
    t = numpy.linspace(0.0, 5.0, 11)
    tol = numpy.log(1e-13)
    res, err = oracle_integrate_dxdv(pt, pa, numpy.hstack([y0, d0]), t, 6, dim=dim, rtol=tol, atol=tol)
    assert numpy.all(err == 0)
    eps = 1e-5
    kw = dict(rtol=tol, atol=tol, planar=dim == 2)
    hi, *_ = oracle_integrate(pt, pa, to_rows(y0 + eps * d0), t, 6, **kw)
    lo, *_ = oracle_integrate(pt, pa, to_rows(y0 - eps * d0), t, 6, **kw)
    fd = (rows_to_rect(hi) - rows_to_rect(lo)) / (2 * eps)
    n = 2 * dim
    assert numpy.abs(res[..., n:] - fd).max() < 1e-6 * max(1.0, numpy.abs(fd).max())
    # and its orbit is DOP853's to the tolerance (the error norm also spans the deviation)
    plain, *_ = oracle_integrate(pt, pa, to_rows(y0), t, 6, **kw)
    assert numpy.abs(res[..., :n] - rows_to_rect(plain)).max() < 1e-9


def test_dxdv_rejects_potentials_without_hessian():
    from galpy_amd.potential import MiyamotoNagaiPotential, SoftenedNeedleBarPotential, parse_pot

    _, pt, pa, _ = parse_pot([MiyamotoNagaiPotential(normalize=1.0), SoftenedNeedleBarPotential(amp=0.1)])
    res, err = oracle_integrate_dxdv(pt, pa, numpy.hstack([cyl_to_rect(ics(2, seed=1)), numpy.ones((2, 6))]),
                                     numpy.linspace(0, 1, 3), 3)
    assert numpy.all(err == -1)
