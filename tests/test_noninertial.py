"""NonInertialFrameForce (pot_type 39 with the caller's time functions, 45
with natural splines of them; NonInertialFrameForce.c), a velocity-dependent
force integrated by the first-order methods.  Pinned by closed forms of a
free particle seen from non-inertial frames: a frame rotating at constant
Omega about z, about a tilted axis, at a linearly growing rate (the Euler
force), and a frame accelerating at constant a0 -- the oracle's DOP853 and
RK6 orbits equal the inertial straight line mapped into the frame."""

import numpy
import pytest

from helpers import oracle_integrate, oracle_tfuncs


def _cyl(xyz, v):
    R = numpy.hypot(xyz[:, 0], xyz[:, 1])
    phi = numpy.arctan2(xyz[:, 1], xyz[:, 0])
    vR = (xyz[:, 0] * v[:, 0] + xyz[:, 1] * v[:, 1]) / R
    vT = (xyz[:, 0] * v[:, 1] - xyz[:, 1] * v[:, 0]) / R
    return numpy.stack([R, vR, vT, xyz[:, 2], v[:, 2], phi], axis=1)


def _xyz(rows):
    return numpy.stack([rows[..., 0] * numpy.cos(rows[..., 5]), rows[..., 0] * numpy.sin(rows[..., 5]),
                        rows[..., 3]], axis=-1)


def _rot(axis, ang):
    """Rotation matrices by angles ang about the unit vector axis (Rodrigues)."""
    k = numpy.asarray(axis, dtype=float) / numpy.linalg.norm(axis)
    K = numpy.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    ang = numpy.asarray(ang)[..., None, None]
    return numpy.eye(3) + numpy.sin(ang) * K + (1 - numpy.cos(ang)) * (K @ K)


def _free(n, seed):
    rng = numpy.random.default_rng(seed)
    X0 = rng.uniform(-1.5, 1.5, (n, 3))
    X0[:, 0] += 2.0
    V0 = rng.uniform(-0.5, 0.5, (n, 3))
    return X0, V0


@pytest.mark.parametrize("code", [6, 2])
@pytest.mark.parametrize("case", ["z", "tilted", "growing", "linear"])
def test_free_particle_in_a_non_inertial_frame(case, code):
    from galpy_amd import potential as P

    n, t = 16, numpy.linspace(0.0, 3.0, 31)
    X0, V0 = _free(n, 7)
    if case == "z":
        Om = 1.3
        f = P.NonInertialFrameForce(Omega=Om)
        ang, axis = Om * t, [0, 0, 1]
        Ovec = numpy.array([0.0, 0.0, Om])
    elif case == "tilted":
        Ovec = numpy.array([0.3, -0.4, 1.1])
        f = P.NonInertialFrameForce(Omega=list(Ovec))
        ang, axis = numpy.linalg.norm(Ovec) * t, Ovec
    elif case == "growing":  # Omega = 0.8 + 0.5 t about z: angle 0.8 t + 0.25 t^2
        f = P.NonInertialFrameForce(Omega=0.8, Omegadot=0.5)
        ang, axis = 0.8 * t + 0.25 * t**2, [0, 0, 1]
        Ovec = numpy.array([0.0, 0.0, 0.8])
    else:
        a0 = numpy.array([0.2, -0.1, 0.3])
        f = P.NonInertialFrameForce(a0=list(a0))
    npot, pt, pa, tf = P.parse_pot([f], t=t)
    if case == "linear":
        want = X0[:, None, :] + V0[:, None, :] * t[None, :, None] - 0.5 * a0 * t[None, :, None] ** 2
        r0, v0 = X0, V0
    else:
        Rm = _rot(axis, -ang)  # inertial -> frame: rotate back by the frame's angle
        xin = X0[:, None, :] + V0[:, None, :] * t[None, :, None]
        want = numpy.einsum("tij,ntj->nti", Rm, xin)
        r0, v0 = X0, V0 - numpy.cross(Ovec, X0)
    with oracle_tfuncs(tf):
        res, err, *_ = oracle_integrate(pt, pa, _cyl(r0, v0), t, code)
    assert numpy.all(err == 0)
    assert numpy.abs(_xyz(res) - want).max() < (1e-9 if code == 6 else 1e-7)


def test_tfunc_and_spline_packings_agree():
    """pot_type 39 (time functions called) and 45 (natural splines of them,
    3000 points) integrate the same orbits to the splines' accuracy."""
    from galpy_amd import potential as P

    t = numpy.linspace(0.0, 4.0, 41)
    Om = lambda tt: 0.9 + 0.2 * numpy.sin(tt)
    Omd = lambda tt: 0.2 * numpy.cos(tt)
    a0 = [lambda tt: 0.1 * numpy.cos(tt), lambda tt: 0.05 * tt, lambda tt: 0.0 * tt]
    mw = P.MWPotential2014_factory()
    kw = dict(Omega=Om, Omegadot=Omd, a0=a0, x0=[lambda tt: 0.0 * tt] * 3, v0=[lambda tt: 0.0 * tt] * 3)
    f39 = P.NonInertialFrameForce(cinterp=False, **kw)
    f45 = P.NonInertialFrameForce(cinterp=True, **kw)
    _, pt39, pa39, tf39 = P.parse_pot(mw + [f39], t=t)
    _, pt45, pa45, tf45 = P.parse_pot(mw + [f45], t=t)
    assert list(pt39[-1:]) == [39] and list(pt45[-1:]) == [45] and len(tf39) == 3 + 6 + 2 and tf45 == []
    X0, V0 = _free(8, 9)
    yo = _cyl(X0 * 0.5, V0 + numpy.array([0.0, 1.0, 0.0]))
    with oracle_tfuncs(tf39):
        r39, e39, *_ = oracle_integrate(pt39, pa39, yo, t, 6)
    r45, e45, *_ = oracle_integrate(pt45, pa45, yo, t, 6)
    assert numpy.all(e39 == 0) and numpy.all(e45 == 0)
    assert numpy.abs(_xyz(r39) - _xyz(r45)).max() < 1e-9
