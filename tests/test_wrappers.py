"""Wrapper potentials on the host and in the oracle: galpy's nested packing
(integrateFullOrbit.py:396-448, integratePlanarOrbit.py:507-565), the oracle's
restatement of the wrappers' C functions (DehnenSmoothWrapperPotential.c,
SolidBodyRotationWrapperPotential.c, CorotatingRotationWrapperPotential.c,
GaussianAmplitudeWrapperPotential.c) against the Python classes, and the
properties that fix each wrapper independently of either restatement: the
ends and midpoint of Dehnen's quintic, grow + decay = 1, the Gaussian's value
one sigma out, the direction of the rotation (forces of the rotated pattern
equal the unrotated forces at phi - omega t - pa), the corotating wrapper's
radial force as -dPhi/dR of its own potential, its reduction to solid-body
rotation at beta = 1, the Jacobi integral E - omega Lz of a pattern rotating
rigidly, and free motion before a wrapped potential switches on."""

import numpy
import pytest

from helpers import npot_of, oracle_forces, oracle_integrate, oracle_planar_forces, oracle_potential


def _P():
    from galpy_amd import potential as P

    return P


def _pk(pot, planar=False):
    from galpy_amd.potential import parse_pot, parse_pot_planar

    n, pt, pa, _ = (parse_pot_planar if planar else parse_pot)(pot)
    assert n == npot_of(pt, pa, planar)
    return pt, pa


def _triax():
    return _P().LogarithmicHaloPotential(q=0.9, b=0.8)


def _points(n=400, seed=0):
    rng = numpy.random.default_rng(seed)
    return (rng.uniform(0.2, 3.0, n), rng.normal(0.0, 0.4, n), rng.uniform(0, 2 * numpy.pi, n),
            rng.uniform(-6.0, 6.0, n))


def _close(a, b, tol=1e-13):
    return numpy.abs(a - b).max() <= tol * max(numpy.abs(b).max(), 1e-300)


def test_nested_packing_layout():
    P = _P()
    inner = P.SolidBodyRotationWrapperPotential(pot=_triax(), omega=0.5, pa=0.2)
    w = P.DehnenSmoothWrapperPotential(amp=0.7, pot=[inner, P.HernquistPotential(a=1.3)], tform=-2.0,
                                       tsteady=1.5, decay=True)
    from galpy_amd.potential import parse_pot

    n, pt, pa, _ = parse_pot([P.MiyamotoNagaiPotential(a=0.5, b=0.1), w])
    assert n == 2
    assert list(pt) == [5, -1, -2, 0, 8]
    lh = _triax().c_args()
    want = ([1.0, 0.5, 0.1] + [2.0] + [1.0] + lh + [1.0, 0.5, 0.2] + [1.0, 1.3]
            + [0.7, -2.0, -0.5, 0.0])
    numpy.testing.assert_array_equal(pa, want)
    assert npot_of(pt, pa) == 2


def test_oracle_matches_python_classes():
    """Two restatements of the same definitions (numpy classes, C++ oracle)
    agree at random points and times, nested wrappers included."""
    P = _P()
    pots = [
        [P.DehnenSmoothWrapperPotential(pot=P.PlummerPotential(b=0.5), tform=-3.0, tsteady=4.0)],
        [P.DehnenSmoothWrapperPotential(pot=_triax(), tform=-3.0, tsteady=4.0, decay=True)],
        [P.GaussianAmplitudeWrapperPotential(amp=1.3, pot=P.HernquistPotential(), to=1.0, sigma=2.0)],
        [P.SolidBodyRotationWrapperPotential(pot=_triax(), omega=0.7, pa=0.3)],
        [P.CorotatingRotationWrapperPotential(pot=_triax(), vpo=1.1, beta=-0.2, to=0.5, pa=0.1)],
        [P.DehnenSmoothWrapperPotential(pot=[P.SolidBodyRotationWrapperPotential(pot=_triax(), omega=0.4),
                                             P.MiyamotoNagaiPotential(a=0.5, b=0.1)], tform=-1.0),
         P.NFWPotential(a=2.0)],
        [P.CorotatingRotationWrapperPotential(
            pot=P.GaussianAmplitudeWrapperPotential(pot=P.DehnenBarPotential(omegab=1.85, rb=0.8, Af=0.01),
                                                    to=0.0, sigma=3.0), vpo=0.9, beta=0.1)],
    ]
    R, z, phi, t = _points()
    for pot in pots:
        pt, pa = _pk(pot)
        FR, Fz, Fp = oracle_forces(pt, pa, R, z, phi, t)
        want = [sum(p.Rforce(R, z, phi, t) for p in pot), sum(p.zforce(R, z, phi, t) for p in pot),
                sum(p.phitorque(R, z, phi, t) for p in pot)]
        for got, w in zip((FR, Fz, Fp), want):
            assert _close(got, w, 1e-12)
        assert _close(oracle_potential(pt, pa, R, z, phi, t), sum(p(R, z, phi, t) for p in pot), 1e-12)
        # the planar packing: the full forces in the plane
        ppt, ppa = _pk(pot, planar=True)
        PR, Pp = oracle_planar_forces(ppt, ppa, R, phi, t)
        FR0, _, Fp0 = oracle_forces(pt, pa, R, 0.0 * R, phi, t)
        assert _close(PR, FR0, 4e-15) and _close(Pp, Fp0, 4e-15)


def test_dehnen_smooth_quintic_and_decay():
    P = _P()
    base = [P.PlummerPotential(b=0.5)]
    tform, tsteady = -2.0, 2.0
    R = numpy.full(5, 1.3)
    z = numpy.full(5, 0.2)
    t = numpy.array([tform - 1.0, tform, 0.0, tsteady, tsteady + 1.0])
    FR0, _, _ = oracle_forces(*_pk(base), R, z, 0.0 * R, t)
    grow = [P.DehnenSmoothWrapperPotential(amp=1.0, pot=base, tform=tform, tsteady=tsteady - tform)]
    decay = [P.DehnenSmoothWrapperPotential(amp=1.0, pot=base, tform=tform, tsteady=tsteady - tform, decay=True)]
    FRg, _, _ = oracle_forces(*_pk(grow), R, z, 0.0 * R, t)
    FRd, _, _ = oracle_forces(*_pk(decay), R, z, 0.0 * R, t)
    numpy.testing.assert_allclose(FRg / FR0, [0.0, 0.0, 0.5, 1.0, 1.0], atol=1e-15)
    numpy.testing.assert_allclose(FRg + FRd, FR0, rtol=1e-15)
    # zero slope at both ends of the ramp
    eps = 1e-4
    te = numpy.array([tform + eps, tsteady - eps])
    FRe, _, _ = oracle_forces(*_pk(grow), R[:2], z[:2], 0.0 * te, te)
    assert FRe[0] / FR0[0] < 1e-11 and 1.0 - FRe[1] / FR0[0] < 1e-11


def test_gaussian_amplitude():
    P = _P()
    base = [P.HernquistPotential(a=0.8)]
    w = [P.GaussianAmplitudeWrapperPotential(amp=2.0, pot=base, to=1.5, sigma=0.5)]
    R, z = numpy.full(3, 1.1), numpy.full(3, -0.3)
    t = numpy.array([1.5, 1.0, 2.0])
    FR0, Fz0, _ = oracle_forces(*_pk(base), R, z, 0.0 * R, t)
    FR, Fz, _ = oracle_forces(*_pk(w), R, z, 0.0 * R, t)
    numpy.testing.assert_allclose(FR / FR0, [2.0, 2.0 * numpy.exp(-0.5), 2.0 * numpy.exp(-0.5)], rtol=1e-14)
    numpy.testing.assert_allclose(Fz / Fz0, FR / FR0, rtol=1e-14)


def test_rotation_direction_and_corotating_reduction():
    P = _P()
    R, z, phi, t = _points(300, seed=1)
    omega, pa = 0.6, 0.25
    rot = [P.SolidBodyRotationWrapperPotential(pot=_triax(), omega=omega, pa=pa)]
    FR, Fz, Fp = oracle_forces(*_pk(rot), R, z, phi, t)
    G = oracle_forces(*_pk([_triax()]), R, z, phi - omega * t - pa, t)
    for a, b in zip((FR, Fz, Fp), G):
        assert _close(a, b, 1e-14)
    # corotating with beta = 1: offset vpo (t - to) + pa, a solid-body rotation
    vpo, to = 0.6, 1.0
    cor = [P.CorotatingRotationWrapperPotential(pot=_triax(), vpo=vpo, beta=1.0, to=to, pa=pa + vpo * to)]
    C = oracle_forces(*_pk(cor), R, z, phi, t)
    for a, b in zip(C, (FR, Fz, Fp)):
        assert _close(a, b, 1e-13)


def test_corotating_radial_force_is_minus_dphi_dR():
    """The chain-rule term of CorotatingRotationWrapperPotential.c:3-15: the
    radial force is -dPhi/dR of the wrapper's own potential (fourth-order
    central differences of the oracle's potential, h = 1e-3)."""
    P = _P()
    cor = [P.CorotatingRotationWrapperPotential(pot=_triax(), vpo=1.2, beta=0.5, to=-1.0, pa=0.4)]
    pt, pa = _pk(cor)
    rng = numpy.random.default_rng(2)
    n = 200
    R, z = rng.uniform(0.5, 3.0, n), rng.normal(0.0, 0.4, n)
    phi, t = rng.uniform(0, 2 * numpy.pi, n), rng.uniform(-2.0, 2.0, n)
    FR, Fz, _ = oracle_forces(pt, pa, R, z, phi, t)
    h = 1e-3

    def d(f):
        return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)

    dR = d(lambda e: oracle_potential(pt, pa, R + e, z, phi, t))
    dz = d(lambda e: oracle_potential(pt, pa, R, z + e, phi, t))
    assert numpy.abs(FR + dR).max() < 1e-9 * numpy.abs(FR).max()
    assert numpy.abs(Fz + dz).max() < 1e-9 * numpy.abs(Fz).max()
    # the term matters: without it the force is off by far more
    off = 1.2 * R**-0.5 * (t + 1.0) + 0.4
    base = oracle_forces(*_pk([_triax()]), R, z, phi - off, t)[0]
    assert numpy.abs(base + dR).max() > 1e-3


@pytest.mark.parametrize("code", [3, 6])
def test_rigidly_rotating_pattern_conserves_jacobi(code):
    """A triaxial halo turning at omega is static in the rotating frame:
    E - omega Lz is conserved (with the opposite sense of rotation it would
    be E + omega Lz)."""
    from galpy_amd.potential import evaluatePotentials
    from helpers import ics

    P = _P()
    omega = 0.5
    rot = [P.SolidBodyRotationWrapperPotential(pot=_triax(), omega=omega, pa=0.3)]
    yo = ics(16, seed=3)
    t = numpy.linspace(0.0, 20.0, 201)
    res, err, _, _ = oracle_integrate(*_pk(rot), yo, t, code, dt=0.001 if code == 3 else None)
    assert numpy.all(err == 0)
    R, vR, vT, z, vz, phi = numpy.moveaxis(res, -1, 0)
    E = evaluatePotentials(rot, R, z, phi, t[None, :]) + 0.5 * (vR**2 + vT**2 + vz**2)
    EJ = E - omega * R * vT
    assert numpy.abs(EJ - EJ[:, :1]).max() < 1e-7 * numpy.abs(EJ).max()
    # neither E alone nor E + omega Lz (the opposite sense) is conserved
    EP = E + omega * R * vT
    assert numpy.abs(E - E[:, :1]).max() > 1e-4
    assert numpy.abs(EP - EP[:, :1]).max() > 1e-4


def test_free_motion_before_switch_on():
    """Before tform a growing wrapper exerts no force: orbits are straight
    lines, also when the wrapped components are MWPotential2014's."""
    from helpers import cyl_to_rect, ics

    P = _P()
    w = [P.DehnenSmoothWrapperPotential(pot=P.MWPotential2014_factory(), tform=50.0)]
    yo = ics(8, seed=4)
    t = numpy.linspace(0.0, 5.0, 11)
    res, err, _, _ = oracle_integrate(*_pk(w), yo, t, 3, dt=0.01)
    x = cyl_to_rect(res.reshape(-1, 6)).reshape(8, 11, 6)
    x0 = cyl_to_rect(yo)
    want = x0[:, None, :3] + x0[:, None, 3:] * t[None, :, None]
    assert numpy.abs(x[..., :3] - want).max() < 1e-12


def test_paramspec_and_unknown_wrappers():
    P = _P()
    from galpy_amd.adjoint import ParamSpec

    with pytest.raises(ValueError):
        ParamSpec([P.DehnenSmoothWrapperPotential(pot=P.HernquistPotential())], [(0, 0)])
    with pytest.raises(ValueError):
        P.DehnenSmoothWrapperPotential()
