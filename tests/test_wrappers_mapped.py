"""The wrappers that move the point where the wrapped potential is
evaluated: RotateAndTiltWrapperPotential (-8), KuzminLikeWrapperPotential
(-10) and CylindricallySeparablePotentialWrapper (-12).  The oracle restates
their C files (recursing into the wrapped components as galpy does); pinned
by closed forms that do not go through any restatement:
  * KuzminLike around a point mass is the Kuzmin disk (b = 0) and the
    Miyamoto-Nagai disk (b > 0);
  * RotateAndTilt about z alone is the static SolidBodyRotation by the same
    angle; a tilt evaluates the potential at the point turned by the rotation
    that takes zvec to the z axis (scipy's Rotation here), and its forces are
    minus the gradient of its potential (finite differences), which pins the
    back-rotation of the forces;
  * CylindricallySeparable is Phi(R, 0) + Phi(Rp, z) - Phi(Rp, 0) of the
    wrapped closed form, its forces the partial derivatives of that."""

import numpy
import pytest

from helpers import oracle_forces, oracle_planar_forces, oracle_potential


def _pack(pot, planar=False):
    from galpy_amd.potential import parse_pot, parse_pot_planar

    _, pt, pa, _ = (parse_pot_planar if planar else parse_pot)(pot)
    return pt, pa


def _pts(seed, n=32):
    rng = numpy.random.default_rng(seed)
    return rng.uniform(0.2, 2.5, n), rng.uniform(-1.0, 1.0, n), rng.uniform(0.0, 2.0 * numpy.pi, n)


def test_kuzminlike_of_a_point_mass_is_kuzmin_and_miyamoto_nagai():
    from galpy_amd import potential as P

    R, z, phi = _pts(1)
    kep = P.KeplerPotential(amp=1.3)
    for b, ref in ((0.0, P.KuzminDiskPotential(amp=1.3, a=0.7)),
                   (0.4, P.MiyamotoNagaiPotential(amp=1.3, a=0.7, b=0.4))):
        w = P.KuzminLikeWrapperPotential(pot=kep, a=0.7, b=b)
        pt, pa = _pack([w])
        rpt, rpa = _pack([ref])
        FR, Fz, Fp = oracle_forces(pt, pa, R, z, phi)
        wR, wz, _ = oracle_forces(rpt, rpa, R, z, phi)
        assert numpy.allclose(FR, wR, rtol=1e-13, atol=0) and numpy.allclose(Fz, wz, rtol=1e-13, atol=1e-300)
        assert numpy.all(Fp == 0.0)
        assert numpy.allclose(oracle_potential(pt, pa, R, z, phi), oracle_potential(rpt, rpa, R, z, phi),
                              rtol=1e-13, atol=0)
        # host class and planar packing
        assert numpy.allclose(w.Rforce(R, z, phi), FR, rtol=1e-13) and numpy.allclose(w.zforce(R, z, phi), Fz,
                                                                                       rtol=1e-13)
        ppt, ppa = _pack([w], planar=True)
        pFR, _ = oracle_planar_forces(ppt, ppa, R, phi)
        assert numpy.allclose(pFR, oracle_forces(pt, pa, R, 0.0 * R, phi)[0], rtol=1e-14, atol=0)
    with pytest.raises(RuntimeError):
        P.KuzminLikeWrapperPotential(pot=P.LogarithmicHaloPotential(b=0.8), a=1.0)


def test_rotate_and_tilt():
    from scipy.spatial.transform import Rotation

    from galpy_amd import potential as P

    R, z, phi = _pts(2)
    halo = P.LogarithmicHaloPotential(normalize=1.0, q=0.8, b=0.7)
    # a turn about z alone is a static solid-body rotation by the same angle
    rt = P.RotateAndTiltWrapperPotential(pot=halo, galaxy_pa=0.4, amp=1.1)
    sb = P.SolidBodyRotationWrapperPotential(pot=halo, omega=0.0, pa=0.4, amp=1.1)
    pt, pa = _pack([rt])
    spt, spa = _pack([sb])
    for a, b in zip(oracle_forces(pt, pa, R, z, phi), oracle_forces(spt, spa, R, z, phi)):
        assert numpy.allclose(a, b, rtol=1e-12, atol=1e-14)
    # a tilt: Phi(x) = Phi_disk(M x + offset), M the rotation taking zvec to e_z
    zvec = numpy.array([0.3, -0.5, 0.8])
    off = numpy.array([0.1, -0.2, 0.05])
    disk = P.MiyamotoNagaiPotential(a=0.5, b=0.1, normalize=1.0)
    rt = P.RotateAndTiltWrapperPotential(pot=disk, zvec=zvec, offset=off, amp=0.9)
    pt, pa = _pack([rt])
    M = Rotation.align_vectors([[0.0, 0.0, 1.0]], [zvec / numpy.linalg.norm(zvec)])[0].as_matrix()
    X = numpy.stack([R * numpy.cos(phi), R * numpy.sin(phi), z], axis=1)
    Xp = X @ M.T + off
    want = 0.9 * disk(numpy.hypot(Xp[:, 0], Xp[:, 1]), Xp[:, 2])
    V = oracle_potential(pt, pa, R, z, phi)
    assert numpy.allclose(V, want, rtol=1e-12, atol=0)
    # forces = -grad Phi (fourth-order differences in R, z, phi)
    h = 1e-4

    def d(i):
        e = numpy.zeros(3)
        e[i] = h
        f = lambda s: oracle_potential(pt, pa, R + s * e[0], z + s * e[1], phi + s * e[2])
        return (-f(2) + 8 * f(1) - 8 * f(-1) + f(-2)) / (12 * h)

    FR, Fz, Fp = oracle_forces(pt, pa, R, z, phi)
    for F, i in ((FR, 0), (Fz, 1), (Fp, 2)):
        assert numpy.allclose(F, -d(i), rtol=1e-8, atol=1e-10)
    # host class
    assert numpy.allclose(rt(R, z, phi), V, rtol=1e-12)
    assert numpy.allclose(rt.Rforce(R, z, phi), FR, rtol=1e-12, atol=1e-14)
    assert numpy.allclose(rt.phitorque(R, z, phi), Fp, rtol=1e-12, atol=1e-14)
    assert numpy.allclose(rt.zforce(R, z, phi), Fz, rtol=1e-12, atol=1e-14)
    # full orbits only
    with pytest.raises(NotImplementedError):
        _pack([rt], planar=True)


def test_cylindrically_separable():
    from galpy_amd import potential as P

    R, z, phi = _pts(3)
    disk = P.MiyamotoNagaiPotential(a=0.5, b=0.2, normalize=1.0)
    w = P.CylindricallySeparablePotentialWrapper(pot=disk, Rp=1.3, amp=1.2)
    pt, pa = _pack([w])
    mn = lambda RR, zz: -disk._amp / numpy.sqrt(RR**2 + (0.5 + numpy.sqrt(zz**2 + 0.04)) ** 2)
    want = 1.2 * (mn(R, 0.0) + mn(1.3, z) - mn(1.3, 0.0))
    V = oracle_potential(pt, pa, R, z, phi)
    assert numpy.allclose(V, want, rtol=1e-13, atol=1e-15)
    h = 1e-4
    dR = (-mn(R + 2 * h, 0) + 8 * mn(R + h, 0) - 8 * mn(R - h, 0) + mn(R - 2 * h, 0)) / (12 * h)
    dz = (-mn(1.3, z + 2 * h) + 8 * mn(1.3, z + h) - 8 * mn(1.3, z - h) + mn(1.3, z - 2 * h)) / (12 * h)
    FR, Fz, Fp = oracle_forces(pt, pa, R, z, phi)
    assert numpy.allclose(FR, -1.2 * dR, rtol=1e-9) and numpy.allclose(Fz, -1.2 * dz, rtol=1e-9, atol=1e-12)
    assert numpy.all(Fp == 0.0)
    assert numpy.allclose(w(R, z, phi), V, rtol=1e-13)
    assert numpy.allclose(w.Rforce(R, z, phi), FR, rtol=1e-13) and numpy.allclose(w.zforce(R, z, phi), Fz,
                                                                                   rtol=1e-13, atol=1e-300)
    ppt, ppa = _pack([w], planar=True)
    pFR, _ = oracle_planar_forces(ppt, ppa, R, phi)
    assert numpy.allclose(pFR, FR, rtol=1e-14, atol=0)


def test_nested_mapped_wrappers():
    """A solid-body rotation around a rotated-and-tilted triaxial halo plus a
    Kuzmin-like disk: the oracle's recursion against the host classes."""
    from galpy_amd import potential as P

    R, z, phi = _pts(4)
    t = numpy.linspace(-1.0, 2.0, R.size)
    halo = P.LogarithmicHaloPotential(normalize=0.6, q=0.8, b=0.7)
    inner = [P.RotateAndTiltWrapperPotential(pot=halo, zvec=[0.2, 0.1, 1.0], galaxy_pa=0.3),
             P.KuzminLikeWrapperPotential(pot=P.HernquistPotential(normalize=0.3), a=0.6, b=0.2)]
    pot = [P.SolidBodyRotationWrapperPotential(pot=inner, omega=0.5, pa=0.2),
           P.CylindricallySeparablePotentialWrapper(pot=P.NFWPotential(normalize=0.2), Rp=1.1)]
    pt, pa = _pack(pot)
    FR, Fz, Fp = oracle_forces(pt, pa, R, z, phi, t)
    wR = sum(p.Rforce(R, z, phi, t) for p in pot)
    wz = sum(p.zforce(R, z, phi, t) for p in pot)
    wp = sum(p.phitorque(R, z, phi, t) for p in pot)
    assert numpy.allclose(FR, wR, rtol=1e-12, atol=1e-14)
    assert numpy.allclose(Fz, wz, rtol=1e-12, atol=1e-14)
    assert numpy.allclose(Fp, wp, rtol=1e-12, atol=1e-14)
    assert numpy.allclose(oracle_potential(pt, pa, R, z, phi, t), P.evaluatePotentials(pot, R, z, phi, t),
                          rtol=1e-12)
