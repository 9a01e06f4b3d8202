"""Synthetic workloads of the benchmark configurations (seeded, in-memory).

Initial conditions follow one law for every configuration: R~U(0.5,2),
z~N(0,0.1), phi~U(0,2 pi), vR,vz~N(0,0.1), vT~N(1,0.1), rows
(R,vR,vT,z,vz,phi) in galpy's natural units; planar workloads keep
(R,vR,vT,phi) of the same draws.
"""

import numpy

from . import potential as P


def ics(n, seed=0):
    rng = numpy.random.default_rng(seed)
    R = rng.uniform(0.5, 2.0, n)
    z = rng.normal(0.0, 0.1, n)
    phi = rng.uniform(0.0, 2.0 * numpy.pi, n)
    vR = rng.normal(0.0, 0.1, n)
    vz = rng.normal(0.0, 0.1, n)
    vT = rng.normal(1.0, 0.1, n)
    return numpy.stack([R, vR, vT, z, vz, phi], axis=1)


def mw2014():
    return P.MWPotential2014_factory()


def mw2014_reordered():
    """MWPotential2014's three components in another order (NFW halo,
    MiyamotoNagai disk, PowerSphericalPotentialwCutoff bulge): the same
    physics through the run-time dispatch kernel instead of the specialised
    one (the sum's order differs, so results agree to rounding)."""
    bulge, disk, halo = P.MWPotential2014_factory()
    return P.CompositePotential(halo, disk, bulge)


def loghalo_bar():
    """LogarithmicHaloPotential(q=0.9) + DehnenBarPotential(omegab=1.85, rb=0.8, Af=0.01)."""
    return P.CompositePotential(
        P.LogarithmicHaloPotential(q=0.9),
        P.DehnenBarPotential(omegab=1.85, rb=0.8, Af=0.01),
    )


def mn_plummer_burkert():
    """MWPotential2014's disk with a Plummer bulge (b = 0.3) and a Burkert
    halo (a = 2) in place of its bulge and NFW halo, carrying 0.6, 0.05 and
    0.35 of the radial force at R = 1 (vc(1) = 1)."""
    return P.CompositePotential(
        P.MiyamotoNagaiPotential(a=3.0 / 8.0, b=0.28 / 8.0, normalize=0.6),
        P.PlummerPotential(b=0.3, normalize=0.05),
        P.BurkertPotential(a=2.0, normalize=0.35),
    )


def mw2014_rotating_triaxial():
    """MWPotential2014 plus a triaxial logarithmic perturbation (amp 0.05,
    q = 0.9, b = 0.8) turning rigidly at omega = 0.6 through
    SolidBodyRotationWrapperPotential: time-dependent in the inertial frame,
    with the Jacobi integral E - 0.6 Lz conserved."""
    return P.CompositePotential(
        *P.MWPotential2014_factory(),
        P.SolidBodyRotationWrapperPotential(pot=P.LogarithmicHaloPotential(amp=0.05, q=0.9, b=0.8),
                                            omega=0.6, pa=0.0),
    )


POTENTIAL_NAMES = {
    mw2014: "MWPotential2014",
    loghalo_bar: "LogarithmicHalo(q=0.9)+DehnenBar",
    mn_plummer_burkert: "MiyamotoNagai(a=3/8,b=0.28/8)+Plummer(b=0.3)+Burkert(a=2)",
    mw2014_rotating_triaxial: "MWPotential2014+SolidBodyRotationWrapper(LogarithmicHalo(amp=0.05,q=0.9,b=0.8),omega=0.6)",
    mw2014_reordered: "MWPotential2014 (components as NFW, MiyamotoNagai, PowerSphericalwCutoff)",
}

WORKLOADS = {
    # 1,000 orbits, symplec4 with per-orbit step estimation, dense outputs
    "mw2014_symplec4_1k": dict(pot=mw2014, n=1000, t=(0.0, 100.0, 10001), method="symplec4_hip",
                              omega=0.0, seed=0),
    # 1,048,576 orbits, DOP853 at galpy's default tolerances
    "mw2014_dop853_1m": dict(pot=mw2014, n=1048576, t=(0.0, 500.0, 1001), method="dop853_hip",
                            omega=0.0, seed=0),
    # 1,048,576 orbits, symplec4 at a fixed step (the force kernel at scale)
    "mw2014_symplec4_1m": dict(pot=mw2014, n=1048576, t=(0.0, 100.0, 101), method="symplec4_hip",
                              omega=0.0, seed=0, dt=0.01),
    # 262,144 orbits in a rotating bar (Jacobi integral instead of energy)
    "loghalo_bar_dop853_256k": dict(pot=loghalo_bar, n=262144, t=(0.0, 500.0, 1001),
                                   method="dop853_hip", omega=1.85, seed=1),
    # planar orbits (integratePlanarOrbit): the DOP853 and fixed-step
    # configurations above in the plane z = 0
    "mw2014_planar_dop853_1m": dict(pot=mw2014, n=1048576, t=(0.0, 500.0, 1001),
                                   method="dop853_hip", omega=0.0, seed=0, planar=True),
    "mw2014_planar_symplec4_1m": dict(pot=mw2014, n=1048576, t=(0.0, 100.0, 101),
                                     method="symplec4_hip", omega=0.0, seed=0, dt=0.01,
                                     planar=True),
    # integrate_dxdv: 65,536 orbits with one phase-space deviation each
    # (standard-normal components, rectangular), symplec4 at dt = 0.01 for
    # 2,000 steps, the orbit and deviation stored every 10 steps
    "mw2014_dxdv_symplec4_65k": dict(pot=mw2014, n=65536, t=(0.0, 20.0, 201),
                                    method="symplec4_hip", omega=0.0, seed=0, dt=0.01, dxdv=True),
    "mw2014_dxdv_dop853_65k": dict(pot=mw2014, n=65536, t=(0.0, 20.0, 201),
                                  method="dop853_hip", omega=0.0, seed=0, dxdv=True),
    # a composition without a specialised kernel (run-time dispatch of the
    # component codes): fixed-step symplec4 at scale, and DOP853 at galpy's
    # default tolerances
    "mpb_symplec4_1m": dict(pot=mn_plummer_burkert, n=1048576, t=(0.0, 100.0, 101),
                           method="symplec4_hip", omega=0.0, seed=0, dt=0.01),
    "mpb_dop853_256k": dict(pot=mn_plummer_burkert, n=262144, t=(0.0, 500.0, 1001),
                           method="dop853_hip", omega=0.0, seed=0),
    # MWPotential2014 through the run-time dispatch kernels: the cost of
    # dispatching on component codes, against the specialised kernels on the
    # same orbits (mw2014_symplec4_1m, mw2014_dop853_1m at 262,144 orbits)
    "mw2014gen_symplec4_1m": dict(pot=mw2014_reordered, n=1048576, t=(0.0, 100.0, 101),
                                 method="symplec4_hip", omega=0.0, seed=0, dt=0.01),
    "mw2014gen_dop853_256k": dict(pot=mw2014_reordered, n=262144, t=(0.0, 500.0, 1001),
                                 method="dop853_hip", omega=0.0, seed=0),
    # the other fixed-step integrators of galpy at the force-kernel workload's
    # size: symplec6 (7 force evaluations per step), RK4 (4) and RK6 (7)
    "mw2014_symplec6_1m": dict(pot=mw2014, n=1048576, t=(0.0, 100.0, 101), method="symplec6_hip",
                              omega=0.0, seed=0, dt=0.01),
    "mw2014_rk4_1m": dict(pot=mw2014, n=1048576, t=(0.0, 100.0, 101), method="rk4_hip",
                         omega=0.0, seed=0, dt=0.01),
    "mw2014_rk6_1m": dict(pot=mw2014, n=1048576, t=(0.0, 100.0, 101), method="rk6_hip",
                         omega=0.0, seed=0, dt=0.01),
    # DOPR54 (galpy's 5(4) Dormand-Prince) at galpy's default tolerances, the
    # DOP853 workload's grid and potential at a quarter of the orbits
    "mw2014_dopr54_256k": dict(pot=mw2014, n=262144, t=(0.0, 500.0, 1001), method="dopr54_hip",
                              omega=0.0, seed=0),
    # IAS15 (galpy's 15th-order Gauss-Radau), the same grid and orbit count
    "mw2014_ias15_256k": dict(pot=mw2014, n=262144, t=(0.0, 500.0, 1001), method="ias15_hip",
                             omega=0.0, seed=0),
    # a wrapped (rotating) component: DOP853, drift of the Jacobi integral
    "mw2014_rottriax_dop853_256k": dict(pot=mw2014_rotating_triaxial, n=262144, t=(0.0, 500.0, 1001),
                                       method="dop853_hip", omega=0.6, seed=0),
}


def make(name, n=None):
    w = dict(WORKLOADS[name])
    if n is not None:
        w["n"] = n
    t0, t1, nt = w["t"]
    w["t"] = numpy.linspace(t0, t1, nt)
    w["yo"] = ics(w["n"], seed=w["seed"])
    w["planar"] = bool(w.get("planar", False))
    w["dxdv"] = bool(w.get("dxdv", False))
    if w["planar"]:
        w["yo"] = numpy.ascontiguousarray(w["yo"][:, [0, 1, 2, 5]])
    if w["dxdv"]:
        # rectangular (x,y,z,vx,vy,vz | dx,dy,dz,dvx,dvy,dvz)
        R, vR, vT, z, vz, phi = w["yo"].T
        c, s = numpy.cos(phi), numpy.sin(phi)
        rect = numpy.stack([R * c, R * s, z, vR * c - vT * s, vR * s + vT * c, vz], axis=1)
        dev = numpy.random.default_rng(w["seed"] + 1).normal(size=rect.shape)
        w["yo"] = numpy.ascontiguousarray(numpy.hstack([rect, dev]))
    w["potential"] = POTENTIAL_NAMES[w["pot"]]
    w["pot"] = w["pot"]()
    w["name"] = name
    return w
