"""NonInertialFrameForce (39: the caller's time functions, tabulated on the
host; 45: natural splines of them) on the GPU, with MWPotential2014: full
orbits through RK4, RK6, DOPR54 and DOP853 (the stage velocities handed to
the force) and planar orbits against the oracle; the symplectic methods are
refused for a velocity-dependent force."""

import numpy
import pytest

from helpers import assert_parity, ics, oracle_integrate, oracle_tfuncs, planar_ics, ulp_sensitivity

pytestmark = pytest.mark.gpu


def _force(which):
    from galpy_amd import potential as P

    if which == "const":
        return P.NonInertialFrameForce(Omega=[0.1, -0.05, 0.9], Omegadot=[0.0, 0.01, 0.02])
    kw = dict(Omega=lambda tt: 0.9 + 0.2 * numpy.sin(tt), Omegadot=lambda tt: 0.2 * numpy.cos(tt),
              a0=[lambda tt: 0.1 * numpy.cos(tt), lambda tt: 0.05 * numpy.sin(0.5 * tt), lambda tt: 0.01 * tt],
              x0=[lambda tt: 0.02 * tt, lambda tt: 0.0 * tt, lambda tt: 0.01 * numpy.sin(tt)],
              v0=[lambda tt: 0.02 + 0.0 * tt, lambda tt: 0.0 * tt, lambda tt: 0.01 * numpy.cos(tt)])
    return P.NonInertialFrameForce(cinterp=(which == "spline"), **kw)


@pytest.mark.parametrize("which", ["const", "tfunc", "spline"])
@pytest.mark.parametrize("method,code", [("rk4_hip", 1), ("rk6_hip", 2), ("dopr54_hip", 5), ("dop853_hip", 6)])
def test_full_orbits(which, method, code):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = P.MWPotential2014_factory() + [_force(which)]
    t = numpy.linspace(0.0, 8.0, 81)
    _, pt, pa, tf = P.parse_pot(pot, t=t)
    yo = ics(128, seed=391)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
    assert numpy.all(err == 0)
    # the rotating frame flings some orbits outwards, where an adaptive step
    # can sit at the accept/reject threshold: the oracle's own result then
    # jumps by ~6e-8 under a few-ulp change of the initial state (seen for one
    # of these orbits in a quarter of such changes), so the sensitivity is
    # sampled with eight sign patterns
    with oracle_tfuncs(tf):
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code, nthreads=1)[0], yo,
                                    nrep=8 if code in (5, 6) else 3)
    bar, factor = (1e-8, 10.0) if code in (5, 6) else (1e-10, 1e3)
    assert_parity(res, ref, bar, sens, factor=factor, label=f"{which} {method}")


@pytest.mark.parametrize("which", ["const", "spline"])
def test_planar_orbits(which):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = P.MWPotential2014_factory() + [_force(which)]
    t = numpy.linspace(0.0, 8.0, 81)
    _, pt, pa, tf = P.parse_pot_planar(pot, t=t)
    yp = planar_ics(128, seed=393)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "dop853_hip")
    assert numpy.all(err == 0)
    with oracle_tfuncs(tf):
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 6, planar=True, nthreads=1)[0], yp)
    assert_parity(res, ref, 1e-8, sens, factor=10.0, label=f"{which} planar")


def test_symplectic_methods_refused():
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = P.MWPotential2014_factory() + [_force("const")]
    with pytest.raises(RuntimeError, match="velocity-dependent"):
        integrateFullOrbit_hip(pot, ics(4, seed=1), numpy.linspace(0.0, 1.0, 11), "symplec4_hip")
