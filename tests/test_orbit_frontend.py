"""The Orbit front end (galpy/orbit/Orbits.py mirror): argument checks on the
CPU, the galpy paper's published numbers and Orbits == single Orbit on GPU."""

import numpy
import pytest

from galpy_amd import MWPotential2014, Orbit


def test_rejects_unknown_method_and_bad_dt():
    o = Orbit([1.0, 0.1, 1.1, 0.0, 0.1, 0.0])
    ts = numpy.linspace(0, 1, 11)
    with pytest.raises(ValueError):
        o.integrate(ts, MWPotential2014, method="symplec4_c")
    with pytest.raises(ValueError):
        o.integrate(ts, MWPotential2014, method="symplec4_hip", dt=0.03)
    with pytest.raises(ValueError):
        o.integrate(numpy.array([0.0, 0.1, 0.3]), MWPotential2014, method="symplec4_hip")
    with pytest.raises(AttributeError):
        o.getOrbit()


def test_unsupported_potential_raises():
    from galpy_amd.potential import Potential

    class Custom(Potential):
        pass

    o = Orbit([1.0, 0.1, 1.1, 0.0, 0.1, 0.0])
    with pytest.raises(NotImplementedError):
        o.integrate(numpy.linspace(0, 1, 11), [Custom()], method="symplec4_hip")


@pytest.mark.gpu
def test_galpypaper_orbit_methods():
    """galpy tests/test_galpypaper.py:293-360 through the Orbit API."""
    o = Orbit([0.8, 0.3, 0.75, 0.0, 0.2, 0.0])
    times = numpy.linspace(0.0, 10.0, 1001)
    o.integrate(times, MWPotential2014, method="symplec4_hip")
    assert numpy.fabs(o.E(0.0) + 1.2547650648697966) < 1e-5
    assert numpy.all(numpy.fabs(o.L(0.0) - numpy.array([0.0, -0.16, 0.6])) < 1e-5)
    assert numpy.fabs(o.Jacobi(0.0, OmegaP=0.65) + 1.64476506) < 1e-5
    assert numpy.fabs(o.ER(times[-1]) + 1.27601734263047) < 1e-5
    assert numpy.fabs(o.Ez(times[-1]) - 0.021252201847851909) < 1e-5
    assert numpy.fabs(o.rperi() - 0.44231993168097) < 1e-5
    assert numpy.fabs(o.rap() - 0.87769030382105) < 1e-5
    assert numpy.fabs(o.zmax() - 0.077452357289016) < 1e-5
    assert numpy.fabs(o.e() - 0.32982348199330563) < 1e-5
    assert numpy.fabs(8.0 * o.R(2.0) - 3.5470772876920007) < 1e-3
    assert numpy.fabs(220.0 * o.vR(5.0) - 45.202530965094553) < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["symplec4_hip", "leapfrog_hip", "dop853_hip"])
def test_orbits_equal_single_orbit(method):
    """galpy tests/test_orbits.py:931-1020: integrating N orbits together
    gives the single-orbit results (here bit-identical: lanes are independent)."""
    from galpy_amd.workloads import ics

    vxvv = ics(5, seed=21)
    times = numpy.linspace(0.0, 10.0, 1001)
    os_ = Orbit(vxvv)
    os_.integrate(times, MWPotential2014, method=method)
    for i in range(len(vxvv)):
        o = Orbit(vxvv[i])
        o.integrate(times, MWPotential2014, method=method)
        assert numpy.array_equal(o.getOrbit(), os_.getOrbit()[i])
    E = os_.E()
    assert numpy.all(numpy.fabs(E / E[:, :1] - 1) < 1e-8)
