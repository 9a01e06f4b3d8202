"""TimeDependentAmplitudeWrapperPotential (-9) on the GPU: A(t) is a Python
function, tabulated by the library on the host (the C-ABI's pot_tfuncs)
and read from the table on the device; orbits through the host-buffer entry
point (full and planar) and through a device-resident table, against the
oracle, which calls A itself as galpy's C does."""

import math

import numpy
import pytest

from helpers import (assert_parity, ics, oracle_integrate, oracle_tfuncs, planar_ics, ulp_sensitivity)

pytestmark = pytest.mark.gpu


# the oracle calls these Python callbacks as galpy's C does; it runs on one
# thread here (callbacks from several threads queue on the interpreter lock)
def A(t):
    return 1.0 + 0.3 * math.sin(0.7 * t) + 0.05 * t


def _pot():
    from galpy_amd import potential as P

    return [P.TimeDependentAmplitudeWrapperPotential(amp=1.2, A=A, pot=P.MWPotential2014_factory()),
            P.SolidBodyRotationWrapperPotential(
                pot=P.TimeDependentAmplitudeWrapperPotential(A=lambda t: math.exp(-0.5 * (t - 4.0) ** 2),
                                                             pot=P.LogarithmicHaloPotential(normalize=0.1, b=0.8)),
                omega=0.4)]


@pytest.mark.parametrize("method,code,bar,factor", [("symplec4_hip", 3, 1e-10, 1e3),
                                                     ("dop853_hip", 6, 1e-8, 10.0)])
def test_full_orbits_host_buffers(method, code, bar, factor):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = _pot()
    _, pt, pa, tf = P.parse_pot(pot)
    assert len(tf) == 2
    t = numpy.linspace(0.0, 10.0, 101)
    yo = ics(256, seed=101)
    res, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
    assert numpy.all(err == 0)
    with oracle_tfuncs(tf):
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code, nthreads=1)[0], yo)
    assert_parity(res, ref, bar, sens, factor=factor, label=method)


def test_planar_orbits_host_buffers():
    from galpy_amd import potential as P
    from galpy_amd.integrate import integratePlanarOrbit_hip

    pot = _pot()
    _, pt, pa, tf = P.parse_pot_planar(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yp = planar_ics(256, seed=103)
    res, err = integratePlanarOrbit_hip(pot, yp.copy(), t, "dop853_hip")
    assert numpy.all(err == 0)
    with oracle_tfuncs(tf):
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 6, planar=True, nthreads=1)[0], yp)
    assert_parity(res, ref, 1e-8, sens, factor=10.0, label="planar dop853")


def test_device_table_and_missing_range():
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import PotentialTable, integrate_device

    pot = _pot()
    with pytest.raises(ValueError):
        PotentialTable(pot)
    _, pt, pa, tf = P.parse_pot(pot)
    t = numpy.linspace(0.0, 10.0, 101)
    yo = ics(256, seed=105)
    tab = PotentialTable(pot, trange=(0.0, 10.0))
    res = integrate_device(tab, torch.tensor(yo, device="cuda"), torch.tensor(t, device="cuda"), "dop853_hip")
    rows = res[0].cpu().numpy() if isinstance(res, tuple) else res.cpu().numpy()
    with oracle_tfuncs(tf):
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, 6, nthreads=1)[0], yo)
    assert_parity(rows, ref, 1e-8, sens, factor=10.0, label="device table")
