"""DOPR54 and IAS15 on persistent lanes claim their orbits longest first
(after a pilot run over the first 2% of the time range) when there are more
orbits than resident lanes.  The ordered launch of 600,000 MWPotential2014
orbits gives, bit for bit, the results of the same orbits integrated in
chunks of 50,000 (fewer orbits than lanes: no pilot, no order), so neither
the pilot nor the order changes an orbit; a sample of the orbits follows
the oracle to the adaptive methods' bar."""

import numpy
import pytest

from helpers import assert_parity, ics, oracle_integrate, ulp_sensitivity

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
@pytest.mark.parametrize("method,code", [("dopr54_hip", 5), ("ias15_hip", 7)])
def test_ordered_launch_matches_chunks(method, code):
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    pot = P.MWPotential2014_factory()
    _, pt, pa, _ = P.parse_pot(pot)
    t = numpy.linspace(0.0, 2.0, 5)
    yo = ics(600000, seed=41)
    print(f"{method}: 600000 orbits in one launch", flush=True)
    big, err = integrateFullOrbit_hip(pot, yo.copy(), t, method)
    assert numpy.all(err == 0)
    for a in range(0, yo.shape[0], 50000):
        part, perr = integrateFullOrbit_hip(pot, yo[a:a + 50000].copy(), t, method)
        assert numpy.all(perr == 0)
        assert numpy.array_equal(part, big[a:a + 50000]), (method, a)
    idx = numpy.random.default_rng(3).choice(yo.shape[0], 32, replace=False)
    ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, t, code)[0], yo[idx])
    assert_parity(big[idx], ref, 1e-8, sens, factor=10.0, label=method)
