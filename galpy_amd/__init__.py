"""MI355X (gfx950) orbit-integration engine with galpy's integration API.

    from galpy_amd import Orbit, MWPotential2014
    o = Orbit([[1.0, 0.1, 1.1, 0.0, 0.1, 0.0]])
    o.integrate(numpy.linspace(0, 10, 1001), MWPotential2014, method="dop853_hip")

Methods: leapfrog_hip, symplec4_hip, dop853_hip (FP64, hand-written HIP).
"""

from .orbit import Orbit, Orbits  # noqa: F401
from .potential import (  # noqa: F401
    DehnenBarPotential,
    HernquistPotential,
    LogarithmicHaloPotential,
    MiyamotoNagaiPotential,
    MWPotential2014,
    NFWPotential,
    PowerSphericalPotentialwCutoff,
    evaluatePotentials,
    evaluateRforces,
    evaluatezforces,
    parse_pot,
)
