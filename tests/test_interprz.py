"""interpRZPotential (pot_type 13): cubic B-spline interpolation on an
(R or log R, |z|) grid, as galpy's C evaluates it (Thevenaz's interpolation
with mirror boundaries).  The oracle's interpolation is pinned against
scipy.ndimage.map_coordinates (order 3, mirror, no prefilter) on random
coefficient grids; the interpolated MWPotential2014 reproduces the exact
forces to interpolation accuracy; the host class equals the oracle."""

import numpy
import pytest

from helpers import oracle_forces, oracle_potential


@pytest.mark.parametrize("logR", [False, True])
def test_interprz(logR):
    from scipy import ndimage

    from galpy_amd import potential as P

    mw = P.MWPotential2014_factory()
    rgrid = (numpy.log(0.05), numpy.log(5.0), 121) if logR else (0.05, 5.0, 121)
    ip = P.interpRZPotential(RZPot=mw, rgrid=rgrid, zgrid=(0.0, 1.0, 81), logR=logR)
    _, pt, pa, _ = P.parse_pot([ip])
    assert list(pt) == [13] and len(pa) == 5 + 121 + 81 + 3 * 121 * 81 + 2
    with pytest.raises(NotImplementedError):
        P.parse_pot_planar([ip])
    rng = numpy.random.default_rng(13)
    n = 200
    R, z = rng.uniform(0.3, 4.0, n), rng.uniform(-0.9, 0.9, n)
    R[:4] = [0.01, 6.0, 0.05, 5.0]  # clamped to the grid
    z[:4] = [0.0, 1.5, -1.2, 0.0]
    FR, Fz, Fp = oracle_forces(pt, pa, R, z)
    V = oracle_potential(pt, pa, R, z)
    assert numpy.all(Fp == 0.0)
    # the host class's scipy interpolation of the same coefficients
    assert numpy.allclose(ip.Rforce(R, z), FR, rtol=1e-12, atol=1e-14)
    assert numpy.allclose(ip.zforce(R, z), Fz, rtol=1e-12, atol=1e-14)
    assert numpy.allclose(ip(R, z), V, rtol=1e-12, atol=1e-14)
    # interpolation accuracy inside the grid
    inside = slice(4, None)
    sc = numpy.abs(mw_R := P.evaluateRforces(mw, R[inside], z[inside])).max()
    assert numpy.abs(FR[inside] - mw_R).max() < 1e-3 * sc
    assert numpy.abs(V[inside] - P.evaluatePotentials(mw, R[inside], z[inside])).max() < 1e-4


def test_bspline_against_scipy():
    """The oracle's interpolation on a random coefficient grid equals
    scipy's map_coordinates at the same fractional indices (a potential
    made of one coefficient grid, read through the R-force slot)."""
    from scipy import ndimage

    rng = numpy.random.default_rng(14)
    nR, nz = 9, 7
    Rg = numpy.linspace(0.5, 2.5, nR)
    zg = numpy.linspace(0.0, 1.2, nz)
    co = rng.normal(size=(3, nR, nz))
    pa = numpy.concatenate([[nR, nz, 0, 0, 0], Rg, zg, co.ravel(), [1.0, 0.0]])
    pt = numpy.array([13], dtype=numpy.int32)
    R, z = rng.uniform(0.5, 2.5, 300), rng.uniform(0.0, 1.2, 300)
    FR, Fz, _ = oracle_forces(pt, pa, R, z)
    V = oracle_potential(pt, pa, R, z)
    xn, yn = (R - 0.5) / (Rg[1] - Rg[0]), z / (zg[1] - zg[0])
    for k, got in ((0, V), (1, FR), (2, Fz)):
        want = ndimage.map_coordinates(co[k], [xn, yn], order=3, mode="mirror", prefilter=False)
        assert numpy.allclose(got, want, rtol=1e-12, atol=1e-12)
