"""Pins the oracle's tangent map (the checker of the adjoint): Hessians
against mpmath, symplecticity of the state-transition matrix, column-wise
agreement with finite differences of the flow, and bit-identity of the base
orbit with a plain run (galpy tests/test_symplectic_dxdv.py:145-290)."""

import json
import os

import numpy
import pytest

from helpers import cyl_to_rect, ics, oracle_hessian, oracle_integrate, oracle_stm

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _mw():
    from galpy_amd.potential import MWPotential2014, parse_pot

    _, pt, pa, _ = parse_pot(MWPotential2014)
    return pt, pa


@pytest.mark.parametrize("name", ["psc", "mn", "nfw", "hernquist", "loghalo_q09", "powersph", "powersph_a2",
                                  "kepler", "jaffe", "isochrone", "plummer", "pseudoiso", "kuzmin", "burkert",
                                  "dehnencore", "dehnen", "dehnen_a05", "homogsphere", "flatpower",
                                  "flatpower_a0"])
def test_hessians_vs_mpmath(name):
    case = json.load(open(os.path.join(GOLDEN, "potentials.json")))[name]
    pts = case["points"]
    R = numpy.array([p["R"] for p in pts])
    z = numpy.array([p["z"] for p in pts])
    R2, z2, Rz = oracle_hessian(case["pot_type"], case["pot_args"], R, z)
    for got, key in ((R2, "R2"), (z2, "z2"), (Rz, "Rz")):
        want = numpy.array([p[key] for p in pts])
        scale = numpy.abs(want).max()
        assert numpy.all(numpy.abs(got - want) <= 1e-10 * numpy.maximum(numpy.abs(want), 1e-4 * scale)), (name, key)


@pytest.mark.parametrize("method", [3, 0])
def test_stm_is_symplectic_and_matches_fd(method):
    pt, pa = _mw()
    w0 = cyl_to_rect(ics(4, seed=31))
    t = numpy.arange(201) * 0.01
    M, wT = oracle_stm(pt, pa, w0, t, method, dt=0.01)
    J = numpy.block([[numpy.zeros((3, 3)), numpy.eye(3)], [-numpy.eye(3), numpy.zeros((3, 3))]])
    for m in M:
        assert abs(numpy.linalg.det(m) - 1.0) < 1e-10
        assert numpy.abs(m.T @ J @ m - J).max() < 1e-10
    # columns against central differences of the flow (through the plain
    # integrator; rows are cylindrical there, so compare the flow in rect form)
    from helpers import oracle_dxdv

    h = 1e-6
    for j in range(6):
        e = numpy.zeros(6)
        e[j] = h
        plus = oracle_dxdv(pt, pa, w0 + e, 0, t, method, dt=0.01)
        minus = oracle_dxdv(pt, pa, w0 - e, 0, t, method, dt=0.01)
        fd = (plus - minus) / (2 * h)
        assert numpy.abs(fd - M[:, :, j]).max() < 1e-6 * max(1.0, numpy.abs(M).max())


def test_base_orbit_bit_identical_to_plain_run():
    """tests/test_symplectic_dxdv.py:145-177: the base trajectory of a
    tangent-map run is the plain run's."""
    from helpers import oracle_dxdv

    pt, pa = _mw()
    yo = ics(3, seed=32)
    t = numpy.linspace(0.0, 3.0, 31)
    for method in (3, 0):
        res, *_ = oracle_integrate(pt, pa, yo, t, method)
        aug = numpy.zeros((3, 12))
        aug[:, :6] = cyl_to_rect(yo)
        aug[:, 6] = 1.0
        out = oracle_dxdv(pt, pa, aug, 1, t, method)
        # plain run's final row, converted back to rectangular
        ref = cyl_to_rect(res[:, -1])
        assert numpy.abs(out[:, :6] - ref).max() < 1e-13
