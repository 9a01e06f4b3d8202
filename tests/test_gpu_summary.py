"""On-device post-processing (gh_orbit_summary_hip): galpy's trajectory
summaries rperi, rap, zmax and e (Orbits.py:3642-4010, analytic=False) and E
and the Jacobi integral at every output (Orbits.py:2902-3115, :3298-3360),
computed from the rows where they are stored.

Bars: rperi, rap, zmax and e are exactly numpy's of the same rows (the
kernel forms r = sqrt(R^2 + z^2) without FMA contraction; the square root
and division are correctly rounded); E and E - omega Lz within 1e-13 of the
oracle's potential plus the kinetic term; galpy's own golden numbers for
MWPotential2014 (tests/test_galpypaper.py:293-360) through the device path."""

import numpy
import pytest

from helpers import ics, oracle_potential, planar_ics

pytestmark = pytest.mark.gpu


def _numpy_summary(rows):
    R = rows[..., 0]
    z = rows[..., 3] if rows.shape[-1] == 6 else 0.0 * R
    r = numpy.sqrt(R**2 + z**2)
    rp, ra = r.min(axis=1), r.max(axis=1)
    return dict(rperi=rp, rap=ra, e=(ra - rp) / (ra + rp), zmax=numpy.abs(z).max(axis=1))


def _integrate(pot, yo, t, method="dop853_hip", dt=None):
    import torch

    from galpy_amd.integrate import PotentialTable, integrate_device

    planar = yo.shape[1] == 4
    tab = PotentialTable(pot, planar=planar)
    tt = torch.tensor(t, device="cuda")
    res, err, _ = integrate_device(tab, torch.tensor(yo, device="cuda"), tt, method, dt=dt)
    torch.cuda.synchronize()
    assert int((err != 0).sum()) == 0
    return tab, tt, res


@pytest.mark.parametrize("planar", [False, True])
def test_summaries_equal_numpy_of_the_same_rows(planar):
    from galpy_amd.integrate import orbit_summary_device
    from galpy_amd.potential import MWPotential2014

    yo = planar_ics(1000, seed=1) if planar else ics(1000, seed=1)
    for nt in (1, 7, 64, 65, 301):  # fewer rows than lanes, exact multiples, ragged
        tt = numpy.linspace(0.0, 5.0, max(nt, 2))
        _, _, res = _integrate(MWPotential2014, yo, tt)
        rows = res[:, :nt].contiguous()
        out = orbit_summary_device(rows)
        want = _numpy_summary(rows.cpu().numpy())
        for k in ("rperi", "rap", "e") + (() if planar else ("zmax",)):
            assert numpy.array_equal(out[k].cpu().numpy(), want[k]), (k, nt)
        assert planar == ("zmax" not in out)


def test_energies_match_oracle_with_wrappers():
    import torch

    from galpy_amd import potential as P
    from galpy_amd.integrate import orbit_summary_device
    from galpy_amd.potential import parse_pot

    omega = 0.6
    pot = [*P.MWPotential2014_factory(),
           P.SolidBodyRotationWrapperPotential(pot=P.LogarithmicHaloPotential(amp=0.05, q=0.9, b=0.8),
                                               omega=omega)]
    yo = ics(256, seed=2)
    t = numpy.linspace(0.0, 20.0, 201)
    tab, tt, res = _integrate(pot, yo, t)
    out = orbit_summary_device(res, tt, tab, omega=omega, energies=True, jacobi=True)
    rows = res.cpu().numpy()
    R, vR, vT, z, vz, phi = numpy.moveaxis(rows, -1, 0)
    _, pt, pa, _ = parse_pot(pot)
    tg = numpy.broadcast_to(t, R.shape)
    Phi = oracle_potential(pt, pa, R.ravel(), z.ravel(), phi.ravel(), tg.ravel()).reshape(R.shape)
    E = Phi + vR**2 / 2 + vT**2 / 2 + vz**2 / 2
    got = out["E"].cpu().numpy()
    assert numpy.abs(got - E).max() <= 1e-13 * numpy.abs(E).max()
    EJ = out["Jacobi"].cpu().numpy()
    assert numpy.abs(EJ - (E - omega * R * vT)).max() <= 1e-13 * numpy.abs(E).max()
    # the rotating frame's integral is conserved along the device rows
    assert numpy.abs(EJ - EJ[:, :1]).max() < 1e-9 * numpy.abs(EJ).max()
    # individual time grids
    tind = torch.tensor(numpy.tile(t, (256, 1)), device="cuda")
    out2 = orbit_summary_device(res, tind, tab, omega=omega, energies=True)
    assert torch.equal(out2["E"], out["E"])


def test_galpy_golden_numbers_on_device():
    """tests/test_galpypaper.py:293-360: rperi, rap, zmax, e and E of the
    MWPotential2014 orbit (0.8, 0.3, 0.75, 0, 0.2, 0) under symplec4."""
    from galpy_amd.integrate import orbit_summary_device
    from galpy_amd.potential import MWPotential2014

    t = numpy.linspace(0.0, 10.0, 1001)
    tab, tt, res = _integrate(MWPotential2014, numpy.array([[0.8, 0.3, 0.75, 0.0, 0.2, 0.0]]), t,
                              method="symplec4_hip")
    out = {k: v.cpu().numpy() for k, v in orbit_summary_device(res, tt, tab, energies=True).items()}
    assert abs(out["rperi"][0] - 0.44231993168097) < 1e-5
    assert abs(out["rap"][0] - 0.87769030382105) < 1e-5
    assert abs(out["zmax"][0] - 0.077452357289016) < 1e-5
    assert abs(out["e"][0] - 0.32982348199330563) < 1e-5
    assert abs(out["E"][0, 0] + 1.2547650648697966) < 1e-5


def test_summary_argument_errors():
    import torch

    from galpy_amd import _lib
    from galpy_amd.integrate import orbit_summary_device

    rows = torch.zeros((4, 3, 5), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        orbit_summary_device(rows)
    rows = torch.ones((4, 3, 6), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        orbit_summary_device(rows, energies=True)
    lib = _lib.load()
    rc = lib.gh_orbit_summary_hip(4, 3, 6, rows.data_ptr(), None, 0, None, 0.0, None, None, None, None,
                                  rows.data_ptr(), None, None)
    assert rc != 0 and "potential" in _lib.last_error()
    assert lib.gh_orbit_summary_hip(4, 0, 6, rows.data_ptr(), None, 0, None, 0.0, None, None, None, None,
                                    None, None, None) != 0
    assert lib.gh_orbit_summary_hip(0, 3, 6, None, None, 0, None, 0.0, None, None, None, None, None, None,
                                    None) == 0
