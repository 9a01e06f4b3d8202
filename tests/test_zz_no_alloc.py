"""integrateFullOrbit_hip_noalloc (the allocation-free entry point) against
integrateFullOrbit_hip through the C-ABI: two calls in a row, the second on
fewer orbits, must give bit-identical rows, error codes and step counts
(the two share their kernels; only where the launch temporaries live
differs).  DOP853 runs with more orbits than resident lanes, so its pilot
order and radix sort use the scratch array too.  A call whose temporaries
exceed the scratch array fails with a message and launches nothing."""

import numpy
import pytest

from helpers import ics

pytestmark = pytest.mark.gpu

DT_SENTINEL = -9999.99
LN1EM12 = -12.0 * numpy.log(10.0)


def _call(symbol, table, yo, t, code, dt=None):
    import torch

    from galpy_amd import _lib

    lib = _lib.load()
    n, nt = yo.shape[0], t.shape[-1]
    dyo = yo.clone()
    res = torch.full((n, nt, 6), -1.0, dtype=torch.float64, device="cuda")
    err = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    stats = torch.full((n, 3), -1.0, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    rc = getattr(lib, symbol)(n, dyo.data_ptr(), nt, t.data_ptr(), 0, table.ptr,
                              DT_SENTINEL if dt is None else dt, LN1EM12, LN1EM12, res.data_ptr(),
                              err.data_ptr(), code, stats.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    return rc, dyo, res, err, stats


@pytest.mark.parametrize("code,sizes,tmax,dt", [(6, (200_000, 3_000), 5.0, None),
                                                (3, (70_000, 1_000), 5.0, 0.01),
                                                (0, (5_000, 640), 5.0, None)])
def test_noalloc_matches_allocating_entry_point(code, sizes, tmax, dt):
    import torch

    from galpy_amd import _lib
    from galpy_amd.integrate import PotentialTable
    from galpy_amd.potential import MWPotential2014

    table = PotentialTable(MWPotential2014)
    t = torch.linspace(0.0, tmax, 11, dtype=torch.float64, device="cuda")
    for n in sizes:
        print(f"noalloc odeint {code}: {n} orbits", flush=True)
        yo = torch.tensor(ics(n, seed=n), device="cuda")
        rc0, y0, r0, e0, s0 = _call("integrateFullOrbit_hip", table, yo, t, code, dt)
        rc1, y1, r1, e1, s1 = _call("integrateFullOrbit_hip_noalloc", table, yo, t, code, dt)
        assert rc0 == 0 and rc1 == 0, _lib.last_error()
        assert int((e0 != 0).sum()) == 0
        for a, b in ((y0, y1), (r0, r1), (e0, e1), (s0, s1)):
            assert torch.equal(a, b)


def test_noalloc_rejects_temporaries_beyond_the_scratch_array():
    import torch

    from galpy_amd import _lib
    from galpy_amd.integrate import PotentialTable
    from galpy_amd.potential import MWPotential2014

    table = PotentialTable(MWPotential2014)
    n = 4_000_000  # 64 MB of pilot order on top of the overflow area
    t = torch.tensor([0.0, 0.5], dtype=torch.float64, device="cuda")
    yo = torch.tensor(ics(n, seed=1), device="cuda")
    rc, _, res, err, _ = _call("integrateFullOrbit_hip_noalloc", table, yo, t, 6)
    assert rc != 0
    assert "exceed the workspace" in _lib.last_error()
    assert int((err != 7).sum()) == 0 and bool((res == -1.0).all())  # nothing launched
    # the array is still usable afterwards
    small = yo[:1000].contiguous()
    rc0, _, r0, _, _ = _call("integrateFullOrbit_hip", table, small, t, 6)
    rc1, _, r1, _, _ = _call("integrateFullOrbit_hip_noalloc", table, small, t, 6)
    assert rc0 == 0 and rc1 == 0 and torch.equal(r0, r1)
