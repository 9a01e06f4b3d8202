"""Continuous-batching server: FIFO memory admission (CPU), and on the GPU
every request's result in a mixed, overlapping stream equals the same request
integrated alone through integrateFullOrbit_hip, within the parity bars
(1e-10 symplectic, 1e-8 DOP853 -- the mixed launch runs the generic batch
kernel, whose rounding can move a DOP853 step decision); memory-limited
admission still completes everything."""

import numpy
import pytest

from galpy_amd.serving import admission_order


def test_admission_is_fifo_and_respects_budget():
    assert admission_order([10, 20, 30], 0, 35) == 2
    assert admission_order([10, 20, 30], 20, 35) == 1
    assert admission_order([40, 1, 1], 0, 35) == 0  # no overtaking
    assert admission_order([], 0, 10) == 0


def _small_requests(seed=0, rate=400.0, nreq=14):
    from galpy_amd.serving import make_requests

    return make_requests(nreq, rate, seed=seed, max_orbits=300, max_len=8.0, max_out=30)


def _alone(r):
    from galpy_amd.integrate import integrateFullOrbit_hip

    res, err = integrateFullOrbit_hip(r.pot, r.yo.copy(), r.t, r.method, dt=r.dt)
    return res


@pytest.mark.gpu
def test_each_result_same_alone_and_in_the_mix():
    from helpers import phase_diff

    from galpy_amd.serving import Server

    reqs = _small_requests()
    srv = Server(budget=50)
    m = srv.run(reqs, realtime=True)
    assert m["completed"] == len(reqs) and m["rejected"] == 0
    assert m["iterations"] > 1
    for r in srv.finished:
        assert (r.err.cpu().numpy() == 0).all()
        tol = 1e-8 if r.method == "dop853_hip" else 1e-10
        assert phase_diff(r.result.cpu().numpy(), _alone(r)) < tol, r.id
    assert m["p99_latency_s"] >= m["p50_latency_s"] > 0


@pytest.mark.gpu
def test_memory_limited_admission_completes_all():
    from galpy_amd.serving import Server

    reqs = _small_requests(seed=3)
    biggest = max(r.nbytes(200) for r in reqs)
    srv = Server(budget=100, hbm_bytes=int(2.5 * biggest))
    m = srv.run(reqs, realtime=False)
    assert m["completed"] == len(reqs)
    assert srv.in_use == 0
