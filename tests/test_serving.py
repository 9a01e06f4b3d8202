"""Continuous-batching server: FIFO memory admission (CPU), and on the GPU

* every request's result in a mixed, overlapping stream equals the same
  request integrated alone through integrateFullOrbit_hip, within the parity
  bars (1e-10 symplectic, 1e-8 DOP853 -- the mixed launch runs the generic
  batch kernel, whose rounding can move a DOP853 step decision);
* every served request of a stream drawn by the serving workload's law
  (MWPotential2014 / Hernquist / LogarithmicHalo / LogarithmicHalo + rotating
  DehnenBar, symplec4 at dt = 0.01 or DOP853 50/50, log-uniform orbit
  counts, lengths and output counts; sizes scaled down so the CPU oracle
  finishes in seconds) matches the oracle: symplec4 to 1e-10, DOP853 to 1e-8
  on every orbit whose oracle twin compiled with fused multiply-adds stays
  within a tenth of that bar, and on the others (chaotic bar orbits) no more
  rows beyond the bar than the twin has;
* memory-limited admission completes everything, and the HBM the server
  counts bounds what it actually allocates, results included, while more
  result bytes than the budget pass through it."""

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


@pytest.mark.gpu
def test_served_results_match_the_oracle():
    from helpers import oracle_integrate

    from galpy_amd.integrate import METHOD_CODES
    from galpy_amd.potential import parse_pot
    from galpy_amd.serving import Server, make_requests

    reqs = make_requests(32, 2000.0, seed=11, max_orbits=2000, max_len=100.0, max_out=101)
    assert {r.method for r in reqs} == {"symplec4_hip", "dop853_hip"}
    srv = Server(budget=200)
    m = srv.run(reqs, realtime=True)
    assert m["completed"] == len(reqs) and m["iterations"] > 1
    kinds = set()
    for r in srv.finished:
        assert not r.result.is_cuda  # drained to the host
        assert (r.err.numpy() == 0).all()
        _, pt, pa, _ = parse_pot(r.pot)
        code = METHOD_CODES[r.method]
        kinds.add((tuple(int(c) for c in pt), code))
        ref, rerr, _, _ = oracle_integrate(pt, pa, r.yo, r.t, code, dt=r.dt)
        assert (rerr == 0).all()
        d = _row_diffs(r.result.numpy(), ref)
        bar = 1e-10 if code == 3 else 1e-8
        twin = _row_diffs(oracle_integrate(pt, pa, r.yo, r.t, code, dt=r.dt, fma=True)[0], ref)
        regular = twin.max(axis=1) <= bar / 10
        assert d[regular].max(initial=0.0) < bar, (r.id, r.method, d[regular].max())
        assert (d > bar).sum() <= 1.25 * (twin > bar).sum() + 10, (r.id, r.method)
    assert (tuple([0, 1]), 6) in kinds or (tuple([0, 1]), 3) in kinds  # the rotating bar was served


def _row_diffs(a, b):
    d = numpy.abs(a - b) / numpy.maximum(numpy.abs(b), 1.0)
    d[..., -1] = numpy.abs((a[..., -1] - b[..., -1] + numpy.pi) % (2 * numpy.pi) - numpy.pi)
    return d.max(axis=-1)


@pytest.mark.gpu
def test_counted_hbm_bounds_allocations_while_results_drain():
    import torch

    from galpy_amd.serving import Server, make_requests

    reqs = make_requests(40, 1e6, seed=5, max_orbits=3000, max_len=20.0, max_out=301)
    total = sum(r.nbytes(200) for r in reqs)
    biggest = max(r.nbytes(200) for r in reqs)
    cap = int(1.5 * biggest)
    assert total > 3 * cap  # more result bytes pass through than the budget holds
    srv = Server(budget=400, hbm_bytes=cap)
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    m = srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    assert m["completed"] == len(reqs)
    assert srv.in_use == 0 and srv.peak_in_use <= cap
    # launch descriptors (tiles, groups, flags) come on top of the counted buffers
    assert torch.cuda.max_memory_allocated() - base <= cap + (4 << 20)
    assert torch.cuda.memory_allocated() - base <= (1 << 20)
    assert all(not r.result.is_cuda for r in srv.finished)
