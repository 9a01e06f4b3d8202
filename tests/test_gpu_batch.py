"""Batched persistent integration: every group of a mixed batch equals the
same group integrated alone through the C-ABI (to 1e-12) and the CPU oracle
(1e-10 symplectic, 1e-8 DOP853); bounded launches and graph replays give the
same results as one launch."""

import numpy
import pytest

from helpers import oracle_integrate, phase_diff

pytestmark = pytest.mark.gpu


def _mixed_groups(seed=0):
    from galpy_amd import workloads
    from galpy_amd.potential import HernquistPotential, LogarithmicHaloPotential

    rng = numpy.random.default_rng(seed)
    pots = [workloads.mw2014(), [HernquistPotential(a=1.7)], [LogarithmicHaloPotential(q=0.8)],
            workloads.loghalo_bar()]
    groups = []
    for i, n in enumerate([1, 5, 64, 65, 130, 3, 200, 17]):
        method = ["symplec4_hip", "leapfrog_hip", "dop853_hip"][i % 3]
        nout = int(rng.integers(2, 40))
        t = numpy.linspace(0.0, float(rng.uniform(1.0, 6.0)), nout)
        dt = None if i % 2 == 0 else (t[1] - t[0]) / 4.0
        groups.append(dict(pot=pots[i % 4], yo=workloads.ics(n, seed=100 + i), t=t, method=method, dt=dt))
    return groups


def _alone(g):
    from galpy_amd.integrate import integrateFullOrbit_hip

    res, err = integrateFullOrbit_hip(g["pot"], g["yo"].copy(), g["t"], g["method"], dt=g.get("dt"))
    return res


@pytest.mark.parametrize("opl", [1, 2])
def test_batch_matches_single_calls_and_oracle(opl):
    import torch

    from galpy_amd.batching import Batch
    from galpy_amd.integrate import METHOD_CODES
    from galpy_amd.potential import parse_pot

    groups = _mixed_groups()
    b = Batch(groups, opl=opl)
    b.run()
    torch.cuda.synchronize()
    for i, g in enumerate(groups):
        got = b.results[i].cpu().numpy()
        assert (b.errs[i].cpu().numpy() == 0).all()
        alone = _alone(g)
        tol = 1e-8 if g["method"] == "dop853_hip" else 1e-10
        assert phase_diff(got, alone) < 1e-12, i
        _, pt, pa, _ = parse_pot(g["pot"])
        ref, *_ = oracle_integrate(pt, pa, g["yo"], g["t"], METHOD_CODES[g["method"]], dt=g.get("dt"))
        assert phase_diff(got, ref) < tol, i


def test_bounded_launches_and_graph_replays_match_one_launch():
    import torch

    from galpy_amd.batching import Batch

    groups = _mixed_groups(seed=1)
    ref = Batch(groups)
    ref.run()
    torch.cuda.synchronize()
    want = [r.cpu().numpy() for r in ref.results]
    b = Batch(groups)
    n = b.run(budget=7)
    torch.cuda.synchronize()
    assert n > 1
    for i in range(len(groups)):
        assert numpy.array_equal(b.results[i].cpu().numpy(), want[i])
    # graph of 3 bounded launches, replayed until done; then reset eagerly and
    # replay the same graph again: every pass must equal the one-launch result
    g = Batch(groups)
    g.run(budget=5, graph_chunks=3)
    torch.cuda.synchronize()
    for i in range(len(groups)):
        assert numpy.array_equal(g.results[i].cpu().numpy(), want[i])
    for _ in range(2):
        for r in g.results:
            r.zero_()
        for y, grp in zip(g.yo, groups):
            y.copy_(torch.as_tensor(grp["yo"], device=y.device))
        g.reset()
        torch.cuda.synchronize()
        while g.remaining() > 0:
            g.graph.replay()
            torch.cuda.synchronize()
        for i in range(len(groups)):
            assert numpy.array_equal(g.results[i].cpu().numpy(), want[i])


def test_autotune_returns_a_candidate(tmp_path, monkeypatch):
    from galpy_amd.batching import autotune, make_groups

    monkeypatch.setenv("GALPY_AMD_CACHE", str(tmp_path))
    groups = make_groups(ngroups=16, seed=3, max_size=300)
    opl, timings = autotune(groups, sample_groups=8, budget=200)
    assert opl in (1, 2) and set(timings) == {1, 2}
    opl2, _ = autotune(groups, sample_groups=8, budget=200)  # cached
    assert opl2 == opl


def test_generic_composition_and_specialised_only_flag():
    """A composition without a specialised force (Hernquist + NFW) runs in
    the batch through the run-time dispatch force and matches the same group
    alone; forcing the specialised-only kernel instead finishes exactly that
    group with err = -1 and leaves the others unchanged."""
    import torch

    from galpy_amd import workloads
    from galpy_amd.batching import GH_BATCH_SPECIALISED, Batch
    from galpy_amd.potential import HernquistPotential, NFWPotential

    t = numpy.linspace(0.0, 3.0, 7)
    odd = dict(pot=[HernquistPotential(a=1.2), NFWPotential(a=4.0, amp=2.0)], yo=workloads.ics(70, seed=51), t=t,
               method="symplec4_hip", dt=0.01)
    mw = dict(pot=workloads.mw2014(), yo=workloads.ics(90, seed=52), t=t, method="leapfrog_hip", dt=0.01)
    b = Batch([odd, mw])
    assert not b.flags & GH_BATCH_SPECIALISED
    b.run()
    torch.cuda.synchronize()
    for i, g in enumerate((odd, mw)):
        assert (b.errs[i].cpu().numpy() == 0).all()
        assert phase_diff(b.results[i].cpu().numpy(), _alone(g)) < 1e-12, i
    b2 = Batch([odd, mw])
    b2.flags |= GH_BATCH_SPECIALISED
    b2.run()
    torch.cuda.synchronize()
    assert (b2.errs[0].cpu().numpy() == -1).all()
    assert (b2.errs[1].cpu().numpy() == 0).all()
    assert phase_diff(b2.results[1].cpu().numpy(), _alone(mw)) < 1e-12
