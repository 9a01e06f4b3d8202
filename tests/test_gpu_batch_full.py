"""The full 4,096-group batching workload (sizes log-uniform in [1, 10,000],
MWPotential2014 / Hernquist / LogarithmicHalo, lengths U(10, 1000) with
2..1001 outputs, symplec4 at dt = 0.01) in one persistent batch: every orbit
err = 0; every orbit takes exactly (long)(step/dt) steps per output interval
(galpy's truncation, integrateFullOrbit.c via bovy_symplecticode.c:288-298);
|dE/E| below 1e-6 for 99.9% of the orbits, and the orbits that drift most
follow the oracle's symplec4 (their drift is the integrator's);
32 sampled groups equal the oracle's integration to 1e-10."""

import numpy
import pytest

from helpers import assert_parity, oracle_integrate, phase_diff, ulp_sensitivity

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_full_size_batch():
    import torch

    from galpy_amd.batching import Batch, PotentialCache, make_groups
    from galpy_amd.integrate import orbit_summary_device
    from galpy_amd.potential import parse_pot

    groups = make_groups(ngroups=4096, seed=0)
    cache = PotentialCache()
    b = Batch(groups, opl=1, potentials=cache)
    print(f"{len(groups)} groups, {sum(g['yo'].shape[0] for g in groups)} orbits", flush=True)
    b.run()
    torch.cuda.synchronize()
    drifts, where = [], []
    for i, g in enumerate(groups):
        assert int((b.errs[i] != 0).sum().item()) == 0, i
        t = g["t"]
        k = int((t[1] - t[0]) / g["dt"])
        steps = b.stats[i][:, 0]
        assert bool((steps == float(k * (len(t) - 1))).all().item()), i
        tab, _, _ = cache.get(g["pot"])
        E = orbit_summary_device(b.results[i], t=b.t[i], table=tab, energies=True)["E"]
        d = ((E - E[:, :1]).abs() / E[:, :1].abs()).max(dim=1).values.cpu().numpy()
        drifts.append(d)
        where.extend((i, j) for j in range(d.size))
    drifts = numpy.concatenate(drifts)
    q = numpy.quantile(drifts, [0.5, 0.99, 0.999, 1.0])
    print("|dE/E| quantiles (50%, 99%, 99.9%, max): " + ", ".join(f"{v:.2e}" for v in q), flush=True)
    assert q[2] < 1e-6
    # the largest drifts are the integrator's own: those orbits follow the
    # oracle's symplec4 to 1e-10 (or 1e3 x the oracle's own one-ulp
    # sensitivity of the orbit), drift included
    for w in numpy.argsort(drifts)[-6:]:
        i, j = where[w]
        g = groups[i]
        _, pt, pa, _ = parse_pot(g["pot"])
        ref, sens = ulp_sensitivity(lambda y: oracle_integrate(pt, pa, y, g["t"], 3, dt=g["dt"])[0],
                                    g["yo"][j:j + 1])
        assert_parity(b.results[i][j:j + 1].cpu().numpy(), ref, 1e-10, sens, factor=1e3,
                      label=f"group {i} orbit {j} drift {drifts[w]:.1e}")
    # the oracle on 32 groups of at most 200 orbits and 300 outputs
    small = [i for i, g in enumerate(groups) if g["yo"].shape[0] <= 200 and len(g["t"]) <= 300]
    rng = numpy.random.default_rng(7)
    for i in rng.choice(small, size=min(32, len(small)), replace=False):
        g = groups[i]
        _, pt, pa, _ = parse_pot(g["pot"])
        ref, *_ = oracle_integrate(pt, pa, g["yo"], g["t"], 3, dt=g["dt"])
        assert phase_diff(b.results[i].cpu().numpy(), ref) < 1e-10, i
