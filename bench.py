#!/usr/bin/env python
"""Benchmark: FP64 orbit-steps/s of the HIP DOP853 integrator on one MI355X.

Workload (default): 1,048,576 orbits in MWPotential2014, dop853 at galpy's
default tolerances (rtol=atol=1e-12), t=linspace(0,500,1001), seeded initial
conditions; all 50.4 GB of trajectories stay resident in HBM.  One "step" is
one integration of the whole batch.  value = accepted orbit-steps per second
over the K timed steps.  |dE/E| is measured on device afterwards.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

NAME: mw2014_dop853_1m (default), loghalo_bar_dop853_256k, mw2014_symplec4_1m,
mw2014_symplec4_1k, mw2014_planar_dop853_1m, mw2014_planar_symplec4_1m (planar
orbits, integratePlanarOrbit), mw2014_dxdv_symplec4_65k, mw2014_dxdv_dop853_65k
(integrate_dxdv: orbit + one deviation), batching, serving, adjoint, copy
(counter calibration).

Prints one JSON line (see the keys below).  The CPU baseline leg times the
CPU oracle (a C++/OpenMP restatement of galpy's C integrators) on a bounded
sample of the same workload on this host's cores; the same leg counts the
algorithm's FP64 operations per orbit-step, which prices the roofline.
"""

import argparse
import json
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the integration kernel, measured on MI355X with two
# rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) of
#   python bench.py --steps 1 --warmup 0 --no-cpu-baseline
# each scaled by what the counter reads for a plain 2 GiB device copy
# (python bench.py --workload copy): FETCH_SIZE reads 1 GiB (x2), WRITE_SIZE
# 2 GiB (x1).  Valid for the workload at its default size and this build.
MEASURED_TRAFFIC = {
    # 304,802.6 KiB fetched x 2 + 77,530,416.5 KiB written
    "mw2014_dop853_1m": (304802.625 * 2.0 + 77530416.5) * 1024.0,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mw2014_dop853_1m")
    ap.add_argument("--norbits", type=int, default=None, help="override the orbit count")
    ap.add_argument("--cpu-sample", type=int, default=1024, help="orbits in the CPU-baseline sample")
    ap.add_argument("--count-sample", type=int, default=64, help="orbits in the op-count sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ngroups", type=int, default=4096, help="groups of the batching workload")
    ap.add_argument("--nreq", type=int, default=300, help="requests per offered rate (serving)")
    ap.add_argument("--rates", default="50,5000", help="offered request rates, req/s (serving)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (latency) leg")
    ap.add_argument("--cpu-reps", type=int, default=5, help="CPU-baseline runs (median reported)")
    ap.add_argument("--budget", type=int, default=4000, help="steps per orbit per serving iteration")
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_leg(w, sample, count_sample, reps=5):
    """CPU oracle on a bounded sample of the same workload (OpenMP, all host
    cores; the median of `reps` timed runs), plus the op count per
    orbit-step of the same algorithm."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    # the oracle is test infrastructure
    from helpers import oracle_count_ops, oracle_count_ops_dxdv, oracle_integrate, oracle_integrate_dxdv

    from galpy_amd.integrate import METHOD_CODES
    from galpy_amd.potential import parse_pot, parse_pot_planar

    planar = w["planar"]
    _, pt, pa, _ = (parse_pot_planar if planar else parse_pot)(w["pot"])
    code = METHOD_CODES[w["method"]]
    nthreads = len(os.sched_getaffinity(0))
    yo = w["yo"][:sample]
    times = []
    if w["dxdv"]:
        for _ in range(reps):
            nsteps = numpy.zeros(len(yo), dtype=numpy.int64)
            t0 = time.perf_counter()
            oracle_integrate_dxdv(pt, pa, yo, w["t"], code, dim=3, dt=w.get("dt"), nthreads=nthreads,
                                  steps=nsteps)
            times.append(time.perf_counter() - t0)
        el = float(numpy.median(times))
        steps = float(nsteps.sum())
        ops = oracle_count_ops_dxdv(pt, pa, w["yo"][:count_sample], w["t"], code, dim=3, dt=w.get("dt"))
        ops["rhs"] = None
    else:
        for _ in range(reps):
            t0 = time.perf_counter()
            _, err, stats, _ = oracle_integrate(pt, pa, yo, w["t"], code, dt=w.get("dt"),
                                                nthreads=nthreads, planar=planar)
            times.append(time.perf_counter() - t0)
        el = float(numpy.median(times))
        steps = stats[:, 0].sum()
        ops = oracle_count_ops(pt, pa, w["yo"][:count_sample], w["t"], code, dt=w.get("dt"),
                               planar=planar)
    flops = ops["add"] + ops["mul"] + ops["div"] + ops["sqrt"] + ops["special"]
    return {
        "value": steps / el,
        "unit": "orbit-steps/s",
        "cores": nthreads,
        "kind": "port",
        "sample": (f"first {sample} of the {w['n']} orbits, full t grid ({int(steps)} orbit-steps); "
                   f"median of {reps} runs: {el:.2f} s (runs {', '.join(f'{x:.2f}' for x in times)})"),
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
    }, {
        "flop_per_orbit_step": flops / ops["steps"],
        "special_per_orbit_step": ops["special"] / ops["steps"],
        "rhs_per_orbit_step": None if ops["rhs"] is None else ops["rhs"] / ops["steps"],
        "count_sample": count_sample,
    }


def bench_batching(args):
    """4,096 groups (log-uniform sizes in [1, 1e4], MWPotential2014 /
    Hernquist / LogarithmicHalo, lengths U(10,1000), 2-1001 outputs,
    symplec4 at dt=0.01) in one persistent launch, against the same groups
    launched one integrateFullOrbit_hip call per group on a sample."""
    import torch

    from galpy_amd.batching import Batch, autotune, make_groups
    from galpy_amd.integrate import integrate_device

    groups = make_groups(ngroups=args.ngroups, seed=0)
    opl, timings = autotune(groups)
    # plain path on a sample of groups: one launch per group, in sequence
    rng = numpy.random.default_rng(1)
    sample = [groups[i] for i in sorted(rng.choice(len(groups), size=min(128, len(groups)), replace=False))]
    b_s = Batch(sample, opl=opl)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b_s.run()
    torch.cuda.synchronize()
    t_batch_sample = time.perf_counter() - t0
    steps_sample = b_s.orbit_steps()
    del b_s
    stats_plain = 0.0
    t0 = time.perf_counter()
    for g in sample:
        from galpy_amd.integrate import PotentialTable

        tab = PotentialTable(g["pot"])
        yo = torch.tensor(g["yo"], device="cuda")
        tt = torch.tensor(g["t"], device="cuda")
        st = torch.zeros((len(g["yo"]), 3), dtype=torch.float64, device="cuda")
        integrate_device(tab, yo, tt, g["method"], dt=g["dt"], stats=st)
        torch.cuda.synchronize()
        stats_plain += st[:, 0].sum().item()
    t_plain = time.perf_counter() - t0
    # full batch
    b = Batch(groups, opl=opl)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    steps = b.orbit_steps()
    nerr = sum(int((e != 0).sum().item()) for e in b.errs)
    out = {
        "metric": "FP64 orbit-steps/s on 1 MI355X, batched groups",
        "value": steps / el,
        "unit": "orbit-steps/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": el * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batching_{len(groups)}_groups", "orbits": b.norbits,
                   "tiles": b.ntiles, "distinct_potential_tables": len(b.pots), "method": "symplec4_hip",
                   "dt": 0.01, "opl": opl, "autotune_orbit_steps_per_s": timings},
        "orbit_steps_per_pass": steps,
        "orbits_with_error": nerr,
        "plain_one_launch_per_group": {
            "sample_groups": len(sample),
            "orbit_steps_per_s": stats_plain / t_plain,
            "batched_same_sample_orbit_steps_per_s": steps_sample / t_batch_sample,
            "batched_full_over_plain": (steps / el) / (stats_plain / t_plain),
        },
    }
    print(json.dumps(out), flush=True)


def serving_leg(nreq, rates, budget):
    """Poisson request streams drawn by the serving law (N log-uniform in
    [1, 1e5], potentials from MWPotential2014 / Hernquist / LogarithmicHalo /
    LogarithmicHalo + rotating DehnenBar, lengths U(1, 1000) with 2..501
    outputs, symplec4 at dt = 0.01 or DOP853 50/50) through the
    continuous-batching server, one stream per offered rate; results drain
    to host memory, latency = result on the host - arrival."""
    import torch

    from galpy_amd.serving import Server, make_requests

    # warm-up: code load and first-launch set-up outside the measured streams
    Server(budget=budget).run(make_requests(8, 1000.0, seed=99, max_orbits=64, max_len=2.0, max_out=5),
                              realtime=False)
    sweep = []
    for k, rate in enumerate(rates):
        reqs = make_requests(nreq, rate, seed=k)
        srv = Server(budget=budget)
        m = srv.run(reqs, realtime=True)
        m["rate"] = rate
        sweep.append(m)
        del srv, reqs
        torch.cuda.empty_cache()
    return sweep


def bench_serving(args):
    """Poisson request streams at several offered rates through the
    continuous-batching server; reports p50/p99 latency and orbit-steps/s."""
    rates = [float(r) for r in args.rates.split(",")]
    sweep = serving_leg(args.nreq, rates, args.budget)
    best = max(sweep, key=lambda m: m["orbit_steps_per_s"])
    out = {
        "metric": "p50/p99 request latency and FP64 orbit-steps/s, continuous batching",
        "value": best["orbit_steps_per_s"],
        "unit": "orbit-steps/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": best["makespan_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "serving_poisson", "requests_per_rate": args.nreq, "rates": rates,
                   "orbits_per_request": "log-uniform [1, 1e5]", "methods": "symplec4 (dt=0.01) / dop853 50/50",
                   "budget_steps_per_iteration": args.budget},
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """Differentiation workload: 65,536 MWPotential2014 orbits, symplec4 and
    leapfrog, 2,000 steps of dt = 0.01; loss = sum of squared distances of the
    final rectangular states to targets drawn as perturbed forward solutions;
    gradients w.r.t. every initial condition and NFW amp/a, MiyamotoNagai a/b
    through the reverse-time adjoint (orbit_flow's backward).  value = orbit-
    steps of the differentiated problem per second (one gradient = forward
    flow + reverse pass over n x 2,000 orbit-steps), symplec4; the line also
    carries each method's central finite-difference check (h = 1e-6) of the
    four parameter gradients."""
    import torch

    from galpy_amd import workloads
    from galpy_amd.adjoint import ParamSpec, _RawTable, cyl_to_rect_torch, flow_raw, orbit_flow
    from galpy_amd.potential import MWPotential2014

    n, nsteps, dt, h = args.norbits or 65536, 2000, 0.01, 1e-6
    w0 = cyl_to_rect_torch(torch.tensor(workloads.ics(n, seed=0), device="cuda"))
    spec = ParamSpec(MWPotential2014, [(2, 0), (2, 1), (1, 1), (1, 2)])
    theta0 = spec.theta0()
    tab = _RawTable(spec.pot_type, spec.pot_args)
    pert = torch.tensor(numpy.random.default_rng(7).normal(0.0, 0.01, (n, 6)), device="cuda")
    per = {}
    for method in ("symplec4_hip", "leapfrog_hip"):
        target = flow_raw(tab, w0 + pert, dt, nsteps, method)

        def gradient():
            theta = torch.tensor(theta0, device="cuda", requires_grad=True)
            w = w0.clone().requires_grad_(True)
            loss = ((orbit_flow(w, theta, spec, dt, nsteps, method) - target) ** 2).sum()
            loss.backward()
            return loss, theta.grad, w.grad

        for _ in range(args.warmup):
            gradient()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            loss, gth, gw = gradient()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / args.steps

        def L(th):
            return ((flow_raw(_RawTable(*spec.packed(th)), w0, dt, nsteps, method) - target) ** 2).sum().item()

        rel = []
        for k in range(len(theta0)):
            e = numpy.zeros_like(theta0)
            e[k] = h
            fd = (L(theta0 + e) - L(theta0 - e)) / (2 * h)
            rel.append(abs(gth[k].item() - fd) / abs(fd))
        per[method] = {"orbit_steps_per_s": n * nsteps / el, "ms_per_gradient": 1e3 * el, "loss": loss.item(),
                       "fd_max_rel_err_params": max(rel), "grad_params": gth.tolist()}
    v = per["symplec4_hip"]
    print(json.dumps({"metric": "FP64 orbit-steps/s of loss + gradient (forward flow + reverse-time adjoint)",
                      "value": v["orbit_steps_per_s"], "unit": "orbit-steps/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": v["ms_per_gradient"], "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "adjoint_mw2014_65536x2000", "orbits": n, "steps_per_orbit": nsteps,
                                 "dt": dt, "params": "NFW amp, a; MiyamotoNagai a, b", "fd_h": h},
                      "methods": per}), flush=True)


def bench_summary(args):
    """On-device post-processing: rperi, rap, zmax and e (galpy's
    Orbit.rperi/rap/zmax/e, analytic=False) over the trajectories of the
    headline integration (MWPotential2014, DOP853, 1,048,576 orbits x 1001
    outputs, 50.4 GB resident in HBM; integrated once, untimed).  A step is
    one gh_orbit_summary_hip pass over all rows; the kernel reads every row
    once, so it is priced against HBM bandwidth."""
    import torch

    from galpy_amd import _lib, workloads
    from galpy_amd.integrate import PotentialTable, integrate_device, orbit_summary_device

    _lib.load()
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    w = workloads.make("mw2014_dop853_1m", n=args.norbits)
    n, nt = w["n"], len(w["t"])
    table = PotentialTable(w["pot"])
    tt = torch.tensor(w["t"], dtype=torch.float64, device="cuda")
    rows, err, _ = integrate_device(table, torch.tensor(w["yo"], device="cuda"), tt, "dop853_hip")
    torch.cuda.synchronize()
    assert int((err != 0).sum()) == 0
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        orbit_summary_device(rows, stream=stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        out = orbit_summary_device(rows, stream=stream)
        b.record(stream)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kms = [a.elapsed_time(b) for a, b in ev]
    read = n * nt * 6 * 8
    written = 4 * n * 8
    achieved = (read + written) / (numpy.mean(kms) / 1e3) / 1e9
    # the reference computes these with numpy on the host (Orbits.py:3642-4010):
    # the same expressions on a bounded sample of the rows, one core
    sample = min(n, 16384)
    host = rows[:sample].cpu().numpy()
    c0 = time.perf_counter()
    R, z = host[..., 0], host[..., 3]
    r = numpy.sqrt(R**2 + z**2)
    rp, ra = numpy.amin(r, axis=-1), numpy.amax(r, axis=-1)
    _ = ((ra - rp) / (ra + rp), numpy.amax(numpy.fabs(z), axis=-1))
    cel = time.perf_counter() - c0
    same = numpy.array_equal(out["rperi"][:sample].cpu().numpy(), rp) and numpy.array_equal(
        out["rap"][:sample].cpu().numpy(), ra)
    print(json.dumps({
        "metric": "trajectory rows summarised/s (rperi, rap, zmax, e) on 1 MI355X",
        "value": n * nt * args.steps / el, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "summary_mw2014_dop853_1m", "orbits": n, "outputs": nt,
                   "row_bytes": 48, "source": "MWPotential2014 DOP853 t=linspace(0,500,1001)"},
        "kernel_ms": kms, "matches_numpy_on_sample": bool(same),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": None, "kernel": "gh_summary_kernel<6>",
                     "algorithmic_bytes": read + written,
                     "note": "algorithmic bytes = every row read once (48 B) + 4 doubles written per orbit"},
        "cpu_baseline": {"value": sample * nt / cel, "unit": "rows/s", "cores": 1, "kind": "reference",
                         "sample": f"galpy's numpy expressions on the first {sample} orbits' rows ({cel:.2f} s)"},
    }), flush=True)


def bench_copy(args):
    """A plain device-to-device copy of known size: the calibration run for
    the FETCH_SIZE / WRITE_SIZE hardware counters (bytes read and written by
    one copy launch are known exactly)."""
    import torch

    nbytes = 1 << 31
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"metric": "device copy bandwidth (counter calibration)", "value": 2 * nbytes * args.steps / el / 1e9,
                      "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak",
                      "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "copy", "bytes_read_per_launch": nbytes,
                                 "bytes_written_per_launch": nbytes}}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload.startswith("batching"):
        bench_batching(args)
        return
    if args.workload.startswith("serving"):
        bench_serving(args)
        return
    if args.workload == "copy":
        bench_copy(args)
        return
    if args.workload.startswith("summary"):
        bench_summary(args)
        return
    if args.workload.startswith("adjoint"):
        bench_adjoint(args)
        return

    import torch

    from galpy_amd import _lib, workloads
    from galpy_amd.integrate import (
        PotentialTable,
        energy_drift_device,
        integrate_device,
        integrate_dxdv_device,
    )

    _lib.load()  # fail loudly when the HIP library is missing
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    dev = torch.device("cuda:0")
    w = workloads.make(args.workload, n=args.norbits)
    n, t_np = w["n"], w["t"]
    nt = len(t_np)
    ncol = 12 if w["dxdv"] else 4 if w["planar"] else 6  # doubles per row
    table = PotentialTable(w["pot"], planar=w["planar"])
    yo0 = torch.tensor(w["yo"], dtype=torch.float64, device=dev)
    yo = torch.empty_like(yo0)
    tt = torch.tensor(t_np, dtype=torch.float64, device=dev)
    result = torch.empty((n, nt, ncol), dtype=torch.float64, device=dev)
    err = torch.zeros(n, dtype=torch.int32, device=dev)
    stats = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def one_step(ev=None):
        yo.copy_(yo0)
        if ev is not None:
            ev[0].record(stream)
        if w["dxdv"]:
            integrate_dxdv_device(table, yo, tt, w["method"], dt=w.get("dt"), result=result, err=err,
                                  stats=stats, stream=stream)
        else:
            integrate_device(table, yo, tt, w["method"], dt=w.get("dt"), result=result, err=err,
                             stats=stats, stream=stream)
        if ev is not None:
            ev[1].record(stream)

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        one_step(events[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in events]

    st = stats.cpu().numpy()
    steps_per_pass = float(st[:, 0].sum())
    rejected = float(st[:, 1].sum())
    nerr = int((err != 0).sum().item())
    if w["dxdv"]:  # the orbit's energy along the rectangular rows
        q = result[..., :6]
        Rr = torch.sqrt(q[..., 0] ** 2 + q[..., 1] ** 2)
        ph = torch.atan2(q[..., 1], q[..., 0])
        cp, sp = torch.cos(ph), torch.sin(ph)
        cyl = torch.stack([Rr, q[..., 3] * cp + q[..., 4] * sp, q[..., 4] * cp - q[..., 3] * sp,
                           q[..., 2], q[..., 5], ph], dim=-1).contiguous()
        e0, demax = energy_drift_device(table, cyl, tt, omega=w["omega"])
        del q, Rr, ph, cp, sp, cyl
    else:
        e0, demax = energy_drift_device(table, result, tt, omega=w["omega"])
    dem = demax.cpu().numpy()
    torch.cuda.synchronize()

    value = steps_per_pass * args.steps / elapsed
    frac_ok = float((dem <= 1e-8).mean())
    out = {
        "metric": "FP64 orbit-steps/s on 1 MI355X at |dE/E|<=1e-8; p50/p99 request latency",
        "value": value,
        "unit": "orbit-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": w["name"],
            "orbits": n,
            "orbit_dim": 2 if w["planar"] else 3,
            "deviations_per_orbit": 1 if w["dxdv"] else 0,
            "method": w["method"],
            "potential": w["potential"],
            "t": [float(t_np[0]), float(t_np[-1]), nt],
            "rtol": 1e-12,
            "atol": 1e-12,
            "dt": w.get("dt"),
            "output_bytes": n * nt * 8 * ncol,
        },
        "orbit_steps_per_pass": steps_per_pass,
        "frac_orbits_dE_le_1e-8": frac_ok,
        "rejected_steps_per_pass": rejected,
        "orbits_with_error": nerr,
        "dE_over_E": {
            "integral": "Jacobi" if w["omega"] != 0.0 else "energy",
            "max": float(dem.max()),
            "p50": float(numpy.median(dem)),
            "p99": float(numpy.quantile(dem, 0.99)),
            "frac_le_1e-8": float((dem <= 1e-8).mean()),
        },
        "kernel_ms": kernel_ms,
    }
    kernel_s = numpy.mean(kernel_ms) / 1e3
    del result, err, stats, e0, demax
    torch.cuda.empty_cache()
    if not args.no_serving and args.norbits is None and w["name"] == "mw2014_dop853_1m":
        rates = [float(r) for r in args.rates.split(",")]
        sweep = serving_leg(args.nreq, rates, args.budget)
        out["latency"] = {
            "p50_s": {str(m["rate"]): m.get("p50_latency_s") for m in sweep},
            "p99_s": {str(m["rate"]): m.get("p99_latency_s") for m in sweep},
            "orbit_steps_per_s": {str(m["rate"]): m["orbit_steps_per_s"] for m in sweep},
            "offered_req_per_s": {str(m["rate"]): m.get("offered_req_per_s") for m in sweep},
            "achieved_req_per_s": {str(m["rate"]): m["achieved_req_per_s"] for m in sweep},
            "completed": {str(m["rate"]): m["completed"] for m in sweep},
            "config": {"workload": "serving_poisson", "requests_per_rate": args.nreq, "rates_req_per_s": rates,
                       "orbits_per_request": "log-uniform [1, 1e5]", "length": "U(1, 1000), 2..501 outputs",
                       "potentials": "MWPotential2014 / Hernquist(a~U(0.5,3)) / LogarithmicHalo(q~U(0.7,1)) / "
                                     "LogarithmicHalo(q=0.9)+DehnenBar(omegab=1.85)",
                       "methods": "symplec4 (dt=0.01) / dop853 (rtol=atol=1e-12) 50/50",
                       "budget_steps_per_iteration": args.budget,
                       "latency": "result in host memory - arrival"},
        }
    if not args.no_cpu_baseline:
        cpu, opc = cpu_baseline_leg(w, min(args.cpu_sample, n), min(args.count_sample, n), args.cpu_reps)
        out["cpu_baseline"] = cpu
        achieved = opc["flop_per_orbit_step"] * steps_per_pass / kernel_s / 1e12
        out["roofline"] = {
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": MEASURED_TRAFFIC.get(w["name"]) if args.norbits is None else None,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "algorithmic_bytes": float(n * nt * 8 * ncol + 2 * n * 8 * ncol),
            "kernel": {"dop853_hip": "gh_dop853_kernel", "rk4_hip": "gh_rk_kernel<4>",
                       "rk6_hip": "gh_rk_kernel<6>", "dopr54_hip": "gh_dopr54_persistent",
                       "ias15_hip": "gh_ias15_persistent"}.get(
                           w["method"], "gh_symplectic_kernel"),
            "note": ("FP64 vector ALU bound (no MFMA: nothing in the path is a contraction; "
                     "78.6 TFLOP/s is MI355X's FP64 peak). achieved = counted algorithmic FP64 ops "
                     "per orbit-step x orbit-steps per launch / mean launch time; +,-,*,/,sqrt and "
                     "each exp/log/pow/sin/cos/atan2 call count as 1 op"),
            "flop_per_orbit_step": opc["flop_per_orbit_step"],
            "special_per_orbit_step": opc["special_per_orbit_step"],
            "rhs_per_orbit_step": opc["rhs_per_orbit_step"],
            "output_gbs": n * nt * 8 * ncol / kernel_s / 1e9,
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
