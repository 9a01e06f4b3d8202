"""Serving: continuous batching of orbit-integration requests on one GPU.

Each request is one galpy-style integration (N orbits, a potential, an
output grid, leapfrog/symplec4/dop853; galpy/orbit/Orbits.py:1987-2061).
The server keeps every admitted request's inputs, outputs and per-orbit
integrator state resident in HBM and runs iterations: each iteration is one
bounded persistent launch (``gh_batch_launch``) over the unfinished tiles of
all active requests, every orbit advancing at most ``budget`` steps.
Between iterations finished requests are retired and queued requests are
admitted in arrival order while their buffers fit in the HBM budget -- so
work that arrives while other requests are on the GPU joins the next launch
and shares it.

Retiring a request frees its inputs and integrator state at once; its result
rows, error codes and step counts drain to pinned host memory on a copy
stream that overlaps the following launches, and their HBM stays counted
against the budget until that copy has completed (``drain=False`` keeps
results on the device, counted until ``release``).  Latency = the result on
the host (or, without draining, the last launch done) - arrival.
"""

import collections
import time

import numpy

from . import _lib
from .batching import GROUP_DTYPE, TILE_DTYPE, PotentialCache, _lib_batch, launch_flags
from .integrate import DT_SENTINEL, METHOD_CODES, _parse_tol


class Request:
    """One integration request (rows (R,vR,vT,z,vz,phi), output times t)."""

    _next_id = 0

    def __init__(self, pot, yo, t, method="symplec4_hip", dt=None, rtol=None, atol=None, arrival=0.0):
        self.id = Request._next_id
        Request._next_id += 1
        self.pot = pot
        self.yo = numpy.ascontiguousarray(yo, dtype=numpy.float64)
        self.t = numpy.ascontiguousarray(t, dtype=numpy.float64)
        self.method = method
        self.dt = dt
        self.rtol, self.atol = rtol, atol
        self.arrival = float(arrival)
        self.n = len(self.yo)
        self.nt = len(self.t)
        self.admitted_at = None
        self.done_at = None
        self.result = self.err = self.stats = None
        self.status = "queued"

    def nbytes(self, state_bytes):
        """HBM the request holds while active."""
        return self.input_bytes(state_bytes) + self.output_bytes()

    def input_bytes(self, state_bytes):
        """HBM freed when the request retires: initial states and integrator state."""
        return self.n * (48 + state_bytes)

    def output_bytes(self):
        """HBM held until the results leave the device: rows, err, stats, t."""
        return self.n * self.nt * 48 + self.n * (4 + 24) + self.nt * 8

    @property
    def latency(self):
        return None if self.done_at is None else self.done_at - self.arrival


def admission_order(queue_bytes, in_use, capacity):
    """FIFO admission against a memory budget: how many of the queued
    requests (their byte sizes, in arrival order) are admitted now.  A
    request never jumps ahead of an earlier one that does not fit."""
    k = 0
    for b in queue_bytes:
        if in_use + b > capacity:
            break
        in_use += b
        k += 1
    return k


class Server:
    """Continuous-batching orbit integrator.

    budget: steps per orbit per iteration; max_tiles: tiles per iteration
    (None: all active); hbm_bytes: memory budget for request buffers (None:
    ``hbm_fraction`` of the free device memory at start)."""

    def __init__(self, budget=4000, opl=1, max_tiles=None, hbm_bytes=None, hbm_fraction=0.8,
                 device="cuda", drain=True):
        import torch

        self.torch = torch
        self.device = torch.device(device)
        self.budget = int(budget)
        self.opl = int(opl)
        self.max_tiles = max_tiles
        self.lib = _lib_batch()
        self.state_bytes = int(self.lib.gh_batch_state_bytes())
        if hbm_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_bytes = int(free * hbm_fraction)
        self.capacity = int(hbm_bytes)
        self.in_use = 0
        self.pots = PotentialCache()
        self.queue = collections.deque()
        self.active = []
        self.finished = []
        self.rejected = []
        self.iterations = 0
        self.d_counter = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.per_tile = 64 * self.opl
        self.drain = bool(drain)
        self.draining = []  # (request, copy-done event)
        self.copy_stream = torch.cuda.Stream(device=self.device) if self.drain else None
        self.peak_in_use = 0

    # -- request lifecycle ----------------------------------------------------
    def submit(self, req):
        if req.nbytes(self.state_bytes) > self.capacity:
            req.status = "rejected"
            self.rejected.append(req)
            return
        self.queue.append(req)

    def _admit(self, now):
        k = admission_order([r.nbytes(self.state_bytes) for r in self.queue], self.in_use, self.capacity)
        torch = self.torch
        for _ in range(k):
            r = self.queue.popleft()
            tab, pt, pa = self.pots.get(r.pot)
            r.d_yo = torch.as_tensor(r.yo, device=self.device).clone()
            r.d_t = torch.as_tensor(r.t, device=self.device)
            r.result = torch.empty((r.n, r.nt, 6), dtype=torch.float64, device=self.device)
            r.err = torch.zeros(r.n, dtype=torch.int32, device=self.device)
            r.stats = torch.zeros((r.n, 3), dtype=torch.float64, device=self.device)
            r.d_state = torch.zeros(r.n * self.state_bytes, dtype=torch.uint8, device=self.device)  # status 0
            rtol, atol = _parse_tol(r.rtol, r.atol)
            rec = numpy.zeros(1, dtype=GROUP_DTYPE)
            rec[0] = (self.lib.gh_potential_device_ptr(tab.handle), r.d_t.data_ptr(), r.d_yo.data_ptr(),
                      r.result.data_ptr(), r.err.data_ptr(), r.stats.data_ptr(), r.d_state.data_ptr(),
                      DT_SENTINEL if r.dt is None else float(r.dt), rtol, atol, r.n, r.nt,
                      METHOD_CODES[r.method], 0)
            r.record = rec
            r.table = tab
            r.pot_type = pt
            firsts = numpy.arange(0, r.n, self.per_tile)
            r.tiles = numpy.stack([firsts, numpy.minimum(self.per_tile, r.n - firsts)], axis=1)
            r.tile_open = numpy.ones(len(firsts), dtype=bool)
            r.admitted_at = now
            r.status = "active"
            self.in_use += r.nbytes(self.state_bytes)
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            self.active.append(r)

    def _retire(self, r, now):
        r.d_yo = r.d_state = None
        self.in_use -= r.input_bytes(self.state_bytes)
        if not self.drain:
            r.done_at = now
            r.status = "done"
            self.finished.append(r)
            return
        torch = self.torch
        host = [torch.empty(x.shape, dtype=x.dtype, pin_memory=True) for x in (r.result, r.err, r.stats)]
        self.copy_stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.copy_stream):
            for h, x in zip(host, (r.result, r.err, r.stats)):
                h.copy_(x, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
        r.host = host
        r.status = "draining"
        self.draining.append((r, ev))

    def _poll_drains(self, now_fn, wait=False):
        """Finish the requests whose results reached the host (with wait,
        block for the oldest one first)."""
        if wait and self.draining:
            self.draining[0][1].synchronize()
        keep = []
        for r, ev in self.draining:
            if ev.query():
                # the device buffers were kept alive by these references until now
                r.result, r.err, r.stats = r.host
                r.host = None
                r.d_t = None
                self.in_use -= r.output_bytes()
                r.done_at = now_fn()
                r.status = "done"
                self.finished.append(r)
            else:
                keep.append((r, ev))
        self.draining = keep

    def release(self, r):
        """Drop a finished request's device results (drain=False only)."""
        if self.drain or r.status != "done" or r.result is None or not r.result.is_cuda:
            return
        r.result = r.err = r.stats = r.d_t = None
        self.in_use -= r.output_bytes()

    # -- one iteration ------------------------------------------------------------
    def step(self, now_fn):
        torch = self.torch
        if not self.active:
            return False
        recs = numpy.concatenate([r.record for r in self.active])
        tiles, owners = [], []
        for gi, r in enumerate(self.active):
            idx = numpy.nonzero(r.tile_open)[0]
            for k in idx:
                tiles.append((gi, int(r.tiles[k, 0]), int(r.tiles[k, 1]), 0))
                owners.append((gi, k))
            if self.max_tiles is not None and len(tiles) >= self.max_tiles:
                break
        T = numpy.array(tiles, dtype=TILE_DTYPE)
        d_groups = torch.from_numpy(recs.view(numpy.uint8).copy()).to(self.device)
        d_tiles = torch.from_numpy(T.view(numpy.uint8).copy()).to(self.device)
        d_done = torch.zeros(len(T), dtype=torch.int32, device=self.device)
        flags = launch_flags([METHOD_CODES[r.method] for r in self.active], [r.pot_type for r in self.active])
        s = torch.cuda.current_stream(self.device)
        rc = self.lib.gh_batch_launch(d_groups.data_ptr(), d_tiles.data_ptr(), len(T), d_done.data_ptr(),
                                      self.d_counter.data_ptr(), None, self.budget, self.opl, flags,
                                      s.cuda_stream)
        _lib.check(rc, "gh_batch_launch")
        done = d_done.cpu().numpy()  # waits for the launch
        self.iterations += 1
        now = now_fn()
        for (gi, k), d in zip(owners, done):
            if d:
                self.active[gi].tile_open[k] = False
        still = []
        for r in self.active:
            if r.tile_open.any():
                still.append(r)
            else:
                self._retire(r, now)
        self.active = still
        return True

    # -- driving loop --------------------------------------------------------------
    def run(self, requests, realtime=True):
        """Serve `requests` (sorted by arrival): arrivals are released at
        their arrival times (realtime=True) or all at once."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        now_fn = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))
        clock = lambda: time.perf_counter() - t0
        i = 0
        while True:
            now = now_fn() if realtime else clock()
            while i < len(reqs) and (not realtime or reqs[i].arrival <= now):
                self.submit(reqs[i])
                i += 1
            self._poll_drains(clock)
            self._admit(now)
            if self.active:
                self.step(clock)
            elif self.queue and self.draining:
                self._poll_drains(clock, wait=True)  # admission waits for memory
            elif i < len(reqs):
                time.sleep(max(0.0, min(reqs[i].arrival - now, 0.005)))
            elif self.draining:
                self._poll_drains(clock, wait=True)
            elif not self.queue:
                break
        self.makespan = time.perf_counter() - t0
        return self.metrics(reqs)

    def metrics(self, reqs):
        lat = numpy.array([r.latency for r in self.finished])
        steps = float(sum(r.stats[:, 0].sum().item() for r in self.finished))
        span = max(self.makespan, 1e-9)
        out = {
            "requests": len(reqs),
            "completed": len(self.finished),
            "rejected": len(self.rejected),
            "iterations": self.iterations,
            "makespan_s": span,
            "orbit_steps": steps,
            "orbit_steps_per_s": steps / span,
            "achieved_req_per_s": len(self.finished) / span,
        }
        if len(lat):
            out.update(p50_latency_s=float(numpy.percentile(lat, 50)), p99_latency_s=float(numpy.percentile(lat, 99)),
                       mean_latency_s=float(lat.mean()))
        if len(reqs) > 1:
            arr = numpy.array([r.arrival for r in reqs])
            out["offered_req_per_s"] = (len(reqs) - 1) / max(arr[-1] - arr[0], 1e-9)
        return out


def make_requests(nreq, rate, seed=0, max_orbits=100_000, max_len=1000.0, max_out=501, dt=0.01,
                  size_scale=1.0):
    """Poisson arrivals at `rate` req/s; N log-uniform in [1, max_orbits];
    potential from {MWPotential2014, Hernquist(a~U(0.5,3)),
    LogarithmicHalo(q~U(0.7,1)), LogarithmicHalo(q=0.9)+DehnenBar}; length
    U(1, max_len) with 2..max_out outputs; symplec4 (dt, output grid snapped
    so (long)(init_dt/dt) is exact) or dop853, 50/50."""
    from .potential import HernquistPotential, LogarithmicHaloPotential, MWPotential2014_factory
    from .workloads import ics, loghalo_bar

    rng = numpy.random.default_rng(seed)
    mw, bar = MWPotential2014_factory(), loghalo_bar()
    arrivals = numpy.cumsum(rng.exponential(1.0 / rate, nreq))
    out = []
    for k in range(nreq):
        n = int(numpy.clip(numpy.floor(numpy.exp(rng.uniform(0.0, numpy.log(max_orbits + 1)))), 1, max_orbits))
        n = max(1, int(round(n * size_scale)))
        kind = int(rng.integers(0, 4))
        pot = [mw, [HernquistPotential(a=float(rng.uniform(0.5, 3.0)))],
               [LogarithmicHaloPotential(q=float(rng.uniform(0.7, 1.0)))], bar][kind]
        L = float(rng.uniform(1.0, max_len))
        nout = int(rng.integers(2, max_out + 1))
        method = "symplec4_hip" if rng.uniform() < 0.5 else "dop853_hip"
        if method == "symplec4_hip":
            kk = max(1, int(round(L / (nout - 1) / dt)))
            step = kk * dt
            while int(step / dt) < kk:
                step = numpy.nextafter(step, numpy.inf)
            t = numpy.arange(nout) * step
            rdt = dt
        else:
            t = numpy.linspace(0.0, L, nout)
            rdt = None
        out.append(Request(pot, ics(n, seed=seed * 1000003 + k), t, method=method, dt=rdt,
                           arrival=arrivals[k]))
    return out
