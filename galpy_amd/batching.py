"""Batched integration of many orbit groups in one persistent launch.

A *group* is what one ``integrateFullOrbit`` call integrates in galpy
(galpy/orbit/orbit_c_ext/integrateFullOrbit.c:1046-1150): N orbits, one
potential, one output grid, one method.  ``Batch`` runs many groups with
different potentials, lengths, output counts and methods in one persistent
HIP launch (``gh_batch_launch``): groups are cut into tiles of 64 x OPL
orbits, tiles are ordered longest-first (an LPT schedule over the persistent
wavefronts), and every orbit keeps a resumable state record in HBM, so a
launch may stop after a bounded number of steps and a later launch continues.

``Batch.run(budget=..., graph_chunks=k)`` runs bounded launches; with
``graph_chunks`` the sequence of k bounded launches is captured once as a
HIP graph (through ``torch.cuda.CUDAGraph``) and replayed until every tile
has finished.  ``autotune`` times the orbits-per-lane choices on a sample of
the batch and caches the winner per device and workload signature.
"""

import ctypes
import json
import os
import time

import numpy

from . import _lib
from .integrate import DT_SENTINEL, METHOD_CODES, PotentialTable, _parse_tol
from .potential import parse_pot

GROUP_DTYPE = numpy.dtype(
    [
        ("table", "<u8"), ("t", "<u8"), ("yo", "<u8"), ("result", "<u8"), ("err", "<u8"),
        ("stats", "<u8"), ("state", "<u8"), ("dt", "<f8"), ("rtol", "<f8"), ("atol", "<f8"),
        ("n", "<i4"), ("nt", "<i4"), ("method", "<i4"), ("pad", "<i4"),
    ]
)
TILE_DTYPE = numpy.dtype([("group", "<i4"), ("first", "<i4"), ("count", "<i4"), ("pad", "<i4")])
assert GROUP_DTYPE.itemsize == 96 and TILE_DTYPE.itemsize == 16

# relative cost of one force evaluation, for the longest-first tile order
_POT_COST = {15: 4.0, 5: 1.0, 9: 1.5, 8: 0.6, 0: 0.5, 1: 1.5}
_METHOD_FORCES = {0: 1.0, 3: 3.0, 6: 16.0}

_BATCH_SIGS = {
    "gh_batch_state_bytes": (ctypes.c_size_t, []),
    "gh_batch_reset": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3),
    "gh_batch_launch": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3
        + [ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
    ),
}


def _lib_batch():
    lib = _lib.load()
    for name, (res, args) in _BATCH_SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


# gh_batch_launch flags (include/galpy_hip.h)
GH_BATCH_DOP = 1
GH_BATCH_SPECIALISED = 2
# compositions (packed pot_type codes) with a force specialised in the batch
# kernel: MWPotential2014, LogarithmicHalo + DehnenBar, Hernquist,
# LogarithmicHalo
SPECIALISED_CODES = {(15, 5, 9), (0, 1), (8,), (0,)}


def launch_flags(methods, pot_types):
    """gh_batch_launch flags for groups with these method codes and packed
    pot_type arrays."""
    flags = GH_BATCH_DOP if any(m == 6 for m in methods) else 0
    if all(tuple(int(c) for c in pt) in SPECIALISED_CODES for pt in pot_types):
        flags |= GH_BATCH_SPECIALISED
    return flags


class PotentialCache:
    """Device tables deduplicated by packed content (many groups share a
    potential, e.g. MWPotential2014)."""

    def __init__(self):
        self._tables = {}

    def get(self, pot):
        npot, pt, pa, _ = parse_pot(pot)
        key = (tuple(pt.tolist()), pa.tobytes())
        tab = self._tables.get(key)
        if tab is None:
            tab = PotentialTable(pot)
            self._tables[key] = tab
        return tab, pt, pa

    def __len__(self):
        return len(self._tables)


class Batch:
    """Groups of orbits integrated together.

    groups: iterable of dicts with keys ``pot``, ``yo`` ([n,6] cylindrical),
    ``t`` ([nt], equally spaced), ``method`` ('leapfrog_hip', 'symplec4_hip'
    or 'dop853_hip'), optional ``dt`` (None: estimated per orbit), ``rtol``,
    ``atol``.  Outputs stay on the device: ``results[i]`` is the
    ``[n, nt, 6]`` tensor of group i, ``errs[i]`` and ``stats[i]`` likewise.
    """

    def __init__(self, groups, opl=1, device="cuda", potentials=None, order="longest_first"):
        import torch

        self.torch = torch
        self.device = torch.device(device)
        self.opl = int(opl)
        self.pots = potentials if potentials is not None else PotentialCache()
        self.groups = list(groups)
        lib = _lib_batch()
        self.state_bytes = int(lib.gh_batch_state_bytes())
        G = numpy.zeros(len(self.groups), dtype=GROUP_DTYPE)
        self.yo, self.t, self.results, self.errs, self.stats, self.tables = [], [], [], [], [], []
        self.state = []
        costs = []
        methods, pot_types = [], []
        for i, g in enumerate(self.groups):
            tab, pt, pa = self.pots.get(g["pot"])
            pot_types.append(pt)
            yo = torch.as_tensor(numpy.ascontiguousarray(g["yo"], dtype=numpy.float64), device=self.device)
            t = torch.as_tensor(numpy.ascontiguousarray(g["t"], dtype=numpy.float64), device=self.device)
            n, nt = yo.shape[0], t.shape[0]
            if nt < 2:
                raise ValueError(f"group {i}: need at least two output times")
            res = torch.empty((n, nt, 6), dtype=torch.float64, device=self.device)
            err = torch.zeros(n, dtype=torch.int32, device=self.device)
            st = torch.zeros((n, 3), dtype=torch.float64, device=self.device)
            state = torch.empty(n * self.state_bytes, dtype=torch.uint8, device=self.device)
            method = METHOD_CODES[g.get("method", "symplec4_hip")]
            methods.append(method)
            rtol, atol = _parse_tol(g.get("rtol"), g.get("atol"))
            dt = g.get("dt")
            G[i] = (lib.gh_potential_device_ptr(tab.handle), t.data_ptr(), yo.data_ptr(), res.data_ptr(),
                    err.data_ptr(), st.data_ptr(), state.data_ptr(), DT_SENTINEL if dt is None else float(dt),
                    rtol, atol, n, nt, method, 0)
            self.state.append(state)
            self.yo.append(yo)
            self.t.append(t)
            self.results.append(res)
            self.errs.append(err)
            self.stats.append(st)
            self.tables.append(tab)
            span = abs(float(g["t"][-1] - g["t"][0]))
            step = float(dt) if dt is not None else abs(float(g["t"][1] - g["t"][0]))
            per_step = _METHOD_FORCES[method] * sum(_POT_COST.get(int(c), 1.0) for c in pt)
            costs.append(span / step * per_step if method != 6 else span * per_step * 20.0)
        self.G = G
        self.flags = launch_flags(methods, pot_types)
        tiles = []
        per_tile = 64 * self.opl
        nstate = 0
        for i, g in enumerate(self.groups):
            n = len(g["yo"])
            for first in range(0, n, per_tile):
                cnt = min(per_tile, n - first)
                tiles.append((i, first, cnt, 0, costs[i]))
            nstate += n
        if order == "longest_first":
            tiles.sort(key=lambda x: -x[4])
        T = numpy.zeros(len(tiles), dtype=TILE_DTYPE)
        for k, tl in enumerate(tiles):
            T[k] = tl[:4]
        self.T = T
        self.ntiles = len(tiles)
        self.nstate = nstate
        self.d_groups = torch.from_numpy(G.view(numpy.uint8).copy()).to(self.device)
        self.d_tiles = torch.from_numpy(T.view(numpy.uint8).copy()).to(self.device)
        self.d_tile_done = torch.zeros(max(1, self.ntiles), dtype=torch.int32, device=self.device)
        self.d_remaining = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.d_counter = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.norbits = nstate

    # -- launches -----------------------------------------------------------
    def _stream(self, stream):
        return stream if stream is not None else self.torch.cuda.current_stream(self.device)

    def reset(self, stream=None):
        s = self._stream(stream)
        rc = _lib_batch().gh_batch_reset(self.d_groups.data_ptr(), self.d_tiles.data_ptr(), self.ntiles,
                                         self.d_tile_done.data_ptr(), self.d_remaining.data_ptr(),
                                         s.cuda_stream)
        _lib.check(rc, "gh_batch_reset")

    def launch(self, budget=0, stream=None):
        """One persistent launch (asynchronous); budget <= 0 runs to completion."""
        s = self._stream(stream)
        rc = _lib_batch().gh_batch_launch(
            self.d_groups.data_ptr(), self.d_tiles.data_ptr(), self.ntiles,
            self.d_tile_done.data_ptr(), self.d_counter.data_ptr(), self.d_remaining.data_ptr(),
            int(budget), self.opl, self.flags, s.cuda_stream,
        )
        _lib.check(rc, "gh_batch_launch")

    def remaining(self):
        return int(self.d_remaining.item())

    def run(self, budget=0, graph_chunks=None, stream=None, max_launches=1_000_000):
        """Integrate every group.  budget > 0 splits the work into launches of
        at most `budget` steps per orbit; graph_chunks=k captures k such
        launches in one HIP graph and replays it until done."""
        torch = self.torch
        self.reset(stream)
        if budget <= 0:
            self.launch(0, stream)
            return 1
        launches = 0
        if graph_chunks:
            s = torch.cuda.Stream(self.device)
            s.wait_stream(torch.cuda.current_stream(self.device))
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                for _ in range(graph_chunks):
                    self.launch(budget, s)
            torch.cuda.current_stream(self.device).wait_stream(s)
            while self.remaining() > 0 and launches < max_launches:
                g.replay()
                launches += graph_chunks
            self.graph = g
            return launches
        while launches < max_launches:
            self.launch(budget, stream)
            launches += 1
            if self.remaining() == 0:
                break
        return launches

    def orbit_steps(self):
        return float(sum(s[:, 0].sum().item() for s in self.stats))


# ---------------------------------------------------------------------------
# Autotuner
# ---------------------------------------------------------------------------
def _cache_path():
    base = os.environ.get("GALPY_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "galpy_amd"))
    os.makedirs(base, exist_ok=True)
    return os.path.join(base, "autotune.json")


def _signature(groups):
    methods = sorted({g.get("method", "symplec4_hip") for g in groups})
    codes = sorted({tuple(parse_pot(g["pot"])[1].tolist()) for g in groups})
    return f"{'+'.join(methods)}|{codes}"


def autotune(groups, candidates=(1, 2), sample_groups=32, budget=200, device="cuda", use_cache=True):
    """Pick orbits-per-lane for a batch: run a sample of the groups to
    completion with each candidate (after a bounded warm-up launch), compare
    orbit-steps/s, cache the winner per device name and workload signature,
    and return (opl, timings)."""
    import torch

    dev = torch.device(device)
    key = f"{torch.cuda.get_device_name(dev)}|{_signature(groups)}"
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            cache = json.load(open(path))
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return cache[key]["opl"], cache[key]["timings"]
    rng = numpy.random.default_rng(0)
    idx = rng.choice(len(groups), size=min(sample_groups, len(groups)), replace=False)
    sample = [groups[i] for i in sorted(idx)]
    pots = PotentialCache()
    timings = {}
    for opl in candidates:
        b = Batch(sample, opl=opl, device=device, potentials=pots)
        b.run(budget=budget)  # warm-up: code load, first-launch set-up
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        b.run()  # the sample to completion; per-orbit step counts land at the end
        torch.cuda.synchronize(dev)
        el = time.perf_counter() - t0
        steps = b.orbit_steps()
        timings[opl] = steps / el if el > 0 else 0.0
        del b
    best = max(timings, key=timings.get)
    cache[key] = {"opl": int(best), "timings": {str(k): v for k, v in timings.items()}}
    if use_cache:
        try:
            json.dump(cache, open(path, "w"), indent=1)
        except OSError:
            pass
    return int(best), timings


# ---------------------------------------------------------------------------
# The batching workload: 4,096 groups
# ---------------------------------------------------------------------------
def make_groups(ngroups=4096, seed=0, dt=0.01, max_size=10_000, size_scale=1.0):
    """4,096 groups: sizes log-uniform integers in [1, max_size]; potentials
    drawn per group from {MWPotential2014, HernquistPotential(a~U(0.5,3)),
    LogarithmicHaloPotential(q~U(0.7,1))}; lengths L~U(10,1000) with
    nout~U{2..1001} outputs, snapped so that every output interval is an exact
    multiple of dt (the C code truncates init_dt/dt); symplec4 at dt."""
    from .potential import HernquistPotential, LogarithmicHaloPotential, MWPotential2014_factory
    from .workloads import ics

    rng = numpy.random.default_rng(seed)
    mw = MWPotential2014_factory()
    groups = []
    for gi in range(ngroups):
        n = int(numpy.clip(numpy.floor(numpy.exp(rng.uniform(0.0, numpy.log(max_size + 1)))), 1, max_size))
        n = max(1, int(round(n * size_scale)))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            pot = mw
        elif kind == 1:
            pot = [HernquistPotential(a=float(rng.uniform(0.5, 3.0)))]
        else:
            pot = [LogarithmicHaloPotential(q=float(rng.uniform(0.7, 1.0)))]
        L = float(rng.uniform(10.0, 1000.0))
        nout = int(rng.integers(2, 1002))
        k = max(1, int(round(L / (nout - 1) / dt)))  # steps per output interval
        step = k * dt
        while int(step / dt) < k:  # make (long)(init_dt/dt) == k exactly
            step = numpy.nextafter(step, numpy.inf)
        t = numpy.arange(nout) * step
        groups.append(dict(pot=pot, yo=ics(n, seed=seed * 100003 + gi), t=t, method="symplec4_hip", dt=dt))
    return groups
