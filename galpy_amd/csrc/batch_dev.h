// Batched orbit integration: many independent groups of orbits (each with its
// own potential, time grid, step and method) in one persistent launch.
//
// Work is split into tiles of 64 * OPL orbits of one group.  A wavefront
// claims a tile with one atomic, so the group -- and with it the potential
// table -- is uniform across the wavefront and every parameter comes through
// scalar loads.  Each orbit keeps a resumable state record in HBM: a launch
// advances every claimed tile by at most `budget` steps per orbit, stores the
// state and moves on, so a sequence of bounded launches (captured once in a
// hipGraph) or a serving loop that admits new groups between launches can
// interleave work from many requests.  OPL (orbits per lane, 1 or 2) gives
// the symplectic steppers independent instruction streams to hide FP64
// latency; the autotuner picks it.
//
// The per-orbit arithmetic is that of galpy's integrators
// (bovy_symplecticode.c:132-377 for leapfrog/symplec4, leung_dop853.c:231-517
// for DOP853), restated as one step per loop trip so that state can be saved
// between any two steps.
#pragma once
#include "../../include/galpy_hip.h"
#include "dop853_dev.h"

// Resumable per-orbit state (one record per orbit of the batch).
struct GhLaneState {
  GhDopState d;       // DOP853 state; symplectic: d.y = (q, p), d.t_cur = time,
                      // d.h = dt, d.done = output rows written, d.nacc = steps
  long long jj;       // symplectic: step inside the current output interval
  long long ndt;      // symplectic: steps per output interval
  int status;         // 0 not started, 1 running, 2 finished
  int err;
  int pad[2];
};

__device__ __forceinline__ int gh_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Start one orbit of a group: cylindrical -> rectangular, row 0, step choice.
template <bool DOP, class F>
__device__ void gh_batch_start(const F& Fo, const gh_batch_group& g, long o, GhLaneState& s) {
  double v[6], w[6];
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = g.yo[o * 6 + k];
  gh_cyl_to_rect(v, w);
  double* out = g.result + o * (long)g.nt * 6;
  s.err = 0;
  s.status = 1;
  if (DOP && g.method == 6) {
#pragma unroll
    for (int k = 0; k < 6; k++) s.d.y[k] = w[k];
    gh_dop_init(Fo, g.t, g.nt, exp(g.rtol), exp(g.atol), s.d, out);
    return;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) g.yo[o * 6 + k] = w[k];  // galpy converts yo in place
  gh_store_cyl(out, w[0], w[1], w[2], w[3], w[4], w[5]);
  const GhVec3 q{w[0], w[1], w[2]}, p{w[3], w[4], w[5]};
  const double init_dt = g.t[1] - g.t[0];
  double dt = g.dt;
  if (dt == GH_DT_SENTINEL)
    dt = (g.method == 0) ? gh_leapfrog_estimate_step(Fo, q, p, init_dt, g.t[0], g.rtol, g.atol)
                         : gh_symplec4_estimate_step(Fo, q, p, init_dt, g.t[0], g.rtol, g.atol);
#pragma unroll
  for (int k = 0; k < 6; k++) s.d.y[k] = w[k];
  s.d.h = dt;
  s.d.t_cur = g.t[0];
  s.d.done = 0;
  s.d.nacc = 0;
  s.d.nrej = 0;
  s.jj = 0;
  s.ndt = (long long)(init_dt / dt);
}

// One symplectic step of galpy's leapfrog (METHOD 0) or symplec4 (METHOD 3).
// Interval structure of bovy_symplecticode.c:197-215 / :312-365: the opening
// drift of each output interval, fused (c4+c1) drifts between the steps of an
// interval, and a closing drift before the row is saved.
template <int METHOD, class F>
__device__ __forceinline__ void gh_symp_step(const F& Fo, const gh_batch_group& g, long o,
                                             GhLaneState& s) {
  GhVec3 q{s.d.y[0], s.d.y[1], s.d.y[2]}, p{s.d.y[3], s.d.y[4], s.d.y[5]};
  const double dt = s.d.h;
  double to = s.d.t_cur;
  // ndt = 0 (dt longer than the output spacing) still takes one step per
  // output, as galpy's loop structure does
  const bool last = s.jj >= s.ndt - 1;
  if (METHOD == 0) {
    if (s.jj == 0) q = gh_drift(q, p, dt / 2.0);
    p = gh_drift(p, gh_accel(Fo, to + dt / 2.0, q), dt);
    q = gh_drift(q, p, last ? dt / 2.0 : dt);
    to = to + dt;
  } else {
    const double c1 = GH_S4C1, c2 = GH_S4C2, c3 = GH_S4C3, c4 = GH_S4C4;
    const double d1 = GH_S4D1, d2 = GH_S4D2, d3 = GH_S4D3;
    if (s.jj == 0) {
      q = gh_drift(q, p, c1 * dt);
      to += c1 * dt;
    }
    p = gh_drift(p, gh_accel(Fo, to, q), d1 * dt);
    q = gh_drift(q, p, c2 * dt);
    to += c2 * dt;
    p = gh_drift(p, gh_accel(Fo, to, q), d2 * dt);
    q = gh_drift(q, p, c3 * dt);
    to += c3 * dt;
    p = gh_drift(p, gh_accel(Fo, to, q), d3 * dt);
    if (last) {
      q = gh_drift(q, p, c4 * dt);
      to += c4 * dt;
    } else {
      q = gh_drift(q, p, (c4 + c1) * dt);
      to += (c4 + c1) * dt;
    }
  }
  s.d.nacc++;
  if (last) {
    s.d.done++;
    gh_store_cyl(g.result + (o * (long)g.nt + s.d.done) * 6, q.x, q.y, q.z, p.x, p.y, p.z);
    s.jj = 0;
  } else {
    s.jj++;
  }
  s.d.y[0] = q.x; s.d.y[1] = q.y; s.d.y[2] = q.z;
  s.d.y[3] = p.x; s.d.y[4] = p.y; s.d.y[5] = p.z;
  s.d.t_cur = to;
}

__device__ __forceinline__ void gh_batch_finish(const gh_batch_group& g, long o, GhLaneState& s) {
  if (g.method == 6)
    for (int k = 0; k < 6; k++) g.yo[o * 6 + k] = s.d.y[k];  // galpy's y0 holds the last state
  g.err[o] = s.err;
  if (g.stats) {
    g.stats[o * 3 + 0] = (double)s.d.nacc;
    g.stats[o * 3 + 1] = (double)s.d.nrej;
    g.stats[o * 3 + 2] = (g.method == 6) ? 0.0 : s.d.h;
  }
  s.status = 2;
}

// Advance the orbits of one tile by at most `budget` steps each, with the
// force of composition Spec; returns the number of orbits of the tile (OPL
// per lane) still unfinished in this lane.
template <int OPL, bool DOP, class Spec>
__device__ int gh_batch_run_tile(const gh_batch_group& g, const gh_batch_tile& tile,
                             long long budget, const volatile int* cancel) {
  GhLaneState* __restrict__ state = (GhLaneState*)g.state;
  const GhForce<Spec> Fo{(const GhPotTable*)g.table};
  const int lane = threadIdx.x & 63;
  GhLaneState s[OPL];
  long o[OPL];
  bool act[OPL];
#pragma unroll
  for (int u = 0; u < OPL; u++) {
    const int k = lane + 64 * u;
    act[u] = k < tile.count;
    o[u] = tile.first + k;
    if (act[u]) {
      s[u] = state[tile.first + k];
      if (s[u].status == 0) gh_batch_start<DOP>(Fo, g, o[u], s[u]);
      act[u] = s[u].status == 1;
    }
  }
  const int method = gh_uniform(g.method);
  bool stop = false;
  for (long long it = 0; it < budget; it++) {
    bool any = false;
#pragma unroll
    for (int u = 0; u < OPL; u++) any |= act[u];
    if (!__any(any)) break;
    if (cancel != nullptr && (it & 255) == 0 && *cancel) stop = true;
    if (stop) {
#pragma unroll
      for (int u = 0; u < OPL; u++)
        if (act[u]) {
          s[u].err = -10;
          gh_batch_finish(g, o[u], s[u]);
          act[u] = false;
        }
      break;
    }
    if (DOP && method == 6) {
#pragma unroll
      for (int u = 0; u < OPL; u++)
        if (act[u] && gh_dop_step(Fo, g.t, g.nt, exp(g.rtol), exp(g.atol), s[u].d,
                                  g.result + o[u] * (long)g.nt * 6)) {
          gh_batch_finish(g, o[u], s[u]);
          act[u] = false;
        }
    } else {
#pragma unroll
      for (int u = 0; u < OPL; u++) {
        if (!act[u]) continue;
        if (method == 0)
          gh_symp_step<0>(Fo, g, o[u], s[u]);
        else
          gh_symp_step<3>(Fo, g, o[u], s[u]);
        if (s[u].d.done >= g.nt - 1) {
          gh_batch_finish(g, o[u], s[u]);
          act[u] = false;
        }
      }
    }
  }
  int left = 0;
#pragma unroll
  for (int u = 0; u < OPL; u++) {
    const int k = lane + 64 * u;
    if (k < tile.count) {
      state[tile.first + k] = s[u];
      left += s[u].status != 2;
    }
  }
  return left;
}

// Compositions with a specialised force: the same ones the single-group
// kernels specialise (MWPotential2014, LogarithmicHalo + DehnenBar,
// Hernquist, LogarithmicHalo); 0 = any other; compositions with wrapped
// components run their own run-time dispatch instance.
enum GhSpecCode { GH_SPEC_ANY = 0, GH_SPEC_MW2014, GH_SPEC_LOGBAR, GH_SPEC_HERN, GH_SPEC_LOG,
                  GH_SPEC_WRAPPED };

__device__ __forceinline__ int gh_spec_code(const GhPotTable* __restrict__ T) {
  if (T->has_wrap) return GH_SPEC_WRAPPED;
  const int n = T->npot;
  const int c0 = T->comp[0].type;
  if (n == 3 && c0 == GH_POWERSPHCUTOFF && T->comp[1].type == GH_MIYAMOTONAGAI &&
      T->comp[2].type == GH_NFW)
    return GH_SPEC_MW2014;
  if (n == 2 && c0 == GH_LOGHALO && T->comp[1].type == GH_DEHNENBAR) return GH_SPEC_LOGBAR;
  if (n == 1 && c0 == GH_HERNQUIST) return GH_SPEC_HERN;
  if (n == 1 && c0 == GH_LOGHALO) return GH_SPEC_LOG;
  return GH_SPEC_ANY;
}

// A tile whose composition has no specialised force in a kernel built
// without the generic one: its orbits finish at once with err = -1.
__device__ __forceinline__ int gh_batch_reject_tile(const gh_batch_group& g,
                                                    const gh_batch_tile& tile) {
  GhLaneState* __restrict__ state = (GhLaneState*)g.state;
  for (int k = threadIdx.x & 63; k < tile.count; k += 64) {
    g.err[tile.first + k] = -1;
    state[tile.first + k].status = 2;
  }
  return 0;
}

// Persistent batch kernel: wavefronts claim tiles until the queue is empty.
// tile_done[i] != 0 marks a finished tile (skipped by later launches);
// *remaining counts unfinished tiles.  DOP = the batch has DOP853 groups
// (its register demand is far above the symplectic steppers', so
// symplectic-only batches use the instance without it).  Each tile runs the
// force specialised for its group's composition (a wavefront-uniform
// branch); GEN = the instance also carries the run-time dispatch force for
// other compositions, whose register demand would otherwise set the whole
// kernel's occupancy.  Occupancy targets: 3 waves per SIMD with one orbit
// per lane, 2 with two or with DOP853 (which then spills a few registers).
template <int OPL, bool DOP, bool GEN>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(DOP ? 2 : 4 - OPL, DOP ? 2 : 4 - OPL)))
gh_batch_kernel(const gh_batch_group* __restrict__ groups, const gh_batch_tile* __restrict__ tiles,
                int ntiles, int* __restrict__ tile_done,
                unsigned long long* __restrict__ counter, int* __restrict__ remaining,
                long long budget, const volatile int* cancel) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    unsigned long long ti = 0;
    if (lane == 0) ti = atomicAdd(counter, 1ull);
    const int t = gh_uniform((int)__shfl(ti, 0));
    if (t >= ntiles) break;
    if (tile_done[t]) continue;
    const gh_batch_tile tile = tiles[t];
    const gh_batch_group g = groups[gh_uniform(tile.group)];
    int left;
    switch (gh_uniform(gh_spec_code((const GhPotTable*)g.table))) {
      case GH_SPEC_MW2014:
        left = gh_batch_run_tile<OPL, DOP, GhSpec<GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW>>(
            g, tile, budget, cancel);
        break;
      case GH_SPEC_LOGBAR:
        left = gh_batch_run_tile<OPL, DOP, GhSpec<GH_LOGHALO, GH_DEHNENBAR>>(g, tile, budget, cancel);
        break;
      case GH_SPEC_HERN:
        left = gh_batch_run_tile<OPL, DOP, GhSpec<GH_HERNQUIST>>(g, tile, budget, cancel);
        break;
      case GH_SPEC_LOG:
        left = gh_batch_run_tile<OPL, DOP, GhSpec<GH_LOGHALO>>(g, tile, budget, cancel);
        break;
      case GH_SPEC_WRAPPED:
        if constexpr (GEN)
          left = gh_batch_run_tile<OPL, DOP, GhSpecWrapped>(g, tile, budget, cancel);
        else
          left = gh_batch_reject_tile(g, tile);
        break;
      default:
        if constexpr (GEN)
          left = gh_batch_run_tile<OPL, DOP, GhSpec<>>(g, tile, budget, cancel);
        else
          left = gh_batch_reject_tile(g, tile);
        break;
    }
    if (!__any(left != 0) && lane == 0) {
      tile_done[t] = 1;
      if (remaining) atomicSub(remaining, 1);
    }
  }
}
