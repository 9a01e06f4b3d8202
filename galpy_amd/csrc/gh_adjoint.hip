// MI355X orbit-integration engine: differentiation unit.  Forward flow of
// leapfrog / symplec4 and the reverse-time adjoint that recomputes the
// trajectory by running the steps backwards (C-ABI gh_flow_hip,
// gh_adjoint_hip; include/galpy_hip.h).
#include <hip/hip_runtime.h>

#include "gh_internal.h"
#include "adjoint_dev.h"

using namespace gh;

namespace {

// N full steps of leapfrog (METHOD 0) or symplec4 (METHOD 3) from the
// rectangular state w0 -- galpy's steppers on t = t0 + k dt with dt given,
// one step per output interval (bovy_symplecticode.c:132-377).
template <int METHOD, class Spec>
__global__ void __launch_bounds__(kBlock)
gh_flow_kernel(int n, const double* __restrict__ w0, double* __restrict__ wT, double t0, double dt,
               long long nsteps, const GhPotTable* __restrict__ T) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const GhForce<Spec> F{T};
  GhVec3 q{w0[6 * i], w0[6 * i + 1], w0[6 * i + 2]}, p{w0[6 * i + 3], w0[6 * i + 4], w0[6 * i + 5]};
  double to = t0;
  for (long long s = 0; s < nsteps; s++) {
    if (METHOD == 0) {
      const GhVec3 q12 = gh_drift(q, p, dt / 2.0);
      p = gh_drift(p, gh_accel(F, to + dt / 2.0, q12), dt);
      q = gh_drift(q12, p, dt / 2.0);
      to = to + dt;
    } else {
      q = gh_drift(q, p, GH_S4C1 * dt);
      to += GH_S4C1 * dt;
      p = gh_drift(p, gh_accel(F, to, q), GH_S4D1 * dt);
      q = gh_drift(q, p, GH_S4C2 * dt);
      to += GH_S4C2 * dt;
      p = gh_drift(p, gh_accel(F, to, q), GH_S4D2 * dt);
      q = gh_drift(q, p, GH_S4C3 * dt);
      to += GH_S4C3 * dt;
      p = gh_drift(p, gh_accel(F, to, q), GH_S4D3 * dt);
      q = gh_drift(q, p, GH_S4C4 * dt);
      to += GH_S4C4 * dt;
    }
  }
  wT[6 * i] = q.x; wT[6 * i + 1] = q.y; wT[6 * i + 2] = q.z;
  wT[6 * i + 3] = p.x; wT[6 * i + 4] = p.y; wT[6 * i + 5] = p.z;
}

// Reverse pass: undo the N steps sub-step by sub-step from wT, carrying
// lambda = dL/dw from gT back to t0 and accumulating dL/dtheta per orbit.
template <int METHOD, class Spec>
__global__ void __launch_bounds__(kBlock)
gh_adjoint_kernel(int n, const double* __restrict__ wT, const double* __restrict__ gT, double dt,
                  long long nsteps, const GhPotTable* __restrict__ T, double* __restrict__ lam0,
                  double* __restrict__ w0rec, double* __restrict__ dtheta) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  GhVec3 q{wT[6 * i], wT[6 * i + 1], wT[6 * i + 2]}, p{wT[6 * i + 3], wT[6 * i + 4], wT[6 * i + 5]};
  GhVec3 lq{gT[6 * i], gT[6 * i + 1], gT[6 * i + 2]}, lp{gT[6 * i + 3], gT[6 * i + 4], gT[6 * i + 5]};
  double dth[GH_MAX_POT][GH_ADJ_NPAR] = {};
  auto undo_drift = [&](double h) {
    q = gh_drift(q, p, -h);
    lp = gh_drift(lp, lq, h);
  };
  auto undo_kick = [&](double h) {
    const GhVec3 a = gh_kick_adjoint<Spec>(T, q, h, lp, lq, dth);
    p = gh_drift(p, a, -h);
  };
  for (long long s = 0; s < nsteps; s++) {
    if (METHOD == 0) {
      undo_drift(dt / 2.0);
      undo_kick(dt);
      undo_drift(dt / 2.0);
    } else {
      undo_drift(GH_S4C4 * dt);
      undo_kick(GH_S4D3 * dt);
      undo_drift(GH_S4C3 * dt);
      undo_kick(GH_S4D2 * dt);
      undo_drift(GH_S4C2 * dt);
      undo_kick(GH_S4D1 * dt);
      undo_drift(GH_S4C1 * dt);
    }
  }
  lam0[6 * i] = lq.x; lam0[6 * i + 1] = lq.y; lam0[6 * i + 2] = lq.z;
  lam0[6 * i + 3] = lp.x; lam0[6 * i + 4] = lp.y; lam0[6 * i + 5] = lp.z;
  if (w0rec) {
    w0rec[6 * i] = q.x; w0rec[6 * i + 1] = q.y; w0rec[6 * i + 2] = q.z;
    w0rec[6 * i + 3] = p.x; w0rec[6 * i + 4] = p.y; w0rec[6 * i + 5] = p.z;
  }
  if (dtheta) {
#pragma unroll
    for (int c = 0; c < GH_MAX_POT; c++)
#pragma unroll
      for (int j = 0; j < GH_ADJ_NPAR; j++)
        dtheta[(i * GH_MAX_POT + c) * GH_ADJ_NPAR + j] = dth[c][j];
  }
}


template <class Spec>
void launch_flow(int method, dim3 grid, hipStream_t st, int n, const double* w0, double* wT,
                 double t0, double dt, long long nsteps, const GhPotTable* T) {
  if (method == 0)
    hipLaunchKernelGGL((gh_flow_kernel<0, Spec>), grid, dim3(kBlock), 0, st, n, w0, wT, t0, dt,
                       nsteps, T);
  else
    hipLaunchKernelGGL((gh_flow_kernel<3, Spec>), grid, dim3(kBlock), 0, st, n, w0, wT, t0, dt,
                       nsteps, T);
}

template <class Spec>
void launch_adjoint(int method, dim3 grid, hipStream_t st, int n, const double* wT,
                    const double* gT, double dt, long long nsteps, const GhPotTable* T,
                    double* lam0, double* w0rec, double* dtheta) {
  if (method == 0)
    hipLaunchKernelGGL((gh_adjoint_kernel<0, Spec>), grid, dim3(kBlock), 0, st, n, wT, gT, dt,
                       nsteps, T, lam0, w0rec, dtheta);
  else
    hipLaunchKernelGGL((gh_adjoint_kernel<3, Spec>), grid, dim3(kBlock), 0, st, n, wT, gT, dt,
                       nsteps, T, lam0, w0rec, dtheta);
}

}  // namespace

GH_EXPORT int gh_flow_hip(int n, const double* d_w0, double* d_wT, double t0, double dt,
                          long long nsteps, const void* pot, int method, void* stream) {
  g_errbuf[0] = 0;
  if (n <= 0) return 0;
  if (method != 0 && method != 3) {
    set_error("gh_flow_hip supports leapfrog (0) and symplec4 (3), got %d", method);
    return -1;
  }
  const GhPotHandle* h = (const GhPotHandle*)pot;
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  hipStream_t st = (hipStream_t)stream;
  using MW2014 = GhSpec<GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW>;
  if (types_are(h, {GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW}))
    launch_flow<MW2014>(method, grid, st, n, d_w0, d_wT, t0, dt, nsteps, h->dev);
  else if (h->has_wrap)
    launch_flow<GhSpecWrapped>(method, grid, st, n, d_w0, d_wT, t0, dt, nsteps, h->dev);
  else
    launch_flow<GhSpec<>>(method, grid, st, n, d_w0, d_wT, t0, dt, nsteps, h->dev);
  GH_CHECK(hipGetLastError());
  return 0;
}

GH_EXPORT int gh_adjoint_hip(int n, const double* d_wT, const double* d_gT, double dt,
                             long long nsteps, const void* pot, int method, double* d_lam0,
                             double* d_w0rec, double* d_dtheta, void* stream) {
  g_errbuf[0] = 0;
  if (n <= 0) return 0;
  const GhPotHandle* h = (const GhPotHandle*)pot;
  if (method != 0 && method != 3) {
    set_error("gh_adjoint_hip supports leapfrog (0) and symplec4 (3), got %d", method);
    return -1;
  }
  if (!adjoint_capable(h)) {
    set_error("the adjoint needs time-independent components with a Hessian (every supported "
              "type but DehnenBar, KuzminKutuzovStaeckel, SoftenedNeedleBar and the planar-only "
              "disks, unwrapped)");
    return -1;
  }
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  hipStream_t st = (hipStream_t)stream;
  using MW2014 = GhSpec<GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW>;
  if (types_are(h, {GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW}))
    launch_adjoint<MW2014>(method, grid, st, n, d_wT, d_gT, dt, nsteps, h->dev, d_lam0, d_w0rec,
                           d_dtheta);
  else
    launch_adjoint<GhSpec<>>(method, grid, st, n, d_wT, d_gT, dt, nsteps, h->dev, d_lam0, d_w0rec,
                             d_dtheta);
  GH_CHECK(hipGetLastError());
  return 0;
}