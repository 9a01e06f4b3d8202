// Host-side declarations shared by the engine's translation units.
//
// The library is compiled as several translation units in parallel: one
// host unit (potential decode, C-ABI entry points, host-buffer driver), one
// unit per (force composition, orbit kind) pair holding that pair's
// integrator kernels, and units for the batch and adjoint kernels.  Each
// kernel is instantiated in exactly one unit; launch_spec<Spec, V> is the
// only way in from the host unit.
#pragma once
#include <hip/hip_runtime.h>

#include <initializer_list>

#include "../../include/galpy_hip.h"
#include "integrators_dev.h"

#define GH_EXPORT extern "C" __attribute__((visibility("default")))

namespace gh {

constexpr int kBlock = 64;  // one wavefront per workgroup

// Last error message of the calling thread (galpy_hip_last_error).
extern thread_local char g_errbuf[512];
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

#define GH_CHECK(expr)                                                          \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ::gh::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                      __FILE__, __LINE__);                                      \
      return -1;                                                                \
    }                                                                           \
  } while (0)

// Interrupt flag polled by running kernels: fine-grained pinned host memory,
// set by SIGINT during a blocking host call or by galpy_hip_set_interrupt.
extern volatile int* g_irq_host;
int* irq_device_ptr();

// Next work counter (one u64 in device memory) of device `dev` for a
// persistent kernel; nullptr if the pool cannot be allocated.
unsigned long long* next_counter(int dev);

// Host handle of an uploaded potential: the device table plus the component
// codes, so launches can pick a kernel specialised for the composition.
struct GhPotHandle {
  GhPotTable* dev;
  int npot;
  int types[GH_MAX_POT];
  int nonaxi[GH_MAX_POT];  // component exerts a phi torque (bar, triaxial halo)
  int has_wrap;            // wrapped or heavy components: extended dispatch, no Hessian
  int ndiss;               // velocity-dependent components
};

inline void fill_handle(const GhPotTable& table, GhPotHandle* h) {
  h->npot = table.npot;
  h->has_wrap = table.has_wrap;
  h->ndiss = table.ndiss;
  for (int i = 0; i < GH_MAX_POT; i++) {
    h->types[i] = table.comp[i].type;
    h->nonaxi[i] = table.comp[i].nonaxi;
  }
}

inline bool types_are(const GhPotHandle* h, std::initializer_list<int> codes) {
  if (h->has_wrap || (int)codes.size() != h->npot) return false;
  int i = 0;
  for (int c : codes)
    if (h->types[i++] != c) return false;
  return true;
}

// Components with a Hessian-vector product on the device (integrate_dxdv's
// kick of the deviation): every unwrapped type but KuzminKutuzovStaeckel,
// SoftenedNeedleBar and the planar-only disks and spiral.
inline bool hessian_capable(const GhPotHandle* h) {
  if (h->has_wrap) return false;  // time-dependent (and rotating) wrappers
  for (int i = 0; i < h->npot; i++) {
    switch (h->types[i]) {  // the components with a Hessian-vector product
      case GH_POWERSPHCUTOFF: case GH_MIYAMOTONAGAI: case GH_NFW: case GH_HERNQUIST:
      case GH_LOGHALO: case GH_DEHNENBAR: case GH_POWERSPH: case GH_JAFFE: case GH_ISOCHRONE:
      case GH_PLUMMER: case GH_PSEUDOISO: case GH_KUZMINDISK: case GH_FLATPOWER: case GH_BURKERT:
      case GH_DEHNENCORESPH: case GH_DEHNENSPH: case GH_HOMOGSPHERE: case GH_NULLPOT: break;
      default: return false;
    }
  }
  return true;
}

// The adjoint's: a Hessian and time-independent -- no bar either (its kick
// adjoint carries no time).
inline bool adjoint_capable(const GhPotHandle* h) {
  if (!hessian_capable(h)) return false;
  for (int i = 0; i < h->npot; i++)
    if (h->types[i] == GH_DEHNENBAR) return false;
  return true;
}

// Caller-provided device memory for a launch's temporaries (the DOP853
// dense-output overflow area, the pilot's cost keys and claim order with
// the radix sort's temporary, the work counter).  With a workspace a launch
// allocates nothing: it carves these from `ptr` (256-byte aligned, `bytes`
// long) and fails with an error when they do not fit.  `cancel` is the
// interrupt flag the kernels poll (nullable).
struct GhWorkspace {
  char* ptr;
  size_t bytes;
  const int* cancel;
};

// Compositions with kernels of their own.
using SpecMW2014 = GhSpec<GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW>;
using SpecLogBar = GhSpec<GH_LOGHALO, GH_DEHNENBAR>;
using SpecHern = GhSpec<GH_HERNQUIST>;
using SpecLog = GhSpec<GH_LOGHALO>;
using SpecAny = GhSpec<>;

// Launch of galpy integrator `odeint_type` for orbits of kind V (GhVec3 full,
// GhVec2 planar, GhVec6 / GhVec4 their variational systems) in composition
// Spec.  Buffers are device buffers; `ws` (nullable) provides the
// temporaries, which are otherwise allocated stream-ordered.  Returns 0 or
// -1 with the error set.
template <class Spec, class V>
int launch_spec(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                const GhPotTable* d_table, double dt, double rtol, double atol, double* d_result,
                int* d_err, int odeint_type, double* d_stats, hipStream_t stream,
                const GhWorkspace* ws);

#define GH_LAUNCH_SPEC_INSTANCE(EXT, SPEC, V)                                                   \
  EXT template int launch_spec<SPEC, V>(int, double*, int, const double*, int,                  \
                                        const GhPotTable*, double, double, double, double*,    \
                                        int*, int, double*, hipStream_t, const GhWorkspace*)

#define GH_LAUNCH_SPEC_ALL(EXT)                            \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecMW2014, GhVec3);        \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecLogBar, GhVec3);        \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecHern, GhVec3);          \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecLog, GhVec3);           \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecAny, GhVec3);           \
  GH_LAUNCH_SPEC_INSTANCE(EXT, GhSpecWrapped, GhVec3);     \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecMW2014, GhVec2);        \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecLogBar, GhVec2);        \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecHern, GhVec2);          \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecLog, GhVec2);           \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecAny, GhVec2);           \
  GH_LAUNCH_SPEC_INSTANCE(EXT, GhSpecWrapped, GhVec2);     \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecMW2014, GhVec6);        \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecAny, GhVec6);           \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecMW2014, GhVec4);        \
  GH_LAUNCH_SPEC_INSTANCE(EXT, SpecAny, GhVec4)

GH_LAUNCH_SPEC_ALL(extern);

// Full-orbit launch for an uploaded potential (the composition picks the
// kernel instance), with an optional workspace: integrateFullOrbit_hip's
// body.
int launch_full_orbits(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                       const GhPotHandle* h, double dt, double rtol, double atol, double* d_result,
                       int* d_err, int odeint_type, double* d_stats, hipStream_t stream,
                       const GhWorkspace* ws);

// Interrupt flag's device address if it is already set up, else nullptr
// (never allocates).
const int* irq_device_ptr_if_ready();

}  // namespace gh
