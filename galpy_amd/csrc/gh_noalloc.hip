// MI355X orbit-integration engine: the allocation-free entry point.
//
// integrateFullOrbit_hip_noalloc is integrateFullOrbit_hip for callers that
// cannot allocate device memory (code under graph capture, tight serving
// loops).  The launch temporaries -- DOP853's dense-output overflow area,
// the pilot's cost keys and longest-first order with the radix sort's
// temporary, and the persistent kernel's work counter -- come from the
// file-scope device array `scratch` instead of a stream-ordered allocation.
// Each GPU holds its own copy of the array (GH_NOALLOC_SCRATCH_BYTES) from
// the library's first use of it on that GPU until the process ends.
//
// Every call on a GPU shares that GPU's array and the array keeps the last
// call's contents; nothing in it needs to start at zero (the counter is set
// by a kernel on the call's stream before each use, the other areas are
// written before they are read).  Calls that overlap on one GPU -- on two
// streams, or from two host threads -- race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gh_internal.h"

namespace {

static __device__ __align__(256) char scratch[GH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
char* g_scratch_ptr[kMaxDevices] = {nullptr};
std::mutex g_scratch_mutex;

// Device address of `scratch` on the current device, looked up once per
// device with that device current (each GPU's copy has its own address).
char* scratch_ptr(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    gh::set_error("device %d outside [0, %d)", dev, kMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      gh::set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                    hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

GH_EXPORT int integrateFullOrbit_hip_noalloc(int nobj, double* d_yo, int nt, const double* d_t,
                                             int indiv_t, const void* pot, double dt, double rtol,
                                             double atol, double* d_result, int* d_err,
                                             int odeint_type, double* d_stats, void* stream) {
  gh::g_errbuf[0] = 0;
  if (pot == nullptr) {
    gh::set_error("potential handle is NULL");
    return -1;
  }
  int dev = 0;
  GH_CHECK(hipGetDevice(&dev));
  char* ptr = scratch_ptr(dev);
  if (ptr == nullptr) return -1;
  const gh::GhWorkspace ws{ptr, (size_t)GH_NOALLOC_SCRATCH_BYTES, gh::irq_device_ptr_if_ready()};
  return gh::launch_full_orbits(nobj, d_yo, nt, d_t, indiv_t, (const gh::GhPotHandle*)pot, dt,
                                rtol, atol, d_result, d_err, odeint_type, d_stats,
                                (hipStream_t)stream, &ws);
}
