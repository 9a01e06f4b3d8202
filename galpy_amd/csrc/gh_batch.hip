// MI355X orbit-integration engine: batching unit.  The persistent kernel
// that runs many groups of orbits (each with its own potential, time grid,
// step and method) from one tile queue, and its C-ABI (gh_batch_*;
// include/galpy_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gh_internal.h"
#include "batch_dev.h"

using namespace gh;

namespace {
__global__ void gh_batch_set_counter(unsigned long long* c, unsigned long long v) { *c = v; }
}  // namespace

namespace {

// one workgroup per tile: mark the tile's orbits unstarted, clear its flag
__global__ void gh_batch_reset_kernel(const gh_batch_group* __restrict__ groups,
                                      const gh_batch_tile* __restrict__ tiles, int ntiles,
                                      int* __restrict__ tile_done, int* __restrict__ remaining) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const gh_batch_tile tile = tiles[t];
  GhLaneState* state = (GhLaneState*)groups[tile.group].state;
  for (int k = threadIdx.x; k < tile.count; k += blockDim.x) state[tile.first + k].status = 0;
  if (threadIdx.x == 0) {
    tile_done[t] = 0;
    if (t == 0 && remaining) *remaining = ntiles;
  }
}

template <int OPL, bool DOP, bool GEN>
int batch_resident_blocks(int dev) {
  static int cached[64] = {0};
  if (dev >= 0 && dev < 64 && cached[dev] > 0) return cached[dev];
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gh_batch_kernel<OPL, DOP, GEN>, 64,
                                                   0) !=
          hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  const int n = std::max(1, per_cu) * std::max(1, ncu);
  if (dev >= 0 && dev < 64) cached[dev] = n;
  return n;
}

}  // namespace

GH_EXPORT size_t gh_batch_state_bytes(void) { return sizeof(GhLaneState); }

GH_EXPORT int gh_batch_reset(const gh_batch_group* d_groups, const gh_batch_tile* d_tiles,
                             int ntiles, int* d_tile_done, int* d_remaining, void* stream) {
  g_errbuf[0] = 0;
  irq_device_ptr();  // set up outside any graph capture of the launches
  if (ntiles <= 0) return 0;
  hipLaunchKernelGGL(gh_batch_reset_kernel, dim3((unsigned)ntiles), dim3(128), 0,
                     (hipStream_t)stream, d_groups, d_tiles, ntiles, d_tile_done, d_remaining);
  GH_CHECK(hipGetLastError());
  return 0;
}

GH_EXPORT int gh_batch_launch(const gh_batch_group* d_groups, const gh_batch_tile* d_tiles,
                              int ntiles, int* d_tile_done, unsigned long long* d_counter,
                              int* d_remaining, long long budget, int opl, int flags,
                              void* stream) {
  g_errbuf[0] = 0;
  if (ntiles <= 0) return 0;
  if (opl != 1 && opl != 2) {
    set_error("orbits per lane must be 1 or 2 (got %d)", opl);
    return -1;
  }
  int dev = 0;
  GH_CHECK(hipGetDevice(&dev));
  const bool dop = (flags & GH_BATCH_DOP) != 0;
  const bool gen = (flags & GH_BATCH_SPECIALISED) == 0;
  const int key = (opl - 1) * 4 + (dop ? 2 : 0) + (gen ? 1 : 0);
  int resident = 0;
  switch (key) {
    case 0: resident = batch_resident_blocks<1, false, false>(dev); break;
    case 1: resident = batch_resident_blocks<1, false, true>(dev); break;
    case 2: resident = batch_resident_blocks<1, true, false>(dev); break;
    case 3: resident = batch_resident_blocks<1, true, true>(dev); break;
    case 4: resident = batch_resident_blocks<2, false, false>(dev); break;
    case 5: resident = batch_resident_blocks<2, false, true>(dev); break;
    case 6: resident = batch_resident_blocks<2, true, false>(dev); break;
    default: resident = batch_resident_blocks<2, true, true>(dev); break;
  }
  if (resident <= 0) {
    set_error("occupancy query failed for the batch kernel");
    return -1;
  }
  const unsigned nblocks = (unsigned)std::min(ntiles, resident);
  const long long b = budget > 0 ? budget : (1ll << 62);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(gh_batch_set_counter, dim3(1), dim3(1), 0, st, d_counter, 0ull);
  const int* cancel = irq_device_ptr();
#define GH_BATCH(O, D, G)                                                                    \
  hipLaunchKernelGGL((gh_batch_kernel<O, D, G>), dim3(nblocks), dim3(64), 0, st, d_groups,  \
                     d_tiles, ntiles, d_tile_done, d_counter, d_remaining, b, cancel)
  switch (key) {
    case 0: GH_BATCH(1, false, false); break;
    case 1: GH_BATCH(1, false, true); break;
    case 2: GH_BATCH(1, true, false); break;
    case 3: GH_BATCH(1, true, true); break;
    case 4: GH_BATCH(2, false, false); break;
    case 5: GH_BATCH(2, false, true); break;
    case 6: GH_BATCH(2, true, false); break;
    default: GH_BATCH(2, true, true); break;
  }
#undef GH_BATCH
  GH_CHECK(hipGetLastError());
  return 0;
}