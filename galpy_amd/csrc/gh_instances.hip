// MI355X orbit-integration engine: integrator kernel instances.  This file
// is compiled once per value of GH_INSTANCE (0 .. GH_NUM_INSTANCES - 1),
// each compile instantiating the kernels of one (force composition, orbit
// kind) pair, so that the instances build in parallel.  The list must match
// GH_LAUNCH_SPEC_ALL in gh_internal.h.
#include "gh_kernels.h"

#ifndef GH_INSTANCE
#error "compile with -DGH_INSTANCE=<index>"
#endif

namespace gh {
#if GH_INSTANCE == 0
GH_LAUNCH_SPEC_INSTANCE(, SpecMW2014, GhVec3);
#elif GH_INSTANCE == 1
GH_LAUNCH_SPEC_INSTANCE(, SpecLogBar, GhVec3);
#elif GH_INSTANCE == 2
GH_LAUNCH_SPEC_INSTANCE(, SpecHern, GhVec3);
#elif GH_INSTANCE == 3
GH_LAUNCH_SPEC_INSTANCE(, SpecLog, GhVec3);
#elif GH_INSTANCE == 4
GH_LAUNCH_SPEC_INSTANCE(, SpecAny, GhVec3);
#elif GH_INSTANCE == 5
GH_LAUNCH_SPEC_INSTANCE(, GhSpecWrapped, GhVec3);
#elif GH_INSTANCE == 6
GH_LAUNCH_SPEC_INSTANCE(, SpecMW2014, GhVec2);
#elif GH_INSTANCE == 7
GH_LAUNCH_SPEC_INSTANCE(, SpecLogBar, GhVec2);
#elif GH_INSTANCE == 8
GH_LAUNCH_SPEC_INSTANCE(, SpecHern, GhVec2);
#elif GH_INSTANCE == 9
GH_LAUNCH_SPEC_INSTANCE(, SpecLog, GhVec2);
#elif GH_INSTANCE == 10
GH_LAUNCH_SPEC_INSTANCE(, SpecAny, GhVec2);
#elif GH_INSTANCE == 11
GH_LAUNCH_SPEC_INSTANCE(, GhSpecWrapped, GhVec2);
#elif GH_INSTANCE == 12
GH_LAUNCH_SPEC_INSTANCE(, SpecMW2014, GhVec6);
#elif GH_INSTANCE == 13
GH_LAUNCH_SPEC_INSTANCE(, SpecAny, GhVec6);
#elif GH_INSTANCE == 14
GH_LAUNCH_SPEC_INSTANCE(, SpecMW2014, GhVec4);
#elif GH_INSTANCE == 15
GH_LAUNCH_SPEC_INSTANCE(, SpecAny, GhVec4);
#else
#error "GH_INSTANCE out of range"
#endif
}  // namespace gh
