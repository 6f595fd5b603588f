// One scheme's sweep launchers and their kernels (k_sweep, k_sweep_pack),
// compiled as a translation unit of its own with NUC_SWEEP_SCHEME set to the
// scheme id, so that the nine schemes' kernels build in parallel.
#include "nuc_sweep.h"

#ifndef NUC_SWEEP_SCHEME
#error "NUC_SWEEP_SCHEME (the scheme id, 0-8) must be defined"
#endif

namespace nucsweep {
template void launch_sweep_s<NUC_SWEEP_SCHEME>(int, const SweepArgs &, hipStream_t);
template void launch_sweep_pack_s<NUC_SWEEP_SCHEME>(int, int, const PackLaunch &, hipStream_t);
}  // namespace nucsweep
