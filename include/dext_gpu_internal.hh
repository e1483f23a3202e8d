// The hooks between the backend-independent part of namespace gpu (the
// window policy, the announcements, the per-thread stage times) and a device
// backend: declaration for declaration src/spots/dext_gpu_internal.hh:18-40.
// In the reference build the policy part is the reference's own
// src/spots/dext_gpu.cc and this backend supplies dext_gpu_hip.cc; built on
// its own, the policy part is dext_gpu_policy.cc.
#ifndef SPOTFINDER_DEXT_GPU_INTERNAL_HH
#define SPOTFINDER_DEXT_GPU_INTERNAL_HH

#include "dext_gpu.hh"

namespace gpu {
namespace internal {

// Records the split for the calling thread, which last_stage_times() returns.
void set_stage_times(const StageTimes &times);

// Name the configuration on stderr when it changes (report_windows(false)
// silences them).
void announce_windows(Window stage0, Window stage2);
void announce_fused();

// The backend's measured preference when nothing has asked for a window.
Window default_stage0_window();
Window default_stage2_window();

} // namespace internal
} // namespace gpu

#endif
