// The backend-independent part of namespace gpu for a standalone build:
// which windows were asked for (with the SPOTFINDER_GPU_STAGE0 / _STAGE2 /
// _FUSED environment overrides and a warning on a value that is neither),
// whether to profile, the per-thread stage times, and the announcement on
// stderr whenever the configuration changes.  The behaviour is that of the
// reference's src/spots/dext_gpu.cc:17-169, which replaces this file when the
// backend is built into the reference (dext_gpu_hip.cc then links against
// it); the code is this library's own.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "dext_gpu_internal.hh"

namespace gpu {
namespace {

std::atomic<Window> chosen0{Window::Tile};
std::atomic<Window> chosen2{Window::Tile};
std::atomic<bool> chosen_fused{false};
std::atomic<bool> profiling{false};
std::atomic<bool> reporting{true};
// The configuration last announced: -1 none, 1 fused, else 2 + 2 * stage0 +
// stage2 (an announcement is made whenever this changes).
std::atomic<int> announced{-1};
std::once_flag chosen_once;

StageTimes &thread_times() {
  static thread_local StageTimes times;
  return times;
}

void resolve_one(const char *suffix, std::atomic<Window> &into) {
  char variable[64];
  std::snprintf(variable, sizeof variable, "SPOTFINDER_GPU_%s", suffix);
  const char *asked = std::getenv(variable);
  if (!asked)
    return;
  if (std::strcmp(asked, "direct") == 0)
    into.store(Window::Direct);
  else if (std::strcmp(asked, "tile") == 0)
    into.store(Window::Tile);
  else
    std::fprintf(stderr,
                 "warning: %s=%s is not 'tile' or 'direct'; leaving it alone\n",
                 variable, asked);
}

void resolve() {
  chosen0.store(internal::default_stage0_window());
  chosen2.store(internal::default_stage2_window());
  resolve_one("STAGE0", chosen0);
  resolve_one("STAGE2", chosen2);
  if (const char *asked = std::getenv("SPOTFINDER_GPU_FUSED")) {
    if (std::strcmp(asked, "1") == 0)
      chosen_fused.store(true);
    else if (std::strcmp(asked, "0") == 0)
      chosen_fused.store(false);
    else
      std::fprintf(stderr,
                   "warning: SPOTFINDER_GPU_FUSED=%s is not '1' or '0'; "
                   "leaving it alone\n",
                   asked);
  }
}

} // namespace

const char *name(Window window) {
  return window == Window::Tile ? "tile" : "direct";
}

Window stage0_window() {
  std::call_once(chosen_once, resolve);
  return chosen0.load(std::memory_order_relaxed);
}
Window stage2_window() {
  std::call_once(chosen_once, resolve);
  return chosen2.load(std::memory_order_relaxed);
}
void stage0_window(Window window) {
  std::call_once(chosen_once, resolve);
  chosen0.store(window, std::memory_order_relaxed);
}
void stage2_window(Window window) {
  std::call_once(chosen_once, resolve);
  chosen2.store(window, std::memory_order_relaxed);
}
bool fused() {
  std::call_once(chosen_once, resolve);
  return chosen_fused.load(std::memory_order_relaxed);
}
void fused(bool on) {
  std::call_once(chosen_once, resolve);
  chosen_fused.store(on, std::memory_order_relaxed);
}

void profile_stages(bool on) { profiling.store(on, std::memory_order_relaxed); }
bool profile_stages() { return profiling.load(std::memory_order_relaxed); }
StageTimes last_stage_times() { return thread_times(); }
void report_windows(bool on) { reporting.store(on, std::memory_order_relaxed); }

namespace internal {

void set_stage_times(const StageTimes &times) { thread_times() = times; }

void announce_windows(Window stage0, Window stage2) {
  const int key = 2 + 2 * static_cast<int>(stage0) + static_cast<int>(stage2);
  if (announced.exchange(key, std::memory_order_relaxed) == key ||
      !reporting.load(std::memory_order_relaxed))
    return;
  std::fprintf(stderr, "%s: stage0 %s window, stage2 %s window\n", backend(),
               name(stage0), name(stage2));
}

void announce_fused() {
  if (announced.exchange(1, std::memory_order_relaxed) == 1 ||
      !reporting.load(std::memory_order_relaxed))
    return;
  std::fprintf(stderr, "%s: the three stages fused, one kernel\n", backend());
}

} // namespace internal
} // namespace gpu
