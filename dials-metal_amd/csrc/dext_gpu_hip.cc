// gpu:: for HIP: the reference's device backend interface
// (src/spots/dext_gpu.hh, dext_gpu.cc) over this library's C ABI.
//
// The policy part -- which kernel, the environment overrides, the warnings,
// the announcement on stderr when the configuration changes -- follows
// dext_gpu.cc:50-166; the device part is sf_find_run / sf_find_fetch, whose
// per-thread workspace and stream mirror dext_cuda.cu:471-545.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <type_traits>

#include "dext_gpu.hh"

static_assert(sizeof(SignalPixel) == sizeof(sf_signal_pixel),
              "SignalPixel and sf_signal_pixel must agree");

namespace gpu {
namespace {

std::atomic<Window> chosen0{Window::Tile};
std::atomic<Window> chosen2{Window::Tile};
std::atomic<bool> chosen_fused{true};
std::atomic<bool> reporting{true};
std::atomic<int> announced{-1};
std::once_flag chosen_once;

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

// The kernel the current choice means, and the line announcing it.
int kernel_for(Window w0, Window w2, bool f) {
  if (std::getenv("SPOTFINDER_HIP_KERNEL"))
    return sf_get_kernel();
  return (!f && w0 == Window::Direct && w2 == Window::Direct) ? SF_KERNEL_DIRECT
                                                              : SF_KERNEL_FUSED;
}

void announce(int kernel) {
  if (!reporting.load())
    return;
  if (announced.exchange(kernel) == kernel)
    return;
  std::fprintf(stderr, "%s: %s\n", backend(),
               kernel == SF_KERNEL_FUSED
                   ? "the three stages fused, one sweep kernel"
                   : "stage0, stage1 and stage2 direct-window kernels");
}

} // namespace

const char *backend() { return sf_backend(); }

const char *name(Window window) {
  return window == Window::Tile ? "tile" : "direct";
}

Window stage0_window() {
  std::call_once(chosen_once, resolve);
  return chosen0.load();
}
void stage0_window(Window window) {
  std::call_once(chosen_once, resolve);
  chosen0.store(window);
}
Window stage2_window() {
  std::call_once(chosen_once, resolve);
  return chosen2.load();
}
void stage2_window(Window window) {
  std::call_once(chosen_once, resolve);
  chosen2.store(window);
}
bool fused() {
  std::call_once(chosen_once, resolve);
  return chosen_fused.load();
}
void fused(bool on) {
  std::call_once(chosen_once, resolve);
  chosen_fused.store(on);
}

void profile_stages(bool on) { sf_profile_stages(on ? 1 : 0); }
bool profile_stages() { return sf_profiling_stages() != 0; }

StageTimes last_stage_times() {
  const sf_stage_times t = sf_last_stage_times();
  StageTimes s;
  s.upload = t.upload;
  s.stage0 = t.stage0;
  s.stage1 = t.stage1;
  s.stage2 = t.stage2;
  s.fused = t.fused;
  s.copy = t.copy;
  s.sort = t.sort;
  return s;
}

void report_windows(bool on) { reporting.store(on); }

bool available() { return sf_available() != 0; }
std::string unavailable_reason() { return sf_unavailable_reason(); }
std::size_t memory_free() { return sf_memory_free(); }
void *host_alloc(std::size_t bytes) { return sf_host_alloc(bytes); }
void host_free(void *pointer) { sf_host_free(pointer); }

template <typename T>
int find(const T *image_in, std::vector<SignalPixel> &signal_out,
         std::size_t height, std::size_t width) {
  static_assert(std::is_same<T, std::uint16_t>::value ||
                    std::is_same<T, std::uint32_t>::value,
                "uint16_t or uint32_t frames");
  signal_out.clear();
  const int kernel = kernel_for(stage0_window(), stage2_window(), fused());
  sf_set_kernel(kernel);
  announce(kernel);
  std::size_t found = 0;
  const int bits = static_cast<int>(8 * sizeof(T));
  const int rc = sf_find_run(bits, image_in, height, width, &found);
  if (rc == -1 || rc == -2)
    return rc;
  if (rc != 0)
    throw std::runtime_error(sf_last_error());
  signal_out.resize(found);
  const int fetched = sf_find_fetch(
      reinterpret_cast<sf_signal_pixel *>(signal_out.data()), found);
  if (fetched != 0)
    throw std::runtime_error(sf_last_error());
  return 0;
}

template int find<std::uint16_t>(const std::uint16_t *,
                                 std::vector<SignalPixel> &, std::size_t,
                                 std::size_t);
template int find<std::uint32_t>(const std::uint32_t *,
                                 std::vector<SignalPixel> &, std::size_t,
                                 std::size_t);

namespace internal {
Window default_stage0_window() { return Window::Tile; }
Window default_stage2_window() { return Window::Tile; }
} // namespace internal

} // namespace gpu
