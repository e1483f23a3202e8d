// The HIP backend of namespace gpu: what a device backend supplies to the
// reference's interface (src/spots/dext_gpu.hh:25-182), the part the CUDA
// backend defines in dext_cuda.cu:549-869 -- backend(), the measured window
// preference, availability, pinned host memory and find<T> -- over this
// library's C ABI.  The backend-independent part (windows, profiling flag,
// stage times, announcements) is the reference's dext_gpu.cc in the reference
// build and dext_gpu_policy.cc standalone; both reach this file through
// dext_gpu_internal.hh.
//
// The two windows choose between two implementations here: the fused
// single-pass kernel (either window Tile, or fused() on) and the three
// direct-window kernels (both Direct, fused() off).  Both give the same
// answer, as the reference's four window pairs do.  SPOTFINDER_HIP_KERNEL=
// fused|direct overrides the choice.  The per-thread workspace and stream
// behind find() mirror dext_cuda.cu:471-545 (sf_find_run in runtime.hip).
#include <cstdlib>
#include <stdexcept>
#include <type_traits>

#include "dext_gpu_internal.hh"
#include "spotfinder_hip.h"

static_assert(sizeof(SignalPixel) == sizeof(sf_signal_pixel),
              "SignalPixel and sf_signal_pixel must agree");

namespace gpu {

const char *backend() { return sf_backend(); }

namespace internal {
// Measured on MI355X: the fused kernel for both.
Window default_stage0_window() { return Window::Tile; }
Window default_stage2_window() { return Window::Tile; }
} // namespace internal

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
  const Window w0 = stage0_window(), w2 = stage2_window();
  int kernel = (!fused() && w0 == Window::Direct && w2 == Window::Direct)
                   ? SF_KERNEL_DIRECT
                   : SF_KERNEL_FUSED;
  if (std::getenv("SPOTFINDER_HIP_KERNEL"))
    kernel = sf_get_kernel();
  if (kernel == SF_KERNEL_FUSED)
    internal::announce_fused();
  else
    internal::announce_windows(Window::Direct, Window::Direct);
  const bool profiling = profile_stages();

  // The kernel and the timing go with this call only: other threads may be
  // inside find() under another choice.
  std::size_t found = 0;
  const int bits = static_cast<int>(8 * sizeof(T));
  const int rc = sf_find_run_as(bits, image_in, height, width, kernel,
                                profiling ? 1 : 0, &found);
  if (rc == -1 || rc == -2)
    return rc;
  if (rc != 0)
    throw std::runtime_error(sf_last_error());
  signal_out.resize(found);
  const int fetched = sf_find_fetch(
      reinterpret_cast<sf_signal_pixel *>(signal_out.data()), found);
  if (fetched != 0)
    throw std::runtime_error(sf_last_error());
  if (profiling) {
    const sf_stage_times t = sf_last_stage_times();
    StageTimes s;
    s.upload = t.upload;
    s.stage0 = t.stage0;
    s.stage1 = t.stage1;
    s.stage2 = t.stage2;
    s.fused = t.fused;
    s.copy = t.copy;
    s.sort = t.sort;
    internal::set_stage_times(s);
  }
  return 0;
}

template int find<std::uint16_t>(const std::uint16_t *,
                                 std::vector<SignalPixel> &, std::size_t,
                                 std::size_t);
template int find<std::uint32_t>(const std::uint32_t *,
                                 std::vector<SignalPixel> &, std::size_t,
                                 std::size_t);

} // namespace gpu
