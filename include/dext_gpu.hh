// The reference's device backend interface, namespace gpu, implemented for
// HIP on gfx950 (dials-metal_amd/csrc/dext_gpu_hip.cc).  Declaration for
// declaration the interface of src/spots/dext_gpu.hh:25-182 and
// dext_gpu_internal.hh:39-40, so that src/spots/find_spots.cc's single call
// site (threshold<T>, find_spots.cc:396-417) and the reference's device test
// and benchmark build against it unchanged.
//
// Where the HIP backend differs in substance it says so here:
//   * backend() is "HIP".
//   * The two windows choose between two implementations: the fused
//     single-pass kernel (any window set to tile, or fused() on; the default)
//     and the three direct-window kernels (both windows direct and fused()
//     off).  Both give the same answer, as the reference's four pairs do.
//     SPOTFINDER_GPU_STAGE0 / _STAGE2 / _FUSED are honoured as there, and
//     SPOTFINDER_HIP_KERNEL=fused|direct overrides both.
//   * uint32 frames are supported (the reference's CUDA instantiation was
//     never run, docs/outstanding.md:119-123; its Metal one refuses them).
#ifndef SPOTFINDER_DEXT_GPU_HH
#define SPOTFINDER_DEXT_GPU_HH

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "spotfinder_hip.h"

// SignalPixel, src/spots/signal_pixel.hh:20-28; the same layout as
// sf_signal_pixel.
#ifndef SPOTFINDER_SIGNAL_PIXEL_HH
#define SPOTFINDER_SIGNAL_PIXEL_HH
struct SignalPixel {
  std::uint32_t index;
  std::uint32_t value;
  float background;
  std::uint16_t population;
  std::uint16_t reserved;
};
static_assert(sizeof(SignalPixel) == 16, "SignalPixel must pack to 128 bits");
#endif

namespace gpu {

const char *backend();

enum class Window { Direct, Tile };
const char *name(Window window);

Window stage0_window();
void stage0_window(Window window);
Window stage2_window();
void stage2_window(Window window);
bool fused();
void fused(bool on);

struct StageTimes {
  double upload = 0.0;
  double stage0 = 0.0;
  double stage1 = 0.0;
  double stage2 = 0.0;
  double fused = 0.0;
  double copy = 0.0;
  double sort = 0.0;
};

void profile_stages(bool on);
bool profile_stages();
StageTimes last_stage_times();
void report_windows(bool on);

bool available();
std::string unavailable_reason();
std::size_t memory_free();
void *host_alloc(std::size_t bytes);
void host_free(void *pointer);

// As dext(): 0, -1 for a bad argument, -2 when the device list overflows;
// throws std::runtime_error on a device failure.  `signal_out` is cleared and
// refilled ascending by index, identical field for field to dext<T>.
template <typename T>
int find(const T *image_in, std::vector<SignalPixel> &signal_out,
         std::size_t height, std::size_t width);

extern template int find<std::uint16_t>(const std::uint16_t *,
                                        std::vector<SignalPixel> &, std::size_t,
                                        std::size_t);
extern template int find<std::uint32_t>(const std::uint32_t *,
                                        std::vector<SignalPixel> &, std::size_t,
                                        std::size_t);

namespace internal {
// The measured preferences on gfx950: tile (the fused kernel) for both.
Window default_stage0_window();
Window default_stage2_window();
} // namespace internal

} // namespace gpu

#endif
