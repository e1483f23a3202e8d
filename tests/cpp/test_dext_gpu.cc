// gpu::find<T> through the C++ drop-in interface (include/dext_gpu.hh)
// against the oracle's dext<T>, every field of every SignalPixel, in the
// shape of the reference's tests/spots/test_dext_gpu.cc:129-212: a set of
// frames run under every kernel choice, exit status 77 when no device is
// usable (ctest's skip), 1 on any difference.
#include <cstdio>
#include <cstring>
#include <vector>

#include "dext_gpu.hh"
#include "oracle.h"
#include "spotfinder_synth.h"

namespace {

int failures = 0;

template <typename T>
void compare(const char *what, const std::vector<T> &frame, size_t h,
             size_t w) {
  std::vector<SignalPixel> got;
  const int rc = gpu::find<T>(frame.data(), got, h, w);
  // The device list holds one entry per 16 bytes of frame
  // (dext_cuda.cu:736-741); past that both sides report -2.
  const size_t capacity = h * w * sizeof(T) / 16;
  std::vector<oracle_signal_pixel> want(h * w);
  size_t found = 0;
  int orc;
  if (sizeof(T) == 2)
    orc = oracle_dext_u16(reinterpret_cast<const uint16_t *>(frame.data()), h,
                          w, want.data(), want.size(), &found);
  else
    orc = oracle_dext_u32(reinterpret_cast<const uint32_t *>(frame.data()), h,
                          w, want.data(), want.size(), &found);
  if (orc == 0 && found > capacity) {
    orc = -2;
    found = 0;
  }
  if (rc != orc || got.size() != found ||
      (found && std::memcmp(got.data(), want.data(), found * 16) != 0)) {
    std::printf("FAIL %s (%zu x %zu): device %d with %zu pixels, oracle %d "
                "with %zu\n",
                what, h, w, rc, got.size(), orc, found);
    failures++;
  }
}

template <typename T>
std::vector<T> synth(sf_synth_desc d, int64_t frame) {
  std::vector<T> f(static_cast<size_t>(d.width) * d.height);
  sf_synth_frame(&d, frame, f.data());
  return f;
}

void run_all() {
  sf_synth_desc d;
  sf_synth_eiger(&d, 2, 4);
  compare("4M u32", synth<uint32_t>(d, 0), d.height, d.width);
  d.seed = 5;
  d.spots_per_frame = 3000;
  compare("4M u32 dense", synth<uint32_t>(d, 1), d.height, d.width);
  sf_synth_default(&d, 1030, 512);
  d.bits = 16;
  d.ring_radius = 200;
  compare("512 x 1030 u16", synth<uint16_t>(d, 0), d.height, d.width);
  sf_synth_default(&d, 96, 64);
  d.spots_per_frame = 3;
  compare("64 x 96 u32", synth<uint32_t>(d, 0), d.height, d.width);
  // Too dense for the device list on a small frame: -2 from both.
  sf_synth_default(&d, 27, 19);
  d.spots_per_frame = 3;
  compare("19 x 27 u32 over capacity", synth<uint32_t>(d, 0), d.height,
          d.width);
  std::vector<uint32_t> masked(64 * 64, 0xFFFFFFFEu);
  compare("64 x 64 all masked", masked, 64, 64);
}

} // namespace

int main() {
  if (!gpu::available()) {
    std::printf("SKIP: %s\n", gpu::unavailable_reason().c_str());
    return 77;
  }
  gpu::report_windows(true);
  std::printf("backend %s\n", gpu::backend());
  const gpu::Window windows[2] = {gpu::Window::Direct, gpu::Window::Tile};
  for (gpu::Window w0 : windows)
    for (gpu::Window w2 : windows) {
      gpu::stage0_window(w0);
      gpu::stage2_window(w2);
      gpu::fused(false);
      run_all();
    }
  gpu::fused(true);
  run_all();
  // A bad argument is -1, as dext()'s.
  std::vector<SignalPixel> out;
  const uint32_t one = 0;
  if (gpu::find<uint32_t>(nullptr, out, 4, 4) != -1 ||
      gpu::find<uint32_t>(&one, out, 0, 4) != -1) {
    std::printf("FAIL: bad arguments not refused\n");
    failures++;
  }
  std::printf("%s: gpu::find against dext, %d failures\n",
              failures == 0 ? "PASS" : "FAILED", failures);
  return failures == 0 ? 0 : 1;
}
