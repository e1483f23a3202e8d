// The backend-independent part of gpu:: (window choice, fused switch,
// profiling, announcements), which needs no device.  The checks follow the
// interface of src/spots/dext_gpu.hh:50-134: the environment is read before
// the first choice is read or written and never again, so a setter wins over
// it; an unrecognised value leaves the default; a configuration is announced
// on stderr when it changes and only while reporting is on.
//
// Run as `test_policy <expect-stage0> <expect-stage2> <expect-fused>`, each
// "tile"/"direct" or "1"/"0", under whatever SPOTFINDER_GPU_* environment the
// caller sets; prints what it announces between markers so that the caller
// can compare stderr.  Exit 1 on any difference.
#include <cstdio>
#include <cstring>

#include "dext_gpu_internal.hh"

namespace {
int failures = 0;

void expect(bool ok, const char *what) {
  if (!ok) {
    std::printf("FAIL %s\n", what);
    failures++;
  }
}
} // namespace

int main(int argc, char **argv) {
  if (argc != 4) {
    std::printf("usage: test_policy stage0 stage2 fused\n");
    return 2;
  }
  expect(std::strcmp(gpu::name(gpu::stage0_window()), argv[1]) == 0,
         "stage0 window from the environment");
  expect(std::strcmp(gpu::name(gpu::stage2_window()), argv[2]) == 0,
         "stage2 window from the environment");
  expect(gpu::fused() == (std::strcmp(argv[3], "1") == 0),
         "fused switch from the environment");

  // Setters take effect at once and are not undone later.
  gpu::stage0_window(gpu::Window::Direct);
  gpu::stage2_window(gpu::Window::Tile);
  gpu::fused(true);
  expect(gpu::stage0_window() == gpu::Window::Direct, "stage0 setter");
  expect(gpu::stage2_window() == gpu::Window::Tile, "stage2 setter");
  expect(gpu::fused(), "fused setter");
  gpu::fused(false);
  expect(!gpu::fused(), "fused setter off");
  gpu::stage0_window(gpu::Window::Tile);
  expect(gpu::stage0_window() == gpu::Window::Tile &&
             gpu::stage2_window() == gpu::Window::Tile,
         "one window's setter leaves the other");

  gpu::profile_stages(true);
  expect(gpu::profile_stages(), "profiling on");
  gpu::profile_stages(false);
  expect(!gpu::profile_stages(), "profiling off");
  expect(!gpu::fused() && gpu::stage0_window() == gpu::Window::Tile,
         "profiling leaves the windows");

  gpu::StageTimes t;
  t.fused = 1.5;
  t.upload = 0.25;
  gpu::internal::set_stage_times(t);
  expect(gpu::last_stage_times().fused == 1.5 &&
             gpu::last_stage_times().upload == 0.25,
         "stage times kept for the thread");

  // Announcements: expected on stderr between the markers are exactly
  //   fused, direct/direct, fused  (the repeats and the silenced ones not).
  std::fflush(stdout);
  std::fprintf(stderr, "<<announce\n");
  gpu::report_windows(true);
  gpu::internal::announce_fused();
  gpu::internal::announce_fused(); // unchanged: nothing
  gpu::internal::announce_windows(gpu::Window::Direct, gpu::Window::Direct);
  gpu::internal::announce_windows(gpu::Window::Direct, gpu::Window::Direct);
  gpu::report_windows(false);
  gpu::internal::announce_windows(gpu::Window::Tile, gpu::Window::Direct);
  gpu::report_windows(true);
  gpu::internal::announce_fused();
  std::fprintf(stderr, "announce>>\n");

  std::printf("%s: gpu:: policy, %d failures\n",
              failures == 0 ? "PASS" : "FAILED", failures);
  return failures == 0 ? 0 : 1;
}
