// gpu::find<T> called the way mxi_find --gpu -j N calls it
// (src/spots/find_spots.cc:396-417, 732-891): N worker threads, each
// thresholding whole frames on its own stream, dispatched in chunks of 4N+4;
// the main thread groups each finished chunk in ascending frame order with
// the Labeller.  Every frame's list is compared with the oracle's dext<T>
// on the worker that made it, and the grouped spots, pixels and counts with
// the oracle's Labeller fed the oracle's lists (variance within 1e-9,
// everything else exact).  Half the workers go through gpu::find, the other
// half through the C ABI's sf_find_run_as with the direct kernels, at the
// same time, so that a choice made for one call cannot steer another's.
//
// Then the failure path (dext_cuda.cu:46-51, 445-467): a failure forced
// from the host after a frame's work is enqueued must come out of find as
// std::runtime_error, leave no work in flight on the thread's stream, and
// leave the next frame's result intact.
//
//   test_find_threads [N]   exit 0 pass, 1 fail, 77 no device
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "dext_gpu.hh"
#include "oracle.h"
#include "spotfinder_hip.h"
#include "spotfinder_synth.h"

namespace {

std::atomic<int> failures{0};

void fail(const std::string &what) {
  std::printf("FAIL %s\n", what.c_str());
  std::fflush(stdout);
  failures++;
}

std::vector<oracle_signal_pixel> oracle_list(const std::vector<uint32_t> &f,
                                             size_t h, size_t w) {
  std::vector<oracle_signal_pixel> out(h * w / 4 + 1);
  size_t found = 0;
  if (oracle_dext_u32(f.data(), h, w, out.data(), out.size(), &found) != 0)
    fail("oracle refused a frame");
  out.resize(found);
  return out;
}

// One chunk in flight, as mxi_find's Chunk: the main thread posts frame
// numbers, the workers take them, and the main thread waits for all.
struct Chunk {
  std::mutex m;
  std::condition_variable cv;
  int64_t next = 0, end = 0;
  int64_t outstanding = 0;
  bool closing = false;
};

} // namespace

int main(int argc, char **argv) {
  if (!gpu::available()) {
    std::printf("SKIP: %s\n", gpu::unavailable_reason().c_str());
    return 77;
  }
  const int threads = argc > 1 ? std::atoi(argv[1]) : 8;
  const int64_t chunk_frames = 4 * threads + 4; // find_spots.cc:321-322
  const int64_t frames = 2 * chunk_frames + 5;   // a ragged last chunk
  gpu::report_windows(false);

  sf_synth_desc d;
  sf_synth_eiger(&d, 2, 4); // EIGER2 4M: 2068 x 2162 uint32
  d.persist = 1;
  d.seed = 17;
  d.spots_per_frame = 600;
  d.threads = 1; // each worker makes its own frames
  const size_t h = d.height, w = d.width;

  std::vector<std::vector<SignalPixel>> found(static_cast<size_t>(frames));
  std::vector<std::vector<oracle_signal_pixel>> expected(
      static_cast<size_t>(frames));

  Chunk chunk;
  std::vector<std::thread> workers;
  for (int t = 0; t < threads; t++) {
    workers.emplace_back([&, t] {
      std::vector<uint32_t> frame(h * w);
      for (;;) {
        int64_t z;
        {
          std::unique_lock<std::mutex> lock(chunk.m);
          chunk.cv.wait(lock, [&] { return chunk.closing || chunk.next < chunk.end; });
          if (chunk.next >= chunk.end)
            return;
          z = chunk.next++;
        }
        sf_synth_frame(&d, z, frame.data());
        std::vector<SignalPixel> &got = found[static_cast<size_t>(z)];
        int rc;
        if (t % 2 == 0) {
          rc = gpu::find<uint32_t>(frame.data(), got, h, w);
        } else {
          size_t n = 0;
          rc = sf_find_run_as(32, frame.data(), h, w, SF_KERNEL_DIRECT, 0, &n);
          got.resize(n);
          if (rc == 0 && sf_find_fetch(reinterpret_cast<sf_signal_pixel *>(got.data()), n) != 0)
            rc = -9;
        }
        expected[static_cast<size_t>(z)] = oracle_list(frame, h, w);
        const auto &want = expected[static_cast<size_t>(z)];
        if (rc != 0 || got.size() != want.size() ||
            std::memcmp(got.data(), want.data(), 16 * want.size()) != 0)
          fail("frame " + std::to_string(z) + " on thread " + std::to_string(t) +
               ": status " + std::to_string(rc) + ", " +
               std::to_string(got.size()) + " pixels against " +
               std::to_string(want.size()));
        if (sf_find_idle() != 1)
          fail("work in flight after find returned, thread " + std::to_string(t));
        std::lock_guard<std::mutex> lock(chunk.m);
        if (--chunk.outstanding == 0)
          chunk.cv.notify_all();
      }
    });
  }

  sf_grouping g;
  sf_grouping_default(&g);
  sf_labeller *lab = nullptr;
  if (sf_labeller_create(h, w, &g, &lab) != 0) {
    std::printf("FAIL sf_labeller_create: %s\n", sf_last_error());
    return 1;
  }
  char err[256] = {0};
  oracle_labeller *ref = oracle_labeller_new(h, w, g.min_spot_size, g.max_spot_size,
                                             g.max_separation, 0, err, sizeof err);

  for (int64_t lo = 0; lo < frames; lo += chunk_frames) {
    const int64_t hi = lo + chunk_frames < frames ? lo + chunk_frames : frames;
    {
      std::unique_lock<std::mutex> lock(chunk.m);
      chunk.next = lo;
      chunk.end = hi;
      chunk.outstanding = hi - lo;
      chunk.cv.notify_all();
      chunk.cv.wait(lock, [&] { return chunk.outstanding == 0; });
    }
    std::printf("chunk %lld..%lld thresholded by %d threads\n",
                static_cast<long long>(lo), static_cast<long long>(hi), threads);
    std::fflush(stdout);
    // Grouped in ascending frame order, whatever order the workers finished.
    for (int64_t z = lo; z < hi; z++) {
      const auto &got = found[static_cast<size_t>(z)];
      if (sf_labeller_add(lab, z, reinterpret_cast<const sf_signal_pixel *>(got.data()),
                          got.size()) != 0)
        fail(std::string("sf_labeller_add: ") + sf_last_error());
      const auto &want = expected[static_cast<size_t>(z)];
      if (oracle_labeller_add(ref, z, want.data(), want.size(), err, sizeof err) != 0)
        fail(std::string("oracle_labeller_add: ") + err);
    }
  }
  {
    std::lock_guard<std::mutex> lock(chunk.m);
    chunk.closing = true;
    chunk.cv.notify_all();
  }
  for (auto &t : workers)
    t.join();

  sf_labeller_finish(lab);
  oracle_labeller_finish(ref);
  const size_t ns = sf_labeller_spot_count(lab);
  if (ns != oracle_labeller_spot_count(ref) || ns < 1000) {
    fail("spot count " + std::to_string(ns) + " against " +
         std::to_string(oracle_labeller_spot_count(ref)));
  } else {
    const sf_spot *a = sf_labeller_spots(lab);
    const oracle_spot *b = oracle_labeller_spots(ref);
    for (size_t i = 0; i < ns; i++) {
      bool same = std::memcmp(a[i].bbox, b[i].bbox, sizeof a[i].bbox) == 0 &&
                  std::memcmp(a[i].position, b[i].position, sizeof a[i].position) == 0 &&
                  a[i].intensity == b[i].intensity &&
                  a[i].intensity_variance == b[i].intensity_variance &&
                  a[i].n_signal == b[i].n_signal && a[i].first == b[i].first;
      for (int k = 0; k < 3; k++)
        same = same && std::fabs(a[i].variance[k] - b[i].variance[k]) <= 1e-9;
      if (!same) {
        fail("spot " + std::to_string(i));
        break;
      }
    }
  }
  const size_t np = sf_labeller_pixel_count(lab);
  if (np != oracle_labeller_pixel_count(ref) ||
      std::memcmp(sf_labeller_pixels(lab), oracle_labeller_pixels(ref), 12 * np) != 0)
    fail("spot pixels differ");
  const sf_counts c = sf_labeller_counts(lab);
  const oracle_counts oc = oracle_labeller_counts(ref);
  if (std::memcmp(&c, &oc, sizeof c) != 0)
    fail("counts differ");
  std::printf("%lld frames on %d threads in chunks of %lld: %zu spots\n",
              static_cast<long long>(frames), threads,
              static_cast<long long>(chunk_frames), ns);
  sf_labeller_destroy(lab);
  oracle_labeller_free(ref);

  // The failure path: the second find from now fails after its work is
  // enqueued; it throws, nothing is left in flight, and the frame after it
  // (the same frame again) is thresholded correctly.
  std::vector<uint32_t> frame(h * w);
  sf_synth_frame(&d, 3, frame.data());
  const auto want = oracle_list(frame, h, w);
  std::vector<SignalPixel> got;
  sf_inject_find_failure(2);
  bool threw = false;
  for (int call = 0; call < 3; call++) {
    try {
      const int rc = gpu::find<uint32_t>(frame.data(), got, h, w);
      if (rc != 0 || got.size() != want.size() ||
          std::memcmp(got.data(), want.data(), 16 * want.size()) != 0)
        fail("find around the injected failure, call " + std::to_string(call));
    } catch (const std::runtime_error &e) {
      threw = true;
      if (call != 1 || std::strstr(e.what(), "injected") == nullptr)
        fail(std::string("unexpected failure: ") + e.what());
      if (!got.empty())
        fail("a failed find left pixels in the output");
    }
    if (sf_find_idle() != 1)
      fail("work in flight after find, call " + std::to_string(call));
  }
  if (!threw)
    fail("the injected failure did not throw");
  sf_inject_find_failure(0);

  std::printf("%s: gpu::find from %d threads, %d failures\n",
              failures == 0 ? "PASS" : "FAILED", threads, failures.load());
  return failures == 0 ? 0 : 1;
}
