// The runtime behind the C ABI: device discovery, pinned host memory, device
// memory, tuning, the per-thread workspace gpu::find uses, device-resident
// workspaces and batches, and the grouping's glue.
//
// The contract follows the reference's CUDA backend (dext_cuda.cu:445-869):
// one stream and grow-only buffers per worker thread, the stream idle on
// every exit path, -1 for bad arguments, -2 when the list would overflow,
// and device failures reported with the runtime's own message.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hh"
#include "host_copy.hh"
#include "stitch.hh"
#include "sf_internal.hh"
#include "sf_labeller.hh"
#include "spotfinder_hip.h"

namespace sfi {

namespace {
thread_local std::string t_last_error;
thread_local std::string t_reason;
thread_local sf_stage_times t_times{};
std::atomic<int> g_kernel{-1};
std::atomic<bool> g_profile{false};
std::once_flag g_kernel_once;

void resolve_kernel() {
  int chosen = SF_KERNEL_FUSED;
  if (const char *asked = std::getenv("SPOTFINDER_HIP_KERNEL")) {
    if (std::strcmp(asked, "fused") == 0)
      chosen = SF_KERNEL_FUSED;
    else if (std::strcmp(asked, "direct") == 0)
      chosen = SF_KERNEL_DIRECT;
    else
      std::fprintf(stderr,
                   "warning: SPOTFINDER_HIP_KERNEL=%s is not 'fused' or "
                   "'direct'; leaving it alone\n",
                   asked);
  }
  int expected = -1;
  g_kernel.compare_exchange_strong(expected, chosen);
}

std::mutex g_tuning_mutex;
bool g_tuning_read = false;
// Launch shapes measured for batches of one size: by the batch's strip-rows,
// two buckets an octave (sf_autotune_mix).
std::map<int, sf_tuning> g_mix_shapes;
// Raised while an autotuner measures shapes: every batch then launches with
// the shape being measured, whatever sf_autotune_mix kept.
std::atomic<int> g_mix_bypass{0};
struct MixBypass {
  MixBypass() { g_mix_bypass++; }
  ~MixBypass() { g_mix_bypass--; }
};

int mix_bucket(uint64_t mix_rows) {
  return static_cast<int>(std::lround(2.0 * std::log2(static_cast<double>(
                                                std::max<uint64_t>(mix_rows, 1)))));
}
// Measured on MI355X over 16-frame batches of 16M frames: four waves a block
// (the most resident waves under the LDS each wave's ring takes), segments
// of 128-256 rows; a lone frame gets 40-row segments.
sf_tuning g_tuning = {0, 8192, 4, 32, 4};

bool tuning_valid(const sf_tuning &t) {
  return t.seg_rows >= 0 && t.seg_rows <= (1 << 20) && t.target_waves >= 1 &&
         (t.waves_per_block == 1 || t.waves_per_block == 2 ||
          t.waves_per_block == 4) &&
         t.frames_per_launch >= 1 && t.frames_per_launch <= 4096 &&
         (t.lane_columns == 4 || t.lane_columns == 8);
}

// SPOTFINDER_HIP_TUNE="key=value,...": an unknown key or a bad value warns
// and leaves that setting alone, as the reference's environment does.
void read_tuning_env() {
  const char *env = std::getenv("SPOTFINDER_HIP_TUNE");
  if (!env)
    return;
  const std::string s(env);
  size_t at = 0;
  while (at < s.size()) {
    size_t end = s.find(',', at);
    if (end == std::string::npos)
      end = s.size();
    const std::string item = s.substr(at, end - at);
    at = end + 1;
    const size_t eq = item.find('=');
    sf_tuning t = g_tuning;
    bool known = eq != std::string::npos;
    if (known) {
      const std::string key = item.substr(0, eq);
      const int value = std::atoi(item.c_str() + eq + 1);
      if (key == "seg_rows")
        t.seg_rows = value;
      else if (key == "target_waves")
        t.target_waves = value;
      else if (key == "waves_per_block")
        t.waves_per_block = value;
      else if (key == "frames_per_launch")
        t.frames_per_launch = value;
      else if (key == "lane_columns")
        t.lane_columns = value;
      else
        known = false;
    }
    if (known && tuning_valid(t))
      g_tuning = t;
    else
      std::fprintf(stderr,
                   "warning: SPOTFINDER_HIP_TUNE item '%s' is not understood; "
                   "leaving it alone\n",
                   item.c_str());
  }
}
} // namespace

void set_last_error(const std::string &message) { t_last_error = message; }

Tuning tuning() {
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  if (!g_tuning_read) {
    g_tuning_read = true;
    read_tuning_env();
  }
  return g_tuning;
}

Tuning tuning_for_mix(uint64_t mix_rows) {
  const Tuning global = tuning();
  if (g_mix_bypass.load() != 0)
    return global;
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  const auto it = g_mix_shapes.find(mix_bucket(mix_rows));
  if (it == g_mix_shapes.end())
    return global;
  Tuning t = it->second;
  t.frames_per_launch = global.frames_per_launch; // a batched-mode setting
  return t;
}

} // namespace sfi

using namespace sfi;

// A workspace: an engine whose two slots alternate frame by frame, so the
// previous frame is still there to link the next one to.
struct sf_workspace {
  explicit sf_workspace(void *stream) : engine(stream) {}
  ~sf_workspace() {
    (void)hipStreamSynchronize(engine.stream()); // nothing still reads `stage`
    for (hipEvent_t e : pulled)
      if (e)
        (void)hipEventDestroy(e);
  }
  Engine engine;
  int cur = 1;
  DeviceBuffer upload; // host frames for sf_find_run
  // Pageable host frames pass through these two pinned pieces in turn; each
  // event marks the end of the last kernel that read its piece.
  Pinned stage[2];
  hipEvent_t pulled[2] = {nullptr, nullptr};
  FrameSlot &slot() { return engine.slot(cur); }
  FrameSlot &other() { return engine.slot(cur ^ 1); }
  bool has_frame() { return engine.slot_count() > 0 && slot().thresholded; }
};

// A batch: an engine whose slots hold each batch in slots 0..n-1, but for
// the previous batch's last frame, whose slot is left alone (it links the
// next batch's first frame): the frame that would take it goes to slot n.
struct sf_batch {
  explicit sf_batch(void *stream) : engine(stream) {}
  Engine engine;
  std::vector<int> ids; // slots of the last batch, in frame order
  int last_prev_id = -1; // the previous batch's last slot
};

namespace {

// Leaves the stream idle however a call is left (dext_cuda.cu:445-467).
struct StreamGuard {
  hipStream_t s;
  ~StreamGuard() {
    if (s)
      (void)hipStreamSynchronize(s);
  }
};

template <typename F> int guarded(F &&body) {
  try {
    return body();
  } catch (const device_error &e) {
    set_last_error(e.what());
    return -3;
  } catch (const std::bad_alloc &) {
    set_last_error("out of host memory");
    return -4;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -3;
  }
}

int bad(const char *what) {
  set_last_error(std::string("bad argument: ") + what);
  return -1;
}

bool bad_shape(size_t height, size_t width) {
  return height == 0 || width == 0 || height > 0x7fffffffu / width;
}

std::unique_ptr<sf_workspace> &thread_workspace() {
  static thread_local std::unique_ptr<sf_workspace> ws;
  return ws;
}

// The workspace's next slot takes the frame; it links to the one before.
// With `graph` the fused sequence replays from one capture per frame shape.
void ws_threshold(sf_workspace *ws, int bits, const void *frame, uint32_t h,
                  uint32_t w, int kernel, bool graph = true) {
  ws->cur ^= 1;
  const FrameIn f{frame, h, w, bits};
  ws->engine.threshold(&f, &ws->cur, 1, kernel,
                       graph && kernel == SF_KERNEL_FUSED);
}

int ws_sync(sf_workspace *ws, size_t *found) {
  size_t n = 0;
  int status = 0;
  ws->engine.sync(&ws->cur, 1, &n, &status);
  if (found)
    *found = n;
  return status;
}

bool ascending_in_frame(const sf_signal_pixel *s, size_t n, size_t pixels) {
  for (size_t i = 0; i < n; i++)
    if (s[i].index >= pixels || (i > 0 && s[i].index <= s[i - 1].index))
      return false;
  return true;
}

} // namespace

// The grouping, with the engine state needed to honour frame links: a frame's
// links count only if they were taken against the frame added just before it.
struct sf_labeller {
  explicit sf_labeller(size_t h, size_t w, const sfl::Options &o) : lab(h, w, o) {}
  sfl::Labeller lab;
  std::unique_ptr<sf_workspace> ws; // for host lists (sf_labeller_add)
  const Engine *last_engine = nullptr;
  uint64_t last_serial = 0;
  // A frame's lists come back into pinned memory, which the copy engines
  // write at the link's rate (pageable memory goes through a bounce buffer).
  Pinned list, comp_of, links, comps;

  // Adds engine slot `id` as `frame`; `host_list` when the caller has it.
  // `fresh` false reuses the counters read for an earlier slot of the same
  // labelled batch.
  void add_slot(int64_t frame, Engine &e, int id,
                const sf_signal_pixel *host_list, bool fresh = true) {
    if (ds) {
      add_device(frame, e, &id, 1);
      return;
    }
    size_t n = 0, ncomp = 0, nlinks = 0;
    e.label_counts(id, &n, &ncomp, &nlinks, fresh);
    sf_signal_pixel *fetched = host_list ? nullptr : list.get<sf_signal_pixel>(n);
    uint32_t *c_of = comp_of.get<uint32_t>(n);
    sf_component *cs = comps.get<sf_component>(ncomp);
    uint32_t *ln = links.get<uint32_t>(2 * nlinks);
    e.fetch_labels(id, c_of, n, cs, ncomp, ln, nlinks, fetched);
    const sf_signal_pixel *l = host_list ? host_list : fetched;
    const FrameSlot &s = e.slot(id);
    const bool to_previous = s.linked_serial != 0 && lab.started() &&
                             last_engine == &e && last_serial == s.linked_serial;
    lab.add(frame, l, n, c_of, cs, ncomp, ln, nlinks, to_previous);
    last_engine = &e;
    last_serial = s.serial;
  }

  // Device mode: the grouping itself runs on the device (DeviceStitch); slots
  // ids[0..count) become frames first_frame onward.  A frame's links count
  // only if they were taken against the frame added just before it.
  std::unique_ptr<DeviceStitch> ds;

  void check_order(int64_t frame) const {
    if (ds && ds->started() && frame <= ds->last_frame())
      throw std::runtime_error("dials_spots: frame " + std::to_string(frame) +
                               " arrived after frame " +
                               std::to_string(ds->last_frame()) +
                               "; frames must be added in ascending order");
  }

  void add_device(int64_t first_frame, Engine &e, const int *ids, size_t count) {
    check_order(first_frame);
    std::unique_ptr<bool[]> linked(new bool[count]);
    const Engine *engine = last_engine;
    uint64_t serial = last_serial;
    for (size_t i = 0; i < count; i++) {
      const FrameSlot &s = e.slot(ids[i]);
      linked[i] = s.linked_serial != 0 && (i > 0 || ds->started()) &&
                  engine == &e && serial == s.linked_serial;
      engine = &e;
      serial = s.serial;
    }
    ds->add(e, ids, count, first_frame, linked.get());
    last_engine = &e;
    last_serial = e.slot(ids[count - 1]).serial;
  }


};

extern "C" {

const char *sf_backend(void) { return "HIP"; }

int sf_available(void) {
  int devices = 0;
  return hipGetDeviceCount(&devices) == hipSuccess && devices > 0;
}

const char *sf_unavailable_reason(void) {
  int devices = 0;
  const hipError_t e = hipGetDeviceCount(&devices);
  if (e == hipSuccess && devices > 0) {
    t_reason.clear();
  } else if (e != hipSuccess) {
    t_reason = std::string(hipGetErrorName(e)) + ", " + hipGetErrorString(e);
  } else {
    t_reason = "HIP found no devices";
  }
  if (!t_reason.empty()) {
    if (const char *v = std::getenv("HIP_VISIBLE_DEVICES"))
      t_reason += std::string("; HIP_VISIBLE_DEVICES is \"") + v + "\"";
  }
  return t_reason.c_str();
}

size_t sf_memory_free(void) {
  size_t free_bytes = 0, total = 0;
  if (hipMemGetInfo(&free_bytes, &total) != hipSuccess)
    return 0;
  return free_bytes;
}

void *sf_host_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess)
    return nullptr;
  return p;
}

void sf_host_free(void *pointer) {
  if (pointer)
    (void)hipHostFree(pointer);
}

const char *sf_last_error(void) { return t_last_error.c_str(); }

void *sf_device_alloc(size_t bytes) {
  void *p = nullptr;
  const hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    set_last_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return nullptr;
  }
  return p;
}

void sf_device_free(void *pointer) {
  if (pointer)
    (void)hipFree(pointer);
}

int sf_copy_to_device(void *dst, const void *src, size_t bytes) {
  return guarded([&]() -> int {
    check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    return 0;
  });
}

int sf_copy_to_host(void *dst, const void *src, size_t bytes) {
  return guarded([&]() -> int {
    check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    return 0;
  });
}

int sf_copy_on_device(void *dst, const void *src, size_t bytes) {
  return guarded([&]() -> int {
    check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy");
    return 0;
  });
}

int sf_device_synchronize(void) {
  return guarded([&]() -> int {
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return 0;
  });
}

void sf_set_kernel(int kernel) {
  std::call_once(g_kernel_once, resolve_kernel);
  if (kernel == SF_KERNEL_FUSED || kernel == SF_KERNEL_DIRECT)
    g_kernel.store(kernel);
}

int sf_get_kernel(void) {
  std::call_once(g_kernel_once, resolve_kernel);
  return g_kernel.load();
}

void sf_get_tuning(sf_tuning *t) {
  if (t)
    *t = tuning();
}

int sf_set_tuning(const sf_tuning *t) {
  if (!t || !tuning_valid(*t))
    return bad("tuning out of range");
  (void)tuning(); // the environment is read once, before any explicit set
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  g_tuning = *t;
  return 0;
}

void sf_profile_stages(int on) { g_profile.store(on != 0); }
int sf_profiling_stages(void) { return g_profile.load() ? 1 : 0; }
sf_stage_times sf_last_stage_times(void) { return t_times; }

namespace {
// Whether this thread's last find asked for stage times (its fetch then
// times the copy back).
thread_local bool t_find_timing = false;
// sf_inject_find_failure: calls left before the one that fails (0: none).
std::atomic<uint64_t> g_fail_in{0};

bool injected_failure_now() {
  uint64_t left = g_fail_in.load();
  while (left != 0)
    if (g_fail_in.compare_exchange_weak(left, left - 1))
      return left == 1;
  return false;
}

// Copies a pageable host frame to the workspace's upload buffer: the host's
// copy threads fill one pinned piece while the kernel pulls the other, so the
// transfer runs at the slower of the host copy and the link, not at the
// runtime's pageable path (which stages through its own buffers).
void pull_pageable(sf_workspace *ws, const void *frame, size_t bytes) {
  constexpr size_t kPiece = size_t{16} << 20;
  hipStream_t s = ws->engine.stream();
  const char *src = static_cast<const char *>(frame);
  char *dst = static_cast<char *>(ws->upload.ptr);
  for (size_t off = 0, k = 0; off < bytes; off += kPiece, k++) {
    const size_t n = std::min(kPiece, bytes - off);
    char *piece = ws->stage[k & 1].get<char>(kPiece);
    hipEvent_t &done = ws->pulled[k & 1];
    if (done) // the kernel that last read this piece has finished
      check(hipEventSynchronize(done), "hipEventSynchronize");
    else
      check(hipEventCreateWithFlags(&done, hipEventDisableTiming),
            "hipEventCreate");
    const CopyPiece cp{piece, src + off, n};
    host_copy(&cp, 1);
    const void *mapped = mapped_host(piece);
    if (mapped == nullptr)
      throw device_error("pinned staging has no device mapping");
    launch_pull(mapped, dst + off, n, s);
    check(hipEventRecord(done, s), "hipEventRecord");
  }
}

int find_run(int pixel_bits, const void *frame, size_t height, size_t width,
             int kernel, bool timing, size_t *found) {
  if (found)
    *found = 0;
  if (frame == nullptr || bad_shape(height, width) ||
      (pixel_bits != 16 && pixel_bits != 32))
    return bad("null frame, zero size, more than INT32_MAX pixels, or not "
               "16 or 32 bits");
  if (kernel != SF_KERNEL_FUSED && kernel != SF_KERNEL_DIRECT)
    return bad("kernel is neither SF_KERNEL_FUSED nor SF_KERNEL_DIRECT");
  return guarded([&]() -> int {
    auto &ws = thread_workspace();
    if (!ws)
      ws = std::make_unique<sf_workspace>(nullptr);
    const StreamGuard idle{ws->engine.stream()};
    t_find_timing = timing;
    const size_t bytes = height * width * static_cast<size_t>(pixel_bits / 8);
    ws->upload.ensure(bytes, ws->engine.stream());
    hipEvent_t up[2] = {nullptr, nullptr};
    if (timing) {
      for (hipEvent_t &e : up)
        check(hipEventCreate(&e), "hipEventCreate");
      check(hipEventRecord(up[0], ws->engine.stream()), "hipEventRecord");
    }
    // A pinned frame is pulled by a kernel reading it over the link; the
    // runtime's DMA copy of one ran at half the link's rate on this stream
    // once other streams had come and gone in the process (2.5 ms rather
    // than 1.28 ms for 72 MB), which the kernel does not.
    const void *mapped = mapped_host(frame);
    if (mapped && reinterpret_cast<uintptr_t>(mapped) % 16 == 0)
      launch_pull(mapped, ws->upload.ptr, bytes, ws->engine.stream());
    else if (mapped)
      check(hipMemcpyAsync(ws->upload.ptr, frame, bytes, hipMemcpyHostToDevice,
                           ws->engine.stream()),
            "copying a frame to the device");
    else
      pull_pageable(ws.get(), frame, bytes);
    if (timing)
      check(hipEventRecord(up[1], ws->engine.stream()), "hipEventRecord");
    ws->engine.set_timing(timing);
    // No graph capture here: find runs on many threads at once, and another
    // thread's legacy-stream work during a capture would invalidate it.  A
    // host-fed frame's upload outweighs the launches a replay would save.
    ws_threshold(ws.get(), pixel_bits, ws->upload.ptr,
                 static_cast<uint32_t>(height), static_cast<uint32_t>(width),
                 kernel, false);
    // A failure forced from the host, with the frame's work still in
    // flight: the guard above must leave the stream idle, as after a real
    // device failure.
    if (injected_failure_now())
      throw device_error("injected failure (sf_inject_find_failure)");
    const int rc = ws_sync(ws.get(), found);
    if (timing) {
      sf_stage_times t{};
      float ms = 0;
      check(hipEventElapsedTime(&ms, up[0], up[1]), "hipEventElapsedTime");
      t.upload = ms;
      ws->engine.kernel_ms(&t.fused, &t.sort);
      t_times = t;
      for (hipEvent_t &e : up)
        (void)hipEventDestroy(e);
    }
    return rc;
  });
}
} // namespace

int sf_find_run(int pixel_bits, const void *frame, size_t height, size_t width,
                size_t *found) {
  return find_run(pixel_bits, frame, height, width, sf_get_kernel(),
                  g_profile.load(), found);
}

int sf_find_run_as(int pixel_bits, const void *frame, size_t height,
                   size_t width, int kernel, int timing, size_t *found) {
  return find_run(pixel_bits, frame, height, width, kernel, timing != 0, found);
}

void sf_inject_find_failure(uint64_t nth) { g_fail_in.store(nth); }

int sf_find_idle(void) {
  auto &ws = thread_workspace();
  if (!ws)
    return 1;
  return hipStreamQuery(ws->engine.stream()) == hipSuccess ? 1 : 0;
}

int sf_find_fetch(sf_signal_pixel *out, size_t n) {
  auto &ws = thread_workspace();
  if (!ws || !ws->has_frame())
    return bad("no frame has been thresholded on this thread");
  if (n > ws->slot().capacity || (n > 0 && out == nullptr))
    return bad("more pixels than the frame's list, or a null list");
  return guarded([&]() -> int {
    const auto t0 = std::chrono::steady_clock::now();
    ws->engine.fetch(ws->cur, out, n);
    if (t_find_timing)
      t_times.copy = std::chrono::duration<double, std::milli>(
                         std::chrono::steady_clock::now() - t0)
                         .count();
    return 0;
  });
}

// ---- device-resident workspaces -------------------------------------------

int sf_ws_create(size_t max_pixels, void *stream, sf_workspace **out) {
  (void)max_pixels; // buffers grow on first use to the frame in hand
  if (out == nullptr)
    return bad("null result");
  *out = nullptr;
  return guarded([&]() -> int {
    *out = new sf_workspace(stream);
    return 0;
  });
}

void sf_ws_destroy(sf_workspace *ws) { delete ws; }

void *sf_ws_stream(sf_workspace *ws) {
  return ws ? ws->engine.stream() : nullptr;
}

int sf_ws_threshold(sf_workspace *ws, int pixel_bits, const void *dev_frame,
                    uint32_t height, uint32_t width, int kernel) {
  if (ws == nullptr || dev_frame == nullptr || bad_shape(height, width) ||
      (pixel_bits != 16 && pixel_bits != 32) ||
      (kernel != SF_KERNEL_FUSED && kernel != SF_KERNEL_DIRECT))
    return bad("workspace, frame, shape, bits or kernel");
  return guarded([&]() -> int {
    ws_threshold(ws, pixel_bits, dev_frame, height, width, kernel);
    return 0;
  });
}

int sf_ws_sync(sf_workspace *ws, size_t *found) {
  if (ws == nullptr || !ws->has_frame())
    return bad("no frame has been thresholded");
  return guarded([&]() -> int { return ws_sync(ws, found); });
}

int sf_ws_fetch(sf_workspace *ws, sf_signal_pixel *out, size_t n) {
  if (ws == nullptr || !ws->has_frame() || n > ws->slot().capacity ||
      (n && !out))
    return bad("no frame, or more pixels than its list");
  return guarded([&]() -> int {
    ws->engine.fetch(ws->cur, out, n);
    return 0;
  });
}

void sf_ws_set_timing(sf_workspace *ws, int on) {
  if (ws)
    ws->engine.set_timing(on != 0);
}

int sf_ws_kernel_ms(sf_workspace *ws, double *threshold_ms, double *order_ms) {
  if (ws == nullptr || !ws->engine.timed())
    return bad("no timed frame");
  return guarded([&]() -> int {
    ws->engine.kernel_ms(threshold_ms, order_ms);
    return 0;
  });
}

uint64_t sf_ws_fallbacks(sf_workspace *ws) {
  return ws ? ws->engine.fallbacks() : 0;
}

int sf_ws_label(sf_workspace *ws, int link_previous) {
  if (ws == nullptr || !ws->has_frame())
    return bad("no frame has been thresholded");
  return guarded([&]() -> int {
    FrameSlot &f = ws->slot();
    int prev = -1;
    if (link_previous && ws->engine.slot_count() > 1) {
      FrameSlot &g = ws->other();
      if (g.labelled && g.serial + 1 == f.serial && g.height == f.height &&
          g.width == f.width)
        prev = ws->cur ^ 1;
    }
    ws->engine.label(&ws->cur, &prev, 1);
    return 0;
  });
}

int sf_ws_sync_label(sf_workspace *ws, size_t *n, size_t *ncomp,
                     size_t *nlinks, int *linked) {
  if (ws == nullptr || !ws->has_frame() || !ws->slot().labelled)
    return bad("no labelled frame");
  return guarded([&]() -> int {
    ws->engine.label_counts(ws->cur, n, ncomp, nlinks);
    if (linked)
      *linked = ws->slot().linked_serial != 0 ? 1 : 0;
    return 0;
  });
}

int sf_ws_fetch_labels(sf_workspace *ws, uint32_t *comp_of, size_t n,
                       sf_component *comps, size_t ncomp, uint32_t *links,
                       size_t nlinks) {
  if (ws == nullptr || !ws->has_frame() || !ws->slot().labelled ||
      n > ws->slot().capacity || ncomp > ws->slot().capacity + 1 ||
      nlinks > ws->slot().link_capacity)
    return bad("no labelled frame, or counts beyond its lists");
  return guarded([&]() -> int {
    ws->engine.fetch_labels(ws->cur, comp_of, n, comps, ncomp, links, nlinks);
    return 0;
  });
}

int sf_ws_load_signal(sf_workspace *ws, const sf_signal_pixel *signal,
                      size_t n, uint32_t height, uint32_t width) {
  if (ws == nullptr || bad_shape(height, width) || (n && !signal) ||
      !ascending_in_frame(signal, n, static_cast<size_t>(height) * width))
    return bad("a list that is not ascending or runs off the frame");
  return guarded([&]() -> int {
    ws->cur ^= 1;
    ws->engine.load_signal(ws->cur, signal, n, height, width);
    return 0;
  });
}

const sf_signal_pixel *sf_ws_device_signal(sf_workspace *ws) {
  return ws && ws->has_frame() ? ws->slot().ordered.as<sf_signal_pixel>()
                               : nullptr;
}

const uint32_t *sf_ws_device_mask(sf_workspace *ws, uint32_t *words_per_row) {
  if (!ws || !ws->has_frame())
    return nullptr;
  if (words_per_row)
    *words_per_row = ws->slot().wpr;
  return ws->slot().mask.as<uint32_t>();
}

// ---- batches ---------------------------------------------------------------

int sf_batch_create(void *stream, sf_batch **out) {
  if (out == nullptr)
    return bad("null result");
  *out = nullptr;
  return guarded([&]() -> int {
    *out = new sf_batch(stream);
    return 0;
  });
}

void sf_batch_destroy(sf_batch *b) { delete b; }

size_t sf_batch_device_bytes(const sf_batch *b) {
  return b ? b->engine.device_bytes() : 0;
}

void *sf_batch_stream(sf_batch *b) { return b ? b->engine.stream() : nullptr; }

int sf_batch_threshold(sf_batch *b, const sf_frame_desc *frames, size_t n,
                       int use_graph) {
  if (b == nullptr || frames == nullptr || n == 0 || n > 4096)
    return bad("batch, frames or count");
  for (size_t i = 0; i < n; i++)
    if (frames[i].frame == nullptr ||
        bad_shape(frames[i].height, frames[i].width) ||
        (frames[i].bits != 16 && frames[i].bits != 32))
      return bad("a frame's pointer, shape or bits");
  return guarded([&]() -> int {
    // Slots 0..n-1, except that the previous batch's last frame keeps its
    // slot (this batch's first frame links to it): the frame that would
    // land there takes slot n.  A batch never needs more than n + 1 slots.
    const int prev_last = b->ids.empty() ? -1 : b->ids.back();
    b->ids.resize(n);
    std::vector<FrameIn> in(n);
    for (size_t i = 0; i < n; i++) {
      b->ids[i] = static_cast<int>(i) == prev_last ? static_cast<int>(n)
                                                   : static_cast<int>(i);
      in[i] = FrameIn{frames[i].frame, frames[i].height, frames[i].width,
                      frames[i].bits};
    }
    b->last_prev_id = prev_last;
    b->engine.threshold(in.data(), b->ids.data(), n, sf_get_kernel(),
                        use_graph != 0);
    return 0;
  });
}

int sf_batch_sync(sf_batch *b, size_t *counts, int *status) {
  if (b == nullptr || b->ids.empty())
    return bad("no batch has been thresholded");
  return guarded([&]() -> int {
    b->engine.sync(b->ids.data(), b->ids.size(), counts, status);
    return 0;
  });
}

int sf_batch_fetch(sf_batch *b, size_t i, sf_signal_pixel *out, size_t n) {
  if (b == nullptr || i >= b->ids.size() ||
      n > b->engine.slot(b->ids[i]).capacity || (n && !out))
    return bad("frame index, or more pixels than its list");
  return guarded([&]() -> int {
    b->engine.fetch(b->ids[i], out, n);
    return 0;
  });
}

int sf_batch_label(sf_batch *b, int link_previous) {
  if (b == nullptr || b->ids.empty())
    return bad("no batch has been thresholded");
  return guarded([&]() -> int {
    const size_t n = b->ids.size();
    std::vector<int> prev(n, -1);
    for (size_t i = 0; i < n; i++) {
      const int p = i > 0 ? b->ids[i - 1] : b->last_prev_id;
      if (!link_previous || p < 0)
        continue;
      const FrameSlot &f = b->engine.slot(b->ids[i]);
      const FrameSlot &g = b->engine.slot(p);
      if ((g.labelled || i > 0) && g.height == f.height && g.width == f.width)
        prev[i] = p;
    }
    b->engine.label(b->ids.data(), prev.data(), n);
    return 0;
  });
}

int sf_batch_label_counts(sf_batch *b, size_t i, size_t *n, size_t *ncomp,
                          size_t *nlinks) {
  if (b == nullptr || i >= b->ids.size() || !b->engine.slot(b->ids[i]).labelled)
    return bad("frame index, or a frame not labelled");
  return guarded([&]() -> int {
    b->engine.label_counts(b->ids[i], n, ncomp, nlinks);
    return 0;
  });
}

int sf_batch_fetch_labels(sf_batch *b, size_t i, uint32_t *comp_of, size_t n,
                          sf_component *comps, size_t ncomp, uint32_t *links,
                          size_t nlinks) {
  if (b == nullptr || i >= b->ids.size())
    return bad("frame index");
  const FrameSlot &s = b->engine.slot(b->ids[i]);
  if (!s.labelled || n > s.capacity || ncomp > s.capacity + 1 ||
      nlinks > s.link_capacity)
    return bad("a frame not labelled, or counts beyond its lists");
  return guarded([&]() -> int {
    b->engine.fetch_labels(b->ids[i], comp_of, n, comps, ncomp, links, nlinks);
    return 0;
  });
}

void sf_batch_set_timing(sf_batch *b, int on) {
  if (b)
    b->engine.set_timing(on != 0);
}

int sf_batch_kernel_ms(sf_batch *b, double *threshold_ms, double *order_ms) {
  if (b == nullptr || !b->engine.timed())
    return bad("no timed batch");
  return guarded([&]() -> int {
    b->engine.kernel_ms(threshold_ms, order_ms);
    return 0;
  });
}

uint64_t sf_batch_fallbacks(sf_batch *b) {
  return b ? b->engine.fallbacks() : 0;
}

// ---- the autotuner ---------------------------------------------------------

int sf_autotune(const sf_frame_desc *frames, size_t n, int label,
                sf_autotune_report *report) {
  if (frames == nullptr || n == 0 || report == nullptr)
    return bad("frames or report");
  *report = sf_autotune_report{};
  const sf_tuning start = tuning();
  const MixBypass measuring; // the shapes tried, not a kept one, launch
  sf_batch *b = nullptr;
  int rc = sf_batch_create(nullptr, &b);
  if (rc != 0)
    return rc;
  std::unique_ptr<sf_batch, void (*)(sf_batch *)> own(b, sf_batch_destroy);
  auto fail = [&](int code) {
    (void)sf_set_tuning(&start);
    return code;
  };
  const size_t shape_n = n < 128 ? n : 128;
  std::vector<size_t> counts(n);
  std::vector<int> status(n);

  // 1. The fused kernel's shape, by its own duration over the first frames
  //    (median of three).
  sf_tuning best = start;
  double best_ms = 0.0;
  const int lane_columns[] = {4, 8};
  const int waves_per_block[] = {1, 2, 4};
  const int seg_rows[] = {64, 128, 192, 256, 0, 0, 0};
  const int target_waves[] = {start.target_waves, start.target_waves,
                              start.target_waves, start.target_waves, 4096,
                              8192, 16384};
  sf_batch_set_timing(b, 1);
  std::vector<sf_tuning> shapes;
  for (int kq : lane_columns)
    for (int w : waves_per_block)
      for (size_t k = 0; k < sizeof(seg_rows) / sizeof(seg_rows[0]); k++) {
        sf_tuning t = start;
        t.lane_columns = kq;
        t.waves_per_block = w;
        t.seg_rows = seg_rows[k];
        t.target_waves = target_waves[k];
        shapes.push_back(t);
      }
  for (const sf_tuning &t : shapes) {
    if ((rc = sf_set_tuning(&t)) != 0)
      return fail(rc);
    double ms[3];
    for (double &m : ms) {
      if ((rc = sf_batch_threshold(b, frames, shape_n, 0)) != 0 ||
          (rc = sf_batch_sync(b, counts.data(), status.data())) != 0 ||
          (rc = sf_batch_kernel_ms(b, &m, nullptr)) != 0)
        return fail(rc);
    }
    std::sort(ms, ms + 3);
    report->shapes_tried++;
    if (best_ms == 0.0 || ms[1] < best_ms) {
      best_ms = ms[1];
      best = t;
    }
  }
  sf_batch_set_timing(b, 0);
  report->fused_ms_per_frame = best_ms / static_cast<double>(shape_n);

  // 2. Frames per launch, by the whole step (threshold, ordering and, with
  //    `label`, labelling) per frame, over the sizes that fit: a batch holds
  //    two sets of slots, so only sizes whose slots take at most 60% of the
  //    free device memory are tried.
  size_t slot = 0;
  for (size_t i = 0; i < n; i++) {
    const size_t s = Engine::slot_bytes(frames[i].height, frames[i].width,
                                        frames[i].bits);
    slot = s > slot ? s : slot;
  }
  const double room = 0.6 * static_cast<double>(sf_memory_free());
  if ((rc = sf_set_tuning(&best)) != 0)
    return fail(rc);
  double best_step = 0.0;
  int best_f = best.frames_per_launch;
  for (int f : {16, 32, 64, 128}) {
    if (static_cast<size_t>(f) > n ||
        2.0 * static_cast<double>(f) * static_cast<double>(slot) > room)
      continue;
    const size_t launches = (n < 128 ? n : 128) / static_cast<size_t>(f);
    double secs = 0.0;
    for (int rep = 0; rep < 2; rep++) { // the first pass sizes the slots
      const auto t0 = std::chrono::steady_clock::now();
      for (size_t l = 0; l < launches; l++) {
        if ((rc = sf_batch_threshold(b, frames + l * f, static_cast<size_t>(f),
                                     1)) != 0 ||
            (label && (rc = sf_batch_label(b, 1)) != 0) ||
            (rc = sf_batch_sync(b, counts.data(), status.data())) != 0)
          return fail(rc);
      }
      secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                 .count();
    }
    const double per = 1e3 * secs / static_cast<double>(launches * f);
    report->launch_sizes_tried++;
    if (best_step == 0.0 || per < best_step) {
      best_step = per;
      best_f = f;
    }
  }
  best.frames_per_launch = best_f;
  report->step_ms_per_frame = best_step;
  report->best = best;
  if ((rc = sf_set_tuning(&best)) != 0)
    return fail(rc);
  return 0;
}

namespace {
uint64_t mix_rows_of(const sf_frame_desc *frames, size_t n) {
  uint64_t rows = 0;
  for (size_t i = 0; i < n; i++)
    rows += static_cast<uint64_t>(frames[i].height) * strips_of(frames[i].width, 4);
  return rows;
}
} // namespace

int sf_autotune_mix(const sf_frame_desc *frames, size_t n, sf_autotune_report *report) {
  if (frames == nullptr || n == 0 || n > 4096 || report == nullptr)
    return bad("frames or report");
  *report = sf_autotune_report{};
  const sf_tuning start = tuning();
  const int bucket = mix_bucket(mix_rows_of(frames, n));
  const MixBypass measuring;
  auto restore = [&](int code) {
    (void)sf_set_tuning(&start);
    return code;
  };
  sf_batch *b = nullptr;
  int rc = sf_batch_create(nullptr, &b);
  if (rc != 0)
    return restore(rc);
  std::unique_ptr<sf_batch, void (*)(sf_batch *)> own(b, sf_batch_destroy);
  std::vector<size_t> counts(n);
  std::vector<int> status(n);
  sf_batch_set_timing(b, 1);
  sf_tuning best = start;
  double best_ms = 0.0;
  for (int kq : {4, 8})
    for (int w : {1, 2, 4})
      for (int k = 0; k < 7; k++) {
        const int seg_rows[] = {64, 128, 192, 256, 0, 0, 0};
        const int target[] = {start.target_waves, start.target_waves,
                              start.target_waves, start.target_waves, 4096, 8192, 16384};
        sf_tuning t = start;
        t.lane_columns = kq;
        t.waves_per_block = w;
        t.seg_rows = seg_rows[k];
        t.target_waves = target[k];
        if ((rc = sf_set_tuning(&t)) != 0)
          return restore(rc);
        double ms[3];
        for (double &m : ms)
          if ((rc = sf_batch_threshold(b, frames, n, 0)) != 0 ||
              (rc = sf_batch_sync(b, counts.data(), status.data())) != 0 ||
              (rc = sf_batch_kernel_ms(b, &m, nullptr)) != 0)
            return restore(rc);
        std::sort(ms, ms + 3);
        report->shapes_tried++;
        if (best_ms == 0.0 || ms[1] < best_ms) {
          best_ms = ms[1];
          best = t;
        }
      }
  (void)sf_set_tuning(&start);
  report->best = best;
  report->fused_ms_per_frame = best_ms / static_cast<double>(n);
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  g_mix_shapes[bucket] = best;
  return 0;
}

int sf_tuning_for_mix(const sf_frame_desc *frames, size_t n, sf_tuning *t) {
  if (frames == nullptr || n == 0 || t == nullptr)
    return bad("frames or tuning");
  const uint64_t rows = mix_rows_of(frames, n);
  *t = tuning_for_mix(rows);
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  return g_mix_shapes.count(mix_bucket(rows)) ? 1 : 0;
}

void sf_clear_mix_tunings(void) {
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  g_mix_shapes.clear();
}

// ---- grouping --------------------------------------------------------------

void sf_grouping_default(sf_grouping *g) {
  if (!g)
    return;
  g->min_spot_size = 3;
  g->max_spot_size = 1000;
  g->max_separation = 2.0;
  g->two_d = 0;
}

int sf_labeller_create(size_t height, size_t width, const sf_grouping *g,
                       sf_labeller **out) {
  if (!out)
    return bad("null result");
  *out = nullptr;
  sfl::Options o;
  if (g) {
    o.min_spot_size = g->min_spot_size;
    o.max_spot_size = g->max_spot_size;
    o.max_separation = g->max_separation;
    o.two_d = g->two_d != 0;
  }
  try {
    *out = new sf_labeller(height, width, o);
    return 0;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  }
}

int sf_labeller_create_device(size_t height, size_t width,
                              const sf_grouping *g, sf_labeller **out) {
  const int rc = sf_labeller_create(height, width, g, out);
  if (rc != 0)
    return rc;
  return guarded([&]() -> int {
    try {
      (*out)->ds = std::make_unique<DeviceStitch>(height, width, (*out)->lab.options());
      return 0;
    } catch (...) {
      delete *out;
      *out = nullptr;
      throw;
    }
  });
}

size_t sf_labeller_device_bytes(const sf_labeller *l) {
  return l && l->ds ? l->ds->held_bytes() : 0;
}

void sf_labeller_destroy(sf_labeller *l) { delete l; }

int sf_labeller_add(sf_labeller *l, int64_t frame,
                    const sf_signal_pixel *signal, size_t n) {
  if (!l || (n && !signal))
    return bad("labeller or list");
  try {
    l->lab.check_frame(frame, signal, n);
    l->check_order(frame);
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  }
  return guarded([&]() -> int {
    if (!l->ws)
      l->ws = std::make_unique<sf_workspace>(nullptr);
    sf_workspace *ws = l->ws.get();
    ws->cur ^= 1;
    ws->engine.load_signal(ws->cur, signal, n,
                           static_cast<uint32_t>(l->lab.height()),
                           static_cast<uint32_t>(l->lab.width()));
    FrameSlot &f = ws->slot();
    int prev = -1;
    if (ws->engine.slot_count() > 1) {
      FrameSlot &g = ws->other();
      if (g.labelled && g.serial + 1 == f.serial)
        prev = ws->cur ^ 1;
    }
    ws->engine.label(&ws->cur, &prev, 1);
    l->add_slot(frame, ws->engine, ws->cur, signal);
    return 0;
  });
}

namespace {
int add_checked(sf_labeller *l, const std::function<void()> &body) {
  try {
    body();
    return 0;
  } catch (const device_error &e) {
    set_last_error(e.what());
    return -3;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  }
}
} // namespace

int sf_labeller_add_ws(sf_labeller *l, int64_t frame, sf_workspace *ws) {
  if (!l || !ws || !ws->has_frame() || !ws->slot().labelled ||
      ws->slot().height != l->lab.height() || ws->slot().width != l->lab.width())
    return bad("the workspace holds no labelled frame of this labeller's shape");
  return add_checked(l, [&] { l->add_slot(frame, ws->engine, ws->cur, nullptr); });
}

int sf_labeller_add_batch(sf_labeller *l, int64_t first_frame, sf_batch *b) {
  if (!l || !b || b->ids.empty())
    return bad("labeller or batch");
  for (int id : b->ids) {
    const FrameSlot &s = b->engine.slot(id);
    if (!s.labelled || s.height != l->lab.height() || s.width != l->lab.width())
      return bad("the batch holds a frame not labelled or not of this "
                 "labeller's shape");
  }
  return add_checked(l, [&] {
    if (l->ds) {
      l->add_device(first_frame, b->engine, b->ids.data(), b->ids.size());
      return;
    }
    for (size_t i = 0; i < b->ids.size(); i++)
      l->add_slot(first_frame + static_cast<int64_t>(i), b->engine, b->ids[i],
                  nullptr, i == 0);
  });
}

int sf_labeller_add_labelled(sf_labeller *l, int64_t frame,
                             const sf_signal_pixel *signal, size_t n,
                             const uint32_t *comp_of,
                             const sf_component *comps, size_t ncomp,
                             const uint32_t *links, size_t nlinks,
                             int links_to_previous) {
  if (!l || (n && (!signal || !comp_of)) || (ncomp && !comps) ||
      (nlinks && !links))
    return bad("labeller, list, labels, components or links");
  return add_checked(l, [&] {
    if (l->ds) // its frames' components must be on the device
      throw std::runtime_error("dials_spots: this labeller takes device frames");
    l->lab.add(frame, signal, n, comp_of, comps, ncomp, links, nlinks,
               links_to_previous != 0);
    l->last_engine = nullptr;
  });
}

int sf_labeller_finish(sf_labeller *l) {
  if (!l)
    return bad("labeller");
  return add_checked(l, [&] {
    if (l->ds)
      l->ds->finish();
    else
      l->lab.finish();
  });
}

// In device mode the records come from the device grouping's host copies; a
// device failure then reads as no records (sf_last_error says).
extern "C++" {
namespace {
template <typename F> auto records_of(F &&f) -> decltype(f()) {
  try {
    return f();
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return {nullptr, 0};
  }
}
} // namespace
}

size_t sf_labeller_spot_count(const sf_labeller *l) {
  if (!l)
    return 0;
  if (!l->ds)
    return l->lab.spots().size();
  return records_of([&] { return l->ds->spots(); }).size;
}
const sf_spot *sf_labeller_spots(const sf_labeller *l) {
  if (!l)
    return nullptr;
  if (!l->ds)
    return l->lab.spots().data();
  return records_of([&] { return l->ds->spots(); }).data;
}
size_t sf_labeller_pixel_count(const sf_labeller *l) {
  if (!l)
    return 0;
  if (!l->ds)
    return l->lab.pixels().size();
  return records_of([&] { return l->ds->pixels(); }).size;
}
const sf_pixel *sf_labeller_pixels(const sf_labeller *l) {
  if (!l)
    return nullptr;
  if (!l->ds)
    return l->lab.pixels().data();
  return records_of([&] { return l->ds->pixels(); }).data;
}
sf_counts sf_labeller_counts(const sf_labeller *l) {
  if (!l)
    return sf_counts{};
  return l->ds ? l->ds->counts() : l->lab.counts();
}
size_t sf_labeller_live(const sf_labeller *l) {
  if (!l)
    return 0;
  return l->ds ? l->ds->live_pixels() : l->lab.live();
}

int sf_labeller_write_refl(const sf_labeller *l, const char *path,
                           const sf_refl_options *options) {
  if (l == nullptr)
    return bad("labeller");
  return sf_refl_write(path, sf_labeller_spots(l), sf_labeller_spot_count(l),
                       sf_labeller_pixels(l), sf_labeller_pixel_count(l),
                       l->lab.width(), options);
}

} // extern "C"
