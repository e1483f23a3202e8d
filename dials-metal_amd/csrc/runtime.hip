// The runtime behind the C ABI: device discovery, pinned host memory, the
// per-thread workspace gpu::find uses, and device-resident workspaces.
//
// The contract follows the reference's CUDA backend (dext_cuda.cu:445-869):
// one stream and grow-only buffers per worker thread, the stream idle on
// every exit path, -1 for bad arguments, -2 when the list would overflow,
// and device failures reported with the runtime's own message.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "sf_internal.hh"
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
} // namespace

void set_last_error(const std::string &message) { t_last_error = message; }

// A grow-only device allocation.  Growing waits for the stream first: work in
// flight may still be reading the old buffer (dext_cuda.cu:506-525).
struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
  ~DeviceBuffer() {
    if (ptr)
      (void)hipFree(ptr);
  }
  void ensure(size_t want, hipStream_t stream) {
    if (want <= bytes)
      return;
    if (ptr) {
      check(hipStreamSynchronize(stream), "waiting before a buffer grows");
      check(hipFree(ptr), "hipFree");
      ptr = nullptr;
      bytes = 0;
    }
    check(hipMalloc(&ptr, want), "hipMalloc");
    bytes = want;
  }
  template <typename P> P *as() const { return static_cast<P *>(ptr); }
};

} // namespace sfi

using namespace sfi;

struct sf_workspace {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  DeviceBuffer upload, scr0, scr1, mask, prefix, unordered, ordered, scan_tmp,
      table;
  uint32_t *counter = nullptr;      // device: [0] count, [1] overflow
  uint32_t *host_counter = nullptr; // pinned mirror of both
  hipEvent_t marks[6] = {};
  uint32_t height = 0, width = 0, capacity = 0;
  int bits = 0;
  bool ran = false;
  const void *frame_ptr = nullptr;
  int kernel_used = SF_KERNEL_FUSED;
  uint64_t refused_fast = 0; // frames the fused kernel handed back
  bool timing_on = false;    // events around every sf_ws_threshold
  bool timed = false;        // the last threshold recorded its events

  ~sf_workspace() {
    if (stream)
      (void)hipStreamSynchronize(stream);
    for (hipEvent_t &e : marks)
      if (e)
        (void)hipEventDestroy(e);
    if (counter)
      (void)hipFree(counter);
    if (host_counter)
      (void)hipHostFree(host_counter);
    if (own_stream && stream)
      (void)hipStreamDestroy(stream);
  }

  void init(void *borrowed_stream) {
    check(hipGetDevice(&device), "hipGetDevice");
    if (borrowed_stream) {
      stream = static_cast<hipStream_t>(borrowed_stream);
    } else {
      check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking),
            "hipStreamCreate");
      own_stream = true;
    }
    check(hipMalloc(&counter, 2 * sizeof(uint32_t)),
          "hipMalloc for the counter");
    check(hipHostMalloc(reinterpret_cast<void **>(&host_counter),
                        2 * sizeof(uint32_t), hipHostMallocDefault),
          "hipHostMalloc for the counter");
    // R(sqrt(R(2) * (R(m) - 1))) for m = 0..127: the right-hand side of the
    // stage-0 test (dext.cc:105-106), by the host's correctly rounded sqrt.
    float tf[128];
    double td[128];
    for (int m = 0; m < 128; m++) {
      // Below two pixels the reference makes no test (dext.cc:99); +inf
      // makes the comparison false without a branch.
      tf[m] = m >= 2 ? std::sqrt(2.0f * (static_cast<float>(m) - 1.0f))
                     : HUGE_VALF;
      td[m] = m >= 2 ? std::sqrt(2.0 * (static_cast<double>(m) - 1.0))
                     : HUGE_VAL;
    }
    table.ensure(sizeof(tf) + sizeof(td), stream);
    check(hipMemcpy(table.ptr, tf, sizeof(tf), hipMemcpyHostToDevice),
          "uploading the root table");
    check(hipMemcpy(table.as<char>() + sizeof(tf), td, sizeof(td),
                    hipMemcpyHostToDevice),
          "uploading the root table");
  }

  const void *table_for(int b) const {
    return b == 16 ? table.ptr
                   : static_cast<const void *>(table.as<char>() + 128 * sizeof(float));
  }

  hipEvent_t *events() {
    if (!marks[0])
      for (hipEvent_t &e : marks)
        check(hipEventCreate(&e), "hipEventCreate");
    return marks;
  }

  // Enqueue the threshold and the ordering of a device frame.
  void threshold(int b, const void *frame, uint32_t h, uint32_t w, int kernel,
                 bool timing) {
    const size_t n = static_cast<size_t>(h) * w;
    const uint32_t wpr = mask_words_per_row(w);
    const size_t words = static_cast<size_t>(h) * wpr;
    // The reference's bound: a list the size of the frame's bytes
    // (dext_cuda.cu:736-741), one entry per 4 pixels at 32 bits, 8 at 16.
    const size_t cap = n * (b / 8) / sizeof(sf_signal_pixel);
    mask.ensure(words * 4, stream);
    prefix.ensure(words * 4 + 4, stream);
    unordered.ensure((cap ? cap : 1) * sizeof(sf_signal_pixel), stream);
    ordered.ensure((cap ? cap : 1) * sizeof(sf_signal_pixel), stream);
    scan_tmp.ensure(order_scan_bytes(words), stream);
    if (kernel == SF_KERNEL_DIRECT) {
      scr0.ensure(n, stream);
      scr1.ensure(n, stream);
    }
    height = h;
    width = w;
    bits = b;
    capacity = static_cast<uint32_t>(cap);

    FrameArgs a;
    a.frame = frame;
    a.height = h;
    a.width = w;
    a.bits = b;
    a.out = unordered.as<sf_signal_pixel>();
    a.counter = counter;
    a.capacity = capacity;
    a.mask = mask.as<uint32_t>();
    a.wpr = wpr;
    a.table = table_for(b);
    a.overflow = counter + 1;
    frame_ptr = frame;
    kernel_used = kernel;

    timing = timing || timing_on;
    timed = timing;
    hipEvent_t *ev = timing ? events() : nullptr;
    launch_fill(counter, 2, 0u, stream);
    if (kernel == SF_KERNEL_FUSED)
      launch_fill(a.mask, words, 0u, stream);
    if (ev)
      check(hipEventRecord(ev[1], stream), "hipEventRecord");
    if (kernel == SF_KERNEL_DIRECT)
      launch_direct(a, scr0.as<uint8_t>(), scr1.as<uint8_t>(), stream);
    else
      launch_fused(a, stream);
    if (ev)
      check(hipEventRecord(ev[2], stream), "hipEventRecord");
    launch_order(a, unordered.as<sf_signal_pixel>(),
                 ordered.as<sf_signal_pixel>(), prefix.as<uint32_t>(),
                 scan_tmp.ptr, scan_tmp.bytes, stream);
    if (ev)
      check(hipEventRecord(ev[3], stream), "hipEventRecord");
    ran = true;
  }

  int sync(size_t *found) {
    check(hipMemcpyAsync(host_counter, counter, 2 * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, stream),
          "reading the signal count back");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (host_counter[1] != 0 && kernel_used == SF_KERNEL_FUSED) {
      // A valid pixel beyond the fused kernel's exact range: this frame is
      // thresholded again by the direct kernels, in the reference's types.
      refused_fast++;
      threshold(bits, frame_ptr, height, width, SF_KERNEL_DIRECT, false);
      return sync(found);
    }
    if (found)
      *found = *host_counter;
    return *host_counter > capacity ? -2 : 0;
  }

  void fetch(sf_signal_pixel *out, size_t n) {
    if (n == 0)
      return;
    check(hipMemcpyAsync(out, ordered.ptr, n * sizeof(sf_signal_pixel),
                         hipMemcpyDeviceToHost, stream),
          "copying the signal pixels back");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }
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

bool bad_shape(size_t height, size_t width) {
  return height == 0 || width == 0 || height > 0x7fffffffu / width;
}

std::unique_ptr<sf_workspace> &thread_workspace() {
  static thread_local std::unique_ptr<sf_workspace> ws;
  return ws;
}

double ms_between(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
  return static_cast<double>(ms);
}

} // namespace

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

void sf_set_kernel(int kernel) {
  std::call_once(g_kernel_once, resolve_kernel);
  if (kernel == SF_KERNEL_FUSED || kernel == SF_KERNEL_DIRECT)
    g_kernel.store(kernel);
}

int sf_get_kernel(void) {
  std::call_once(g_kernel_once, resolve_kernel);
  return g_kernel.load();
}

void sf_profile_stages(int on) { g_profile.store(on != 0); }
int sf_profiling_stages(void) { return g_profile.load() ? 1 : 0; }
sf_stage_times sf_last_stage_times(void) { return t_times; }

int sf_find_run(int pixel_bits, const void *frame, size_t height, size_t width,
                size_t *found) {
  if (found)
    *found = 0;
  if (frame == nullptr || bad_shape(height, width) ||
      (pixel_bits != 16 && pixel_bits != 32)) {
    set_last_error("bad argument");
    return -1;
  }
  return guarded([&]() -> int {
    auto &ws = thread_workspace();
    if (!ws) {
      auto fresh = std::make_unique<sf_workspace>();
      fresh->init(nullptr);
      ws = std::move(fresh);
    }
    const StreamGuard idle{ws->stream};
    const size_t bytes = height * width * (pixel_bits / 8);
    ws->upload.ensure(bytes, ws->stream);
    const bool timing = g_profile.load();
    hipEvent_t *ev = timing ? ws->events() : nullptr;
    if (ev)
      check(hipEventRecord(ev[0], ws->stream), "hipEventRecord");
    check(hipMemcpyAsync(ws->upload.ptr, frame, bytes, hipMemcpyHostToDevice,
                         ws->stream),
          "copying a frame to the device");
    ws->threshold(pixel_bits, ws->upload.ptr, static_cast<uint32_t>(height),
                  static_cast<uint32_t>(width), sf_get_kernel(), timing);
    size_t n = 0;
    const int rc = ws->sync(&n);
    if (found)
      *found = n;
    if (timing) {
      sf_stage_times t{};
      t.upload = ms_between(ev[0], ev[1]);
      t.fused = ms_between(ev[1], ev[2]);
      t.sort = ms_between(ev[2], ev[3]);
      t_times = t;
    }
    return rc;
  });
}

int sf_find_fetch(sf_signal_pixel *out, size_t n) {
  return guarded([&]() -> int {
    auto &ws = thread_workspace();
    if (!ws || !ws->ran) {
      set_last_error("nothing to fetch: no frame has been thresholded");
      return -1;
    }
    if (n > ws->capacity || (n > 0 && out == nullptr)) {
      set_last_error("bad argument");
      return -1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    ws->fetch(out, n);
    if (g_profile.load())
      t_times.copy = std::chrono::duration<double, std::milli>(
                         std::chrono::steady_clock::now() - t0)
                         .count();
    return 0;
  });
}

int sf_ws_create(size_t max_pixels, void *stream, sf_workspace **out) {
  if (out == nullptr)
    return -1;
  *out = nullptr;
  return guarded([&]() -> int {
    auto ws = std::make_unique<sf_workspace>();
    ws->init(stream);
    (void)max_pixels; // buffers grow on first use to the frame in hand
    *out = ws.release();
    return 0;
  });
}

void sf_ws_destroy(sf_workspace *ws) { delete ws; }

void *sf_ws_stream(sf_workspace *ws) { return ws ? ws->stream : nullptr; }

int sf_ws_threshold(sf_workspace *ws, int pixel_bits, const void *dev_frame,
                    uint32_t height, uint32_t width, int kernel) {
  if (ws == nullptr || dev_frame == nullptr || bad_shape(height, width) ||
      (pixel_bits != 16 && pixel_bits != 32) ||
      (kernel != SF_KERNEL_FUSED && kernel != SF_KERNEL_DIRECT)) {
    set_last_error("bad argument");
    return -1;
  }
  return guarded([&]() -> int {
    ws->threshold(pixel_bits, dev_frame, height, width, kernel, false);
    return 0;
  });
}

int sf_ws_sync(sf_workspace *ws, size_t *found) {
  if (ws == nullptr || !ws->ran)
    return -1;
  return guarded([&]() -> int { return ws->sync(found); });
}

int sf_ws_fetch(sf_workspace *ws, sf_signal_pixel *out, size_t n) {
  if (ws == nullptr || !ws->ran || n > ws->capacity || (n && !out))
    return -1;
  return guarded([&]() -> int {
    ws->fetch(out, n);
    return 0;
  });
}

void sf_ws_set_timing(sf_workspace *ws, int on) {
  if (ws)
    ws->timing_on = on != 0;
}

int sf_ws_kernel_ms(sf_workspace *ws, double *threshold_ms, double *order_ms) {
  if (ws == nullptr || !ws->timed)
    return -1;
  return guarded([&]() -> int {
    check(hipEventSynchronize(ws->marks[3]), "hipEventSynchronize");
    if (threshold_ms)
      *threshold_ms = ms_between(ws->marks[1], ws->marks[2]);
    if (order_ms)
      *order_ms = ms_between(ws->marks[2], ws->marks[3]);
    return 0;
  });
}

uint64_t sf_ws_fallbacks(sf_workspace *ws) { return ws ? ws->refused_fast : 0; }

const sf_signal_pixel *sf_ws_device_signal(sf_workspace *ws) {
  return ws ? ws->ordered.as<sf_signal_pixel>() : nullptr;
}

const uint32_t *sf_ws_device_mask(sf_workspace *ws, uint32_t *words_per_row) {
  if (!ws)
    return nullptr;
  if (words_per_row)
    *words_per_row = mask_words_per_row(ws->width);
  return ws->mask.as<uint32_t>();
}

} // extern "C"
