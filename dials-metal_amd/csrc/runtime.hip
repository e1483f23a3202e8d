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

// What one thresholded frame leaves on the device.  Two of these alternate,
// so the previous frame's bitmask, ranks and labels are still there when the
// next frame is linked to it.
struct FrameSlot {
  DeviceBuffer mask, prefix, ordered, comp_of;
  uint32_t height = 0, width = 0, capacity = 0;
  int bits = 0;
  bool thresholded = false, labelled = false;
  uint64_t serial = 0; // which thresholded frame of this workspace it holds
};

struct sf_workspace {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  FrameSlot slots[2];
  int cur = 0;
  uint64_t serial = 0;
  DeviceBuffer upload, scr0, scr1, unordered, scan_tmp, table;
  // labelling scratch, shared by the slots
  DeviceBuffer parent, roots, root_prefix, label_tmp, comps, links;
  uint32_t *counter = nullptr;      // device: [0] count, [1] overflow,
                                    // [2] components, [3] links
  uint32_t *host_counter = nullptr; // pinned mirror
  hipEvent_t marks[6] = {};
  const void *frame_ptr = nullptr;
  int kernel_used = SF_KERNEL_FUSED;
  uint64_t refused_fast = 0; // frames the fused kernel handed back
  bool timing_on = false;    // events around every sf_ws_threshold
  bool timed = false;        // the last threshold recorded its events
  bool linked = false;       // the last labelling linked to the slot before
  uint32_t link_capacity = 0;

  FrameSlot &slot() { return slots[cur]; }
  FrameSlot &other() { return slots[cur ^ 1]; }

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
    check(hipMalloc(&counter, 4 * sizeof(uint32_t)),
          "hipMalloc for the counters");
    check(hipHostMalloc(reinterpret_cast<void **>(&host_counter),
                        4 * sizeof(uint32_t), hipHostMallocDefault),
          "hipHostMalloc for the counters");
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

  // Sizes the current slot and the shared buffers for an h x w frame.
  void shape(FrameSlot &f, int b, uint32_t h, uint32_t w) {
    const size_t n = static_cast<size_t>(h) * w;
    const size_t words = static_cast<size_t>(h) * mask_words_per_row(w);
    // The reference's bound: a list the size of the frame's bytes
    // (dext_cuda.cu:736-741), one entry per 4 pixels at 32 bits, 8 at 16.
    const size_t cap = n * (b / 8) / sizeof(sf_signal_pixel);
    f.mask.ensure(words * 4, stream);
    f.prefix.ensure(words * 4 + 4, stream);
    f.ordered.ensure((cap ? cap : 1) * sizeof(sf_signal_pixel), stream);
    unordered.ensure((cap ? cap : 1) * sizeof(sf_signal_pixel), stream);
    scan_tmp.ensure(order_scan_bytes(words), stream);
    f.height = h;
    f.width = w;
    f.bits = b;
    f.capacity = static_cast<uint32_t>(cap);
    f.thresholded = true;
    f.labelled = false;
    f.serial = ++serial;
  }

  FrameArgs args(FrameSlot &f, const void *frame) {
    FrameArgs a;
    a.frame = frame;
    a.height = f.height;
    a.width = f.width;
    a.bits = f.bits;
    a.out = unordered.as<sf_signal_pixel>();
    a.counter = counter;
    a.capacity = f.capacity;
    a.mask = f.mask.as<uint32_t>();
    a.wpr = mask_words_per_row(f.width);
    a.table = table_for(f.bits);
    a.overflow = counter + 1;
    return a;
  }

  // Enqueue the threshold and the ordering of a device frame into a fresh
  // slot.
  void threshold(int b, const void *frame, uint32_t h, uint32_t w, int kernel,
                 bool timing) {
    cur ^= 1;
    shape(slot(), b, h, w);
    enqueue(frame, kernel, timing);
  }

  void enqueue(const void *frame, int kernel, bool timing) {
    FrameSlot &f = slot();
    const size_t n = static_cast<size_t>(f.height) * f.width;
    if (kernel == SF_KERNEL_DIRECT) {
      scr0.ensure(n, stream);
      scr1.ensure(n, stream);
    }
    FrameArgs a = args(f, frame);
    frame_ptr = frame;
    kernel_used = kernel;
    timing = timing || timing_on;
    timed = timing;
    hipEvent_t *ev = timing ? events() : nullptr;
    launch_fill(counter, 2, 0u, stream);
    if (kernel == SF_KERNEL_FUSED)
      launch_fill(a.mask, static_cast<size_t>(f.height) * a.wpr, 0u, stream);
    if (ev)
      check(hipEventRecord(ev[1], stream), "hipEventRecord");
    if (kernel == SF_KERNEL_DIRECT)
      launch_direct(a, scr0.as<uint8_t>(), scr1.as<uint8_t>(), stream);
    else
      launch_fused(a, stream);
    if (ev)
      check(hipEventRecord(ev[2], stream), "hipEventRecord");
    launch_order(a, unordered.as<sf_signal_pixel>(),
                 f.ordered.as<sf_signal_pixel>(), f.prefix.as<uint32_t>(),
                 scan_tmp.ptr, scan_tmp.bytes, stream);
    if (ev)
      check(hipEventRecord(ev[3], stream), "hipEventRecord");
  }

  int sync(size_t *found) {
    check(hipMemcpyAsync(host_counter, counter, 4 * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, stream),
          "reading the signal count back");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (host_counter[1] != 0 && kernel_used == SF_KERNEL_FUSED) {
      // A valid pixel beyond the fused kernel's exact range: this frame is
      // thresholded again, into the same slot, by the direct kernels, whose
      // sums are in the reference's types.
      refused_fast++;
      enqueue(frame_ptr, SF_KERNEL_DIRECT, false);
      return sync(found);
    }
    if (found)
      *found = host_counter[0];
    return host_counter[0] > slot().capacity ? -2 : 0;
  }

  void fetch(sf_signal_pixel *out, size_t n) {
    if (n == 0)
      return;
    check(hipMemcpyAsync(out, slot().ordered.ptr, n * sizeof(sf_signal_pixel),
                         hipMemcpyDeviceToHost, stream),
          "copying the signal pixels back");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  // A host list instead of a thresholded frame: upload, and build the
  // bitmask and ranks the labelling reads.
  void load_signal(const sf_signal_pixel *list, size_t n, uint32_t h,
                   uint32_t w) {
    cur ^= 1;
    FrameSlot &f = slot();
    shape(f, 32, h, w);
    if (n > f.capacity) {
      // Labelling a list never needs the threshold's bound; size for it.
      f.capacity = static_cast<uint32_t>(n);
      f.ordered.ensure(n * sizeof(sf_signal_pixel), stream);
    }
    const uint32_t wpr = mask_words_per_row(w);
    const size_t words = static_cast<size_t>(h) * wpr;
    launch_fill(f.mask.as<uint32_t>(), words, 0u, stream);
    host_counter[0] = static_cast<uint32_t>(n);
    host_counter[1] = 0;
    check(hipMemcpyAsync(counter, host_counter, 2 * sizeof(uint32_t),
                         hipMemcpyHostToDevice, stream),
          "uploading the signal count");
    if (n)
      check(hipMemcpyAsync(f.ordered.ptr, list, n * sizeof(sf_signal_pixel),
                           hipMemcpyHostToDevice, stream),
            "uploading a signal list");
    launch_set_bits(f.ordered.as<sf_signal_pixel>(), counter, f.capacity, w,
                    f.mask.as<uint32_t>(), wpr, stream);
    launch_popcount_scan(f.mask.as<uint32_t>(), words, nullptr, 0,
                         f.prefix.as<uint32_t>(), scan_tmp.ptr, stream);
    // The pinned counter is read by the copy above; wait before reuse.
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }

  // Enqueue the labelling of the current slot; link to the other slot when
  // it holds the frame thresholded just before this one, of the same shape.
  void label(bool link_previous) {
    FrameSlot &f = slot();
    FrameSlot &g = other();
    const uint32_t cap = f.capacity;
    parent.ensure((static_cast<size_t>(cap) + 1) * 4, stream);
    const size_t root_words = static_cast<size_t>(cap) / 32 + 2;
    roots.ensure(root_words * 4, stream);
    root_prefix.ensure(root_words * 4, stream);
    label_tmp.ensure(scan_tmp_bytes(root_words), stream);
    f.comp_of.ensure((static_cast<size_t>(cap) + 1) * 4, stream);
    comps.ensure((static_cast<size_t>(cap) + 1) * sizeof(sf_component), stream);
    link_capacity = cap + 1;
    links.ensure(static_cast<size_t>(link_capacity) * 8, stream);
    linked = link_previous && g.labelled && g.serial + 1 == f.serial &&
             g.height == f.height && g.width == f.width;
    LabelArgs a;
    a.list = f.ordered.as<sf_signal_pixel>();
    a.counter = counter;
    a.capacity = cap;
    a.width = f.width;
    a.mask = f.mask.as<uint32_t>();
    a.prefix = f.prefix.as<uint32_t>();
    a.wpr = mask_words_per_row(f.width);
    a.parent = parent.as<uint32_t>();
    a.roots = roots.as<uint32_t>();
    a.root_prefix = root_prefix.as<uint32_t>();
    a.scan_tmp = label_tmp.ptr;
    a.comp_of = f.comp_of.as<uint32_t>();
    a.comps = comps.as<sf_component>();
    a.ncomp = counter + 2;
    a.prev_mask = linked ? g.mask.as<uint32_t>() : nullptr;
    a.prev_prefix = linked ? g.prefix.as<uint32_t>() : nullptr;
    a.prev_comp_of = linked ? g.comp_of.as<uint32_t>() : nullptr;
    a.links = links.as<uint32_t>();
    a.nlinks = counter + 3;
    a.link_capacity = link_capacity;
    launch_fill(counter + 3, 1, 0u, stream);
    launch_label(a, stream);
    f.labelled = true;
  }

  // Waits for the labelling; counts of pixels, components and links.
  void sync_label(size_t *n, size_t *ncomp, size_t *nlinks) {
    check(hipMemcpyAsync(host_counter, counter, 4 * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, stream),
          "reading the label counts back");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (host_counter[3] > link_capacity)
      throw device_error("more frame links than the link list holds");
    if (n)
      *n = host_counter[0];
    if (ncomp)
      *ncomp = host_counter[2];
    if (nlinks)
      *nlinks = host_counter[3];
  }

  void fetch_labels(uint32_t *comp_of, size_t n, sf_component *c, size_t nc,
                    uint32_t *l, size_t nl) {
    if (comp_of && n)
      check(hipMemcpyAsync(comp_of, slot().comp_of.ptr, n * 4,
                           hipMemcpyDeviceToHost, stream),
            "copying the labels back");
    if (c && nc)
      check(hipMemcpyAsync(c, comps.ptr, nc * sizeof(sf_component),
                           hipMemcpyDeviceToHost, stream),
            "copying the components back");
    if (l && nl)
      check(hipMemcpyAsync(l, links.ptr, nl * 8, hipMemcpyDeviceToHost, stream),
            "copying the frame links back");
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
    if (!ws || !ws->slot().thresholded) {
      set_last_error("nothing to fetch: no frame has been thresholded");
      return -1;
    }
    if (n > ws->slot().capacity || (n > 0 && out == nullptr)) {
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
  if (ws == nullptr || !ws->slot().thresholded)
    return -1;
  return guarded([&]() -> int { return ws->sync(found); });
}

int sf_ws_fetch(sf_workspace *ws, sf_signal_pixel *out, size_t n) {
  if (ws == nullptr || !ws->slot().thresholded || n > ws->slot().capacity ||
      (n && !out))
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
  return ws ? ws->slot().ordered.as<sf_signal_pixel>() : nullptr;
}

const uint32_t *sf_ws_device_mask(sf_workspace *ws, uint32_t *words_per_row) {
  if (!ws)
    return nullptr;
  if (words_per_row)
    *words_per_row = mask_words_per_row(ws->slot().width);
  return ws->slot().mask.as<uint32_t>();
}

} // extern "C"

// ---------------------------------------------------------------------------
// Labelling and grouping.
// ---------------------------------------------------------------------------

#include "sf_labeller.hh"

struct sf_labeller {
  explicit sf_labeller(size_t h, size_t w, const sfl::Options &o) : lab(h, w, o) {}
  sfl::Labeller lab;
  std::unique_ptr<sf_workspace> ws; // for host lists (sf_labeller_add)
  const sf_workspace *last_ws = nullptr;
  uint64_t last_serial = 0;
  std::vector<sf_signal_pixel> list;
  std::vector<uint32_t> comp_of, links;
  std::vector<sf_component> comps;
};

namespace {

bool ascending_in_frame(const sf_signal_pixel *s, size_t n, size_t pixels) {
  for (size_t i = 0; i < n; i++) {
    if (s[i].index >= pixels || (i > 0 && s[i].index <= s[i - 1].index))
      return false;
  }
  return true;
}

// Fetches what the workspace's current slot holds and hands it to the
// labeller.
void add_from_workspace(sf_labeller *l, int64_t frame, sf_workspace *ws,
                        const sf_signal_pixel *host_list) {
  size_t n = 0, ncomp = 0, nlinks = 0;
  ws->sync_label(&n, &ncomp, &nlinks);
  const sf_signal_pixel *list = host_list;
  if (!list) {
    l->list.resize(n);
    ws->fetch(l->list.data(), n);
    list = l->list.data();
    l->lab.check_frame(frame, list, n);
  }
  l->comp_of.resize(n);
  l->comps.resize(ncomp);
  l->links.resize(2 * nlinks);
  ws->fetch_labels(l->comp_of.data(), n, l->comps.data(), ncomp,
                   l->links.data(), nlinks);
  const bool to_previous = ws->linked && l->lab.started() &&
                           l->last_ws == ws &&
                           l->last_serial == ws->other().serial;
  l->lab.add(frame, list, n, l->comp_of.data(), l->comps.data(), ncomp,
             l->links.data(), nlinks, to_previous);
  l->last_ws = ws;
  l->last_serial = ws->slot().serial;
}

} // namespace

extern "C" {

int sf_ws_label(sf_workspace *ws, int link_previous) {
  if (ws == nullptr || !ws->slot().thresholded)
    return -1;
  return guarded([&]() -> int {
    ws->label(link_previous != 0);
    return 0;
  });
}

int sf_ws_sync_label(sf_workspace *ws, size_t *n, size_t *ncomp,
                     size_t *nlinks, int *linked) {
  if (ws == nullptr || !ws->slot().labelled)
    return -1;
  return guarded([&]() -> int {
    ws->sync_label(n, ncomp, nlinks);
    if (linked)
      *linked = ws->linked ? 1 : 0;
    return 0;
  });
}

int sf_ws_fetch_labels(sf_workspace *ws, uint32_t *comp_of, size_t n,
                       sf_component *comps, size_t ncomp, uint32_t *links,
                       size_t nlinks) {
  if (ws == nullptr || !ws->slot().labelled || n > ws->slot().capacity ||
      ncomp > ws->slot().capacity + 1 || nlinks > ws->link_capacity)
    return -1;
  return guarded([&]() -> int {
    ws->fetch_labels(comp_of, n, comps, ncomp, links, nlinks);
    return 0;
  });
}

int sf_ws_load_signal(sf_workspace *ws, const sf_signal_pixel *signal,
                      size_t n, uint32_t height, uint32_t width) {
  if (ws == nullptr || bad_shape(height, width) || (n && !signal) ||
      !ascending_in_frame(signal, n, static_cast<size_t>(height) * width)) {
    set_last_error("bad argument");
    return -1;
  }
  return guarded([&]() -> int {
    ws->load_signal(signal, n, height, width);
    return 0;
  });
}

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
    return -1;
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

void sf_labeller_destroy(sf_labeller *l) { delete l; }

int sf_labeller_add(sf_labeller *l, int64_t frame,
                    const sf_signal_pixel *signal, size_t n) {
  if (!l || (n && !signal)) {
    set_last_error("bad argument");
    return -1;
  }
  try {
    l->lab.check_frame(frame, signal, n);
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  }
  return guarded([&]() -> int {
    if (!l->ws) {
      auto fresh = std::make_unique<sf_workspace>();
      fresh->init(nullptr);
      l->ws = std::move(fresh);
    }
    sf_workspace *ws = l->ws.get();
    ws->load_signal(signal, n, static_cast<uint32_t>(l->lab.height()),
                    static_cast<uint32_t>(l->lab.width()));
    ws->label(true);
    add_from_workspace(l, frame, ws, signal);
    return 0;
  });
}

int sf_labeller_add_ws(sf_labeller *l, int64_t frame, sf_workspace *ws) {
  if (!l || !ws || !ws->slot().labelled ||
      ws->slot().height != l->lab.height() || ws->slot().width != l->lab.width()) {
    set_last_error("bad argument: the workspace holds no labelled frame of "
                   "this labeller's shape");
    return -1;
  }
  try {
    add_from_workspace(l, frame, ws, nullptr);
    return 0;
  } catch (const device_error &e) {
    set_last_error(e.what());
    return -3;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  }
}

int sf_labeller_add_labelled(sf_labeller *l, int64_t frame,
                             const sf_signal_pixel *signal, size_t n,
                             const uint32_t *comp_of,
                             const sf_component *comps, size_t ncomp,
                             const uint32_t *links, size_t nlinks,
                             int links_to_previous) {
  if (!l || (n && (!signal || !comp_of)) || (ncomp && !comps) ||
      (nlinks && !links)) {
    set_last_error("bad argument");
    return -1;
  }
  try {
    l->lab.add(frame, signal, n, comp_of, comps, ncomp, links, nlinks,
               links_to_previous != 0);
    l->last_ws = nullptr;
    return 0;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  }
}

int sf_labeller_finish(sf_labeller *l) {
  if (!l)
    return -1;
  l->lab.finish();
  return 0;
}

size_t sf_labeller_spot_count(const sf_labeller *l) {
  return l ? l->lab.spots().size() : 0;
}
const sf_spot *sf_labeller_spots(const sf_labeller *l) {
  return l ? l->lab.spots().data() : nullptr;
}
size_t sf_labeller_pixel_count(const sf_labeller *l) {
  return l ? l->lab.pixels().size() : 0;
}
const sf_pixel *sf_labeller_pixels(const sf_labeller *l) {
  return l ? l->lab.pixels().data() : nullptr;
}
sf_counts sf_labeller_counts(const sf_labeller *l) {
  return l ? l->lab.counts() : sf_counts{};
}
size_t sf_labeller_live(const sf_labeller *l) { return l ? l->lab.live() : 0; }

} // extern "C"
