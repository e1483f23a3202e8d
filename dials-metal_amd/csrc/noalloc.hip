// sf_threshold_noalloc: the threshold and ordering of one device frame with
// no device allocation at all, for callers that cannot allocate (code under
// stream capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Everything the kernels need beyond the caller's frame and output list
// lives in `scratch`, a file-scope device array of kScratchBytes that each
// device holds from the library's load until the process ends:
//
//   counters (8 words) | item queue | root tables (float, double) |
//   descriptors (FusedParams, FrameDev, ScanSeg) | bitmask | word prefix |
//   block totals | unordered list
//
// Host code never passes `scratch` itself (its name is only a host-side
// stand-in); it passes the device address hipGetSymbolAddress gives for the
// current device, looked up once per device.  Descriptors and the root
// tables are written by a one-thread kernel from its by-value argument, so
// no host-to-device copy or staging buffer is involved; the counters come
// back into the caller's stack with one copy.  The exact re-run of a frame
// the fused kernel hands back (a valid pixel at or above 2^18) uses the
// unordered list's memory as its first byte mask and the caller's output
// list as its second: each is dead while the other is in use.
//
// Calls on one device share its scratch: calls that overlap in time on one
// device race, and must be serialised by the caller.
#include <cmath>
#include <cstring>
#include <mutex>

#include "sf_internal.hh"

constexpr size_t kScratchBytes = size_t{80} << 20; // 80 MiB a device

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

size_t up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

struct Descriptors {
  sfi::FusedParams fused;
  sfi::FrameDev dev;
  sfi::ScanSeg seg;
  float tf[128];
  double td[128];
};

__global__ void put_descriptors(const Descriptors d, Descriptors *to) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *to = d;
}

// Where each part of the scratch sits for one frame, and what it needs.
struct Layout {
  size_t counters = 0, queue = 0, desc = 0, mask = 0, prefix = 0, totals = 0,
         unordered = 0, end = 0;
  size_t words = 0, capacity = 0;
  uint32_t wpr = 0;
};

Layout layout_for(uint32_t h, uint32_t w, int bits) {
  Layout L;
  const size_t n = static_cast<size_t>(h) * w;
  L.wpr = sfi::mask_words_per_row(w);
  L.words = static_cast<size_t>(h) * L.wpr;
  // The reference's bound: one entry per 16 bytes of frame
  // (dext_cuda.cu:736-741).
  L.capacity = n * static_cast<size_t>(bits / 8) / sizeof(sf_signal_pixel);
  L.counters = 0;
  L.queue = up(L.counters + sfi::kCounters * sizeof(uint32_t));
  L.desc = up(L.queue + sfi::kQueueBytes);
  L.mask = up(L.desc + sizeof(Descriptors));
  L.prefix = up(L.mask + 4 * L.words);
  L.totals = up(L.prefix + 4 * L.words);
  L.unordered = up(L.totals + 4 * sfi::scan_tmp_words(L.words));
  // The list, and never less than a byte a pixel: it is also the exact
  // re-run's first byte mask.
  const size_t list = std::max(std::max<size_t>(L.capacity, 1) *
                                   sizeof(sf_signal_pixel), n);
  L.end = up(L.unordered + list);
  return L;
}

// The output list a caller passes: the frame's list capacity, and at least a
// byte a pixel (the exact re-run's second byte mask).
size_t output_records(uint32_t h, uint32_t w, int bits) {
  const size_t n = static_cast<size_t>(h) * w;
  const size_t cap = n * static_cast<size_t>(bits / 8) / sizeof(sf_signal_pixel);
  return std::max(cap, (n + sizeof(sf_signal_pixel) - 1) / sizeof(sf_signal_pixel));
}

std::once_flag g_scratch_once[kMaxDevices];
unsigned char *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

// The device address of this device's scratch, looked up once per device
// with that device current (the caller's current device).
unsigned char *scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw sfi::device_error("device ordinal out of range for the scratch table");
  std::call_once(g_scratch_once[device], [device] {
    void *p = nullptr;
    g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = static_cast<unsigned char *>(p);
  });
  sfi::check(g_scratch_err[device], "hipGetSymbolAddress (the scratch array)");
  if (g_scratch_ptr[device] == nullptr)
    throw sfi::device_error("hipGetSymbolAddress gave no address for the scratch");
  return g_scratch_ptr[device];
}

int refuse(const std::string &what) {
  sfi::set_last_error("bad argument: " + what);
  return -1;
}

} // namespace

extern "C" size_t sf_noalloc_scratch_bytes(void) { return kScratchBytes; }

extern "C" size_t sf_noalloc_output_records(uint32_t height, uint32_t width,
                                            int pixel_bits) {
  if (pixel_bits != 16 && pixel_bits != 32)
    return 0;
  return output_records(height, width, pixel_bits);
}

extern "C" int sf_threshold_noalloc(int pixel_bits, const void *dev_frame,
                                    uint32_t height, uint32_t width,
                                    sf_signal_pixel *dev_out,
                                    size_t out_records, size_t *found,
                                    void *stream) {
  if (found)
    *found = 0;
  if (dev_frame == nullptr || dev_out == nullptr || found == nullptr ||
      height == 0 || width == 0 ||
      height > 0x7fffffffu / width || (pixel_bits != 16 && pixel_bits != 32))
    return refuse("null frame, list or count, zero size, more than INT32_MAX "
                  "pixels, or not 16 or 32 bits");
  const Layout L = layout_for(height, width, pixel_bits);
  if (L.end > kScratchBytes)
    return refuse("frame of " + std::to_string(height) + " x " +
                  std::to_string(width) + " at " + std::to_string(pixel_bits) +
                  " bits needs " + std::to_string(L.end) +
                  " bytes of scratch; the allocation-free threshold holds " +
                  std::to_string(kScratchBytes));
  if (out_records < output_records(height, width, pixel_bits))
    return refuse("output list of " + std::to_string(out_records) +
                  " records; this frame needs " +
                  std::to_string(output_records(height, width, pixel_bits)));
  try {
    int device = 0;
    sfi::check(hipGetDevice(&device), "hipGetDevice");
    unsigned char *base = scratch_on(device);
    hipStream_t s = static_cast<hipStream_t>(stream);

    uint32_t *counter = reinterpret_cast<uint32_t *>(base + L.counters);
    uint32_t *queue = reinterpret_cast<uint32_t *>(base + L.queue);
    Descriptors *desc_d = reinterpret_cast<Descriptors *>(base + L.desc);
    uint32_t *mask = reinterpret_cast<uint32_t *>(base + L.mask);
    uint32_t *prefix = reinterpret_cast<uint32_t *>(base + L.prefix);
    uint32_t *totals = reinterpret_cast<uint32_t *>(base + L.totals);
    sf_signal_pixel *unordered =
        reinterpret_cast<sf_signal_pixel *>(base + L.unordered);

    Descriptors d;
    std::memset(&d, 0, sizeof d);
    // R(sqrt(R(2) * (R(m) - 1))) for m = 0..127 by the host's correctly
    // rounded sqrt, +inf below two pixels (dext.cc:99-106), as every other
    // entry point uses.
    for (int m = 0; m < 128; m++) {
      d.tf[m] = m >= 2 ? std::sqrt(2.0f * (static_cast<float>(m) - 1.0f)) : HUGE_VALF;
      d.td[m] = m >= 2 ? std::sqrt(2.0 * (static_cast<double>(m) - 1.0)) : HUGE_VAL;
    }
    const float *tf_d = desc_d->tf;
    const double *td_d = desc_d->td;

    sfi::FrameArgs a;
    a.frame = dev_frame;
    a.height = height;
    a.width = width;
    a.bits = pixel_bits;
    a.out = unordered;
    a.counter = counter;
    a.capacity = static_cast<uint32_t>(L.capacity);
    a.mask = mask;
    a.wpr = L.wpr;
    a.table = pixel_bits == 16 ? static_cast<const void *>(tf_d)
                               : static_cast<const void *>(td_d);
    a.overflow = counter + 1;
    a.totals = totals;

    const sfi::Tuning t = sfi::tuning();
    sfi::FusedPlan plan;
    const uint32_t seg = sfi::segment_rows(
        static_cast<uint64_t>(height) * sfi::strips_of(width, t.lane_columns), t);
    const uint32_t items = sfi::plan_fused(plan, height, width, seg, t.lane_columns);
    d.fused = sfi::to_params(a, plan, 0);

    d.dev.unordered = unordered;
    d.dev.ordered = dev_out;
    d.dev.runs = reinterpret_cast<uint4 *>(unordered);
    d.dev.counter = counter;
    d.dev.capacity = a.capacity;
    d.dev.width = width;
    d.dev.height = height;
    d.dev.wpr = L.wpr;
    d.dev.mask = mask;
    d.dev.prefix = prefix;
    d.dev.totals = totals;
    d.seg = sfi::ScanSeg{mask, prefix, totals, L.words, nullptr, 0, 1};

    put_descriptors<<<1, 64, 0, s>>>(d, desc_d);
    sfi::check(hipGetLastError(), "launching the descriptor write");
    const sfi::FrameDev *fd = &desc_d->dev;
    const sfi::ScanSeg *sg = &desc_d->seg;
    sfi::launch_clear(fd, 1, true, 0, s, queue);
    sfi::launch_fused_batch(&desc_d->fused, 1, items, tf_d, td_d, queue,
                            t.lane_columns, t.waves_per_block, s);
    sfi::launch_popcount_scan(sg, 1, L.words, s, true);
    sfi::launch_scatter(fd, 1, s);
    uint32_t host[2] = {0, 0};
    sfi::check(hipMemcpyAsync(host, counter, sizeof host, hipMemcpyDeviceToHost, s),
               "reading the counters back");
    sfi::check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (host[1] != 0 || sfi::fused_narrow(width, t.lane_columns)) {
      // A valid pixel beyond the fused kernel's exact range (or a frame too
      // narrow for it): the direct kernels, in the reference's wide sums.
      sfi::launch_clear(fd, 1, false, 0, s, nullptr);
      sfi::launch_direct(a, reinterpret_cast<uint8_t *>(unordered),
                         reinterpret_cast<uint8_t *>(dev_out), s);
      sfi::launch_popcount_scan(sg, 1, L.words, s, false);
      sfi::launch_scatter(fd, 1, s);
      sfi::check(hipMemcpyAsync(host, counter, sizeof host, hipMemcpyDeviceToHost, s),
                 "reading the counters back");
      sfi::check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    *found = host[0];
    return host[0] > L.capacity ? -2 : 0;
  } catch (const std::exception &e) {
    // Nothing is left in flight on the caller's stream.
    (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
    sfi::set_last_error(e.what());
    return -3;
  }
}
