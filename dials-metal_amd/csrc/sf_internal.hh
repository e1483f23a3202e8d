// Host-side declarations shared by the translation units of the library: the
// kernel launchers and the error plumbing.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_copy.hh"
#include "spotfinder_hip.h"

namespace sfi {

// A HIP failure, carried as an exception inside the library and turned into
// status -3 plus sf_last_error() at the C boundary.
struct device_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw device_error(std::string(what) + ": " + hipGetErrorName(e) + ", " +
                       hipGetErrorString(e));
}

void set_last_error(const std::string &message);


// The fused kernel's launch shape (sf_tuning in the public header).
using Tuning = sf_tuning;
Tuning tuning();
// The launch shape for a batch whose frames make `mix_rows` strip-rows (the
// sum over its frames of height x strips at four columns a lane): the shape
// sf_autotune_mix measured for batches of that size, where there is one,
// else tuning().
Tuning tuning_for_mix(uint64_t mix_rows);

// A grow-only device allocation.  Growing waits for the stream first: work in
// flight may still be reading the old buffer (dext_cuda.cu:506-525).
// Pinned host memory that only grows.
class Pinned {
public:
  Pinned() = default;
  Pinned(const Pinned &) = delete;
  Pinned &operator=(const Pinned &) = delete;
  ~Pinned() {
    if (p_)
      (void)hipHostFree(p_);
  }
  template <typename T> T *get(size_t count) {
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    if (bytes > cap_) {
      const size_t next = std::max(bytes, 2 * cap_);
      void *q = nullptr;
      check(hipHostMalloc(&q, next, hipHostMallocDefault),
            "hipHostMalloc (pinned staging)");
      if (p_)
        (void)hipHostFree(p_);
      p_ = q;
      cap_ = next;
    }
    return static_cast<T *>(p_);
  }

private:
  void *p_ = nullptr;
  size_t cap_ = 0;
};

struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  ~DeviceBuffer() {
    if (ptr)
      (void)hipFree(ptr);
  }
  // True when the buffer moved, which invalidates any graph holding it.
  bool ensure(size_t want, hipStream_t stream) {
    if (want <= bytes)
      return false;
    if (ptr) {
      check(hipStreamSynchronize(stream), "waiting before a buffer grows");
      check(hipFree(ptr), "hipFree");
      ptr = nullptr;
      bytes = 0;
    }
    check(hipMalloc(&ptr, want), "hipMalloc");
    bytes = want;
    return true;
  }
  // Grows keeping the contents (the stream's work on the old buffer is
  // waited for); true when the buffer moved.
  bool ensure_keep(size_t want, hipStream_t stream) {
    if (want <= bytes)
      return false;
    void *next = nullptr;
    check(hipMalloc(&next, want), "hipMalloc");
    if (ptr) {
      check(hipMemcpyAsync(next, ptr, bytes, hipMemcpyDeviceToDevice, stream),
            "keeping a buffer's contents");
      check(hipStreamSynchronize(stream), "waiting before a buffer grows");
      check(hipFree(ptr), "hipFree");
    }
    ptr = next;
    bytes = want;
    return true;
  }
  template <typename P> P *as() const { return static_cast<P *>(ptr); }
};

// Words per row of the strong-pixel bitmask: one bit per pixel, rows padded
// to whole 32-bit words, bit x % 32 of word y * wpr + x / 32.
inline uint32_t mask_words_per_row(uint32_t width) { return (width + 31) / 32; }

// Geometry and buffers of one frame's threshold.
struct FrameArgs {
  const void *frame;      // device, height * width pixels
  uint32_t height, width;
  int bits;               // 16 or 32
  sf_signal_pixel *out;   // unordered list, `capacity` entries
  uint32_t *counter;      // device, cleared by the caller
  uint32_t capacity;
  uint32_t *mask;         // strong-pixel bitmask, height * wpr words
  uint32_t wpr;
  const void *table;      // R(sqrt(2 (m - 1))) for m = 0..127, R per bits
  uint32_t *overflow;     // device word, set by the fused kernel when a
                          // valid pixel is too large for its fast sums
  uint32_t *totals;       // the bitmask's set bits per kScanPerBlock words,
                          // counted by the fused kernel (cleared before it)
};

// The three direct-window kernels (the cross-check form): scr0 and scr1 are
// byte masks of height * width.
void launch_direct(const FrameArgs &a, uint8_t *scr0, uint8_t *scr1,
                   hipStream_t stream);

// One frame as the fused kernel sees it; a batch is an array of these in
// device memory, with item_base the frame's first work item.
struct FusedParams {
  const void *frame;
  uint32_t height, width;
  sf_signal_pixel *out;
  uint32_t *counter;
  uint32_t capacity;
  uint32_t *mask;
  uint32_t wpr;
  const void *table;
  uint32_t strips, segments, seg_rows;
  uint32_t *overflow;
  uint32_t *totals;
  uint32_t bits;
  uint32_t item_base;
};

// A frame's cut into work items (strips x segments) under a tuning.
struct FusedPlan {
  uint32_t strips = 0, segments = 0, seg_rows = 0;
  int lane_columns = 4;
};
// Strips of a frame's width with `lane_columns` columns a lane; rows per
// segment for a batch of `strip_rows` (the sum over its frames of height x
// strips) under a tuning; the plan and its work items, none for a frame
// narrower than a lane's columns (fused_narrow), which the exact kernels take.
uint32_t strips_of(uint32_t width, int lane_columns);
uint32_t segment_rows(uint64_t strip_rows, const Tuning &t);
uint32_t plan_fused(FusedPlan &p, uint32_t height, uint32_t width,
                    uint32_t seg_rows, int lane_columns);
FusedParams to_params(const FrameArgs &a, const FusedPlan &plan,
                      uint32_t item_base);
inline bool fused_narrow(uint32_t width, int lane_columns) {
  return width < static_cast<uint32_t>(lane_columns);
}
// The fused launch's item counters: one per XCD of the device, each in its
// own 256-byte line (kQueueStride words apart), kQueueBytes in all.
constexpr int kQueueParts = 8;
constexpr int kQueueStride = 64;
constexpr size_t kQueueBytes = kQueueParts * kQueueStride * sizeof(uint32_t);

// One persistent launch over every item of `nframes` frames (descriptors in
// device memory, planned with `lane_columns`), mixed sizes and pixel types;
// each frame's mask and overflow word cleared beforehand, and `queue`, the
// kQueueBytes of counters from which waves take items, zeroed.
void launch_fused_batch(const FusedParams *frames_device, uint32_t nframes,
                        uint32_t total_items, const float *tf, const double *td,
                        uint32_t *queue, int lane_columns, int waves_per_block,
                        hipStream_t stream);

// Clears `n` 32-bit words with a kernel, which unlike a memset node stays
// correct when the sequence is captured into a graph and replayed.
void launch_fill(uint32_t *p, size_t n, uint32_t value, hipStream_t stream);

// The device's address for `p` when it is pinned host memory mapped into
// the device's address space (hipHostMalloc'd), else null (pageable memory).
const void *mapped_host(const void *p);

// Copies `bytes` from pinned host memory, through its device-side mapping
// `mapped_src`, to device memory `dst`, with a kernel; both addresses are
// 16-byte aligned.
void launch_pull(const void *mapped_src, void *dst, size_t bytes,
                 hipStream_t stream);

// One frame's bit array to scan: the exclusive prefix of per-word popcounts
// (frame-local, from zero).  With `live` (a device count of bits in use,
// refused as zero above `capacity`), words past the live bits read as zero.
// `totals` holds scan_tmp_words(words) words.
constexpr int kScanPerBlock = 4096; // words per scan block
struct ScanSeg {
  const uint32_t *bits;
  uint32_t *prefix;
  uint32_t *totals;
  size_t words;
  const uint32_t *live;
  uint32_t capacity;
  // Nonzero: only words with a set bit get their prefix (every reader of a
  // strong-pixel mask's prefix reads it for such words alone).
  uint32_t sparse;
};
size_t scan_tmp_words(size_t words);
// One scan per frame, all frames in three launches (two when the block
// totals are already counted); descriptors on device.
void launch_popcount_scan(const ScanSeg *segs_device, uint32_t nframes,
                          size_t max_words, hipStream_t stream,
                          bool totals_ready = false);

// Words of counters per frame: [0] signal pixels [1] the fused kernel's
// overflow flag [2] components [3] frame links [4] row runs.
constexpr int kCounters = 8;

// One frame's device state after its threshold, for ordering and labelling.
struct FrameDev {
  sf_signal_pixel *unordered; // the threshold's emit
  sf_signal_pixel *ordered;   // ascending by index
  // The labelling's row runs {first rank, x0, y, length}, in the unordered
  // list's memory (free once the list is ordered; runs <= pixels).
  uint4 *runs;
  uint32_t *counter;          // kCounters words, as above
  uint32_t capacity;
  uint32_t width, height, wpr;
  uint32_t *mask, *prefix;    // strong-pixel bitmask and its word prefix
  uint32_t *totals;           // the bitmask's set bits per kScanPerBlock words
  // labelling
  uint32_t *parent, *roots, *root_prefix, *comp_of;
  uint32_t *row_runs; // height words: runs per row, then each row's first run
  // components + 1 words: runs per component, then each component's first
  // run in component order (and the run count after the last)
  uint32_t *comp_runs;
  sf_component *comps;
  // the previous frame of the same shape, for three-dimensional links; its
  // counters say whether it was labelled (a refused, overfull frame has a
  // bitmask but no list, so it is not linked to)
  const uint32_t *prev_mask, *prev_prefix, *prev_comp_of, *prev_counter;
  uint32_t prev_capacity;
  uint32_t *links;            // 2 * link_capacity words
  uint32_t link_capacity;
};

// Clears each frame's counters from `first` to the last and, with `masks`,
// its bitmask and block totals; zeroes `queue` when given (the fused
// kernel's item queue).
void launch_clear(const FrameDev *frames_device, uint32_t nframes, bool masks,
                  int first, hipStream_t stream, uint32_t *queue = nullptr);
// Places every frame's unordered emit at its rank (blockIdx.y = frame).
void launch_scatter(const FrameDev *frames_device, uint32_t nframes,
                    hipStream_t stream);
// Sets the bitmask bit of every pixel of each frame's ordered list (a list
// loaded from the host); the masks start clear.
void launch_set_bits(const FrameDev *frames_device, uint32_t nframes,
                     hipStream_t stream);
// Labels every frame: components, moments, and links where prev_mask is set.
void launch_label(const FrameDev *frames_device, uint32_t nframes,
                  hipStream_t stream);

// The device decoder behind sf_decoder_*, for the serving loop: load and run
// as sf_decoder_load/run (throwing instead of returning a status), and the
// last run's status per chunk once `after` has fired, read on `read_stream`
// (0 or -1, with each chunk's message).  check_chunk throws, with the
// reference's message, for a chunk that does not describe its frame.
void decoder_load(sf_decoder *d, const sf_chunk *chunks, size_t n);
void decoder_run(sf_decoder *d);
int decoder_status_after(sf_decoder *d, hipStream_t read_stream, hipEvent_t after,
                         int *status, std::vector<std::string> *messages);
void check_chunk(const sf_chunk &c);

} // namespace sfi
