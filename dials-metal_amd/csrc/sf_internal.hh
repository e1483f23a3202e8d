// Host-side declarations shared by the translation units of the library: the
// kernel launchers and the error plumbing.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

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
};

// The three direct-window kernels (the cross-check form): scr0 and scr1 are
// byte masks of height * width.
void launch_direct(const FrameArgs &a, uint8_t *scr0, uint8_t *scr1,
                   hipStream_t stream);

// The fused single-kernel threshold.
void launch_fused(const FrameArgs &a, hipStream_t stream);

// Ordering of an unordered list by the bitmask's ranks: prefix is
// height * wpr + 1 words of scratch; `scan_tmp` is grown as needed.
size_t order_scan_bytes(size_t words);
void launch_order(const FrameArgs &a, const sf_signal_pixel *unordered,
                  sf_signal_pixel *ordered, uint32_t *prefix, void *scan_tmp,
                  size_t scan_tmp_bytes, hipStream_t stream);

// Clears `n` 32-bit words with a kernel, which unlike a memset node stays
// correct when the sequence is captured into a graph and replayed.
void launch_fill(uint32_t *p, size_t n, uint32_t value, hipStream_t stream);

// Exclusive prefix of per-word popcounts over `words` words.  With `live`
// (a device count of bits in use, refused as zero above `capacity`), words
// past the live bits are read as zero.  `tmp` needs scan_tmp_bytes(words).
size_t scan_tmp_bytes(size_t words);
void launch_popcount_scan(const uint32_t *bits, size_t words,
                          const uint32_t *live, uint32_t capacity,
                          uint32_t *prefix, void *tmp, hipStream_t stream);

// One frame's labelling: the ordered list and its bitmask in, components out.
struct LabelArgs {
  const sf_signal_pixel *list; // ordered, `*counter` entries
  const uint32_t *counter;
  uint32_t capacity;
  uint32_t width;
  const uint32_t *mask, *prefix; // the frame's bitmask and word prefix
  uint32_t wpr;
  uint32_t *parent;      // capacity
  uint32_t *roots;       // capacity / 32 + 2 words
  uint32_t *root_prefix; // same
  void *scan_tmp;        // scan_tmp_bytes(capacity / 32 + 2)
  uint32_t *comp_of;     // capacity
  sf_component *comps;   // capacity
  uint32_t *ncomp;       // device word
  // The previous frame of the same shape, for three-dimensional links, or
  // null.
  const uint32_t *prev_mask, *prev_prefix, *prev_comp_of;
  uint32_t *links;       // 2 * link_capacity words
  uint32_t *nlinks;      // device word
  uint32_t link_capacity;
};

void launch_label(const LabelArgs &a, hipStream_t stream);

// Sets the bitmask bit of every pixel of an ordered list of `*counter`
// entries (at most `capacity`); the mask starts clear.
void launch_set_bits(const sf_signal_pixel *list, const uint32_t *counter,
                     uint32_t capacity, uint32_t width, uint32_t *mask,
                     uint32_t wpr, hipStream_t stream);

} // namespace sfi
