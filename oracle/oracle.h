/*
 * The oracle: a CPU restatement of the reference's spot-finding path, used
 * only as the checker in tests, in smoke() and as the CPU baseline leg of
 * bench.py.  Nothing in the product links, loads or calls it.
 *
 * What it restates (reference paths relative to the reference repository):
 *   oracle_dext_u16 / oracle_dext_u32  src/spots/dext.cc:8-225 (dext<T>)
 *   oracle_sort_by_index              src/spots/signal_order.cc:19-72
 *   oracle_labeller_*                 src/spots/dials_spots.cc:36-360
 *   oracle_narrow_to_16, oracle_scan_* src/spots/find_spots.cc:162-182,
 *                                     apps/mxi_max.cc:45-68
 *   oracle_decompress_image etc.      src/spots/decompress.cc:1-136, with the
 *                                     bitshuffle and lz4 codecs it calls
 *                                     (oracle/decompress.c says how those are
 *                                     restated and pinned)
 *
 * Every function here is pinned against oracle/_ref (the reference's own
 * dext.cc, signal_order.cc and dials_spots.cc compiled from their sources) and
 * against the reference tests' hand-derived cases; tests/test_oracle.py holds
 * the checks.
 *
 * Record layouts are the reference's: SignalPixel (signal_pixel.hh:20-28),
 * dials_spots::Pixel, Spot, Counts (dials_spots.hh:50-106).
 */
#ifndef SF_ORACLE_H
#define SF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t index;
  uint32_t value;
  float background;
  uint16_t population;
  uint16_t reserved;
} oracle_signal_pixel;

typedef struct {
  int32_t frame;
  uint32_t index;
  uint32_t value;
} oracle_pixel;

typedef struct {
  int32_t bbox[6];
  double position[3];
  double variance[3];
  double intensity;
  double intensity_variance;
  uint32_t n_signal;
  uint64_t first;
} oracle_spot;

typedef struct {
  uint64_t signal_pixels, groups, too_small, too_large, separated, accepted;
} oracle_counts;

/* dext<T>: 0 on success, -1 on a bad argument, -2 when more than `capacity`
 * pixels were found (the first `capacity` are written, *found is the total). */
int oracle_dext_u16(const uint16_t *image, size_t height, size_t width,
                    oracle_signal_pixel *out, size_t capacity, size_t *found);
int oracle_dext_u32(const uint32_t *image, size_t height, size_t width,
                    oracle_signal_pixel *out, size_t capacity, size_t *found);

/* Ascending by index, in place; duplicates end up adjacent. */
void oracle_sort_by_index(oracle_signal_pixel *pixels, size_t n);

typedef struct oracle_labeller oracle_labeller;

/* NULL on bad options; the message goes to err. */
oracle_labeller *oracle_labeller_new(size_t height, size_t width,
                                     size_t min_spot_size, size_t max_spot_size,
                                     double max_separation, int two_d,
                                     char *err, size_t errlen);
void oracle_labeller_free(oracle_labeller *l);
/* 0 on success, -1 with a message (frame order, pixel order, off-frame). */
int oracle_labeller_add(oracle_labeller *l, int64_t frame,
                        const oracle_signal_pixel *signal, size_t n, char *err,
                        size_t errlen);
void oracle_labeller_finish(oracle_labeller *l);
size_t oracle_labeller_spot_count(const oracle_labeller *l);
const oracle_spot *oracle_labeller_spots(const oracle_labeller *l);
size_t oracle_labeller_pixel_count(const oracle_labeller *l);
const oracle_pixel *oracle_labeller_pixels(const oracle_labeller *l);
oracle_counts oracle_labeller_counts(const oracle_labeller *l);
size_t oracle_labeller_live(const oracle_labeller *l);

/* The CPU baseline: `total` frames of height*width taken in turn from the
 * `nframes` given (frame k is frames[k % nframes]), thresholded by `threads`
 * workers, one frame per worker at a time.  Returns total signal pixels and the
 * wall time in seconds through *seconds. */
uint64_t oracle_dext_u32_threads(const uint32_t *frames, size_t nframes,
                                 size_t total, size_t height, size_t width,
                                 int threads, double *seconds);

/* Decompression (oracle/decompress.c).  oracle_lz4_decode_block: bytes
 * written, or < 0 for a malformed block.  oracle_bslz4_decode: bytes consumed,
 * or bitshuffle's negative code.  oracle_decompress_image: 0, or -1 with the
 * reference's message in err. */
long oracle_lz4_decode_block(const uint8_t *src, size_t n, uint8_t *dst,
                             size_t cap);
void oracle_bitshuffle(const uint8_t *in, uint8_t *out, size_t n, size_t elem);
void oracle_bitunshuffle(const uint8_t *in, uint8_t *out, size_t n, size_t elem);
size_t oracle_bslz4_default_block(size_t elem);
long oracle_bslz4_decode(const uint8_t *in, size_t n, uint8_t *out,
                         size_t count, size_t elem, size_t block);
int oracle_decompress_image(const uint8_t *data, size_t n, const char *algorithm,
                            unsigned bit_depth, size_t height, size_t width,
                            uint8_t *out, size_t out_size, char *err,
                            size_t err_len);

/* --gpu-force's narrowing and mxi_max's scan (oracle/narrow.c). */
typedef struct {
  uint64_t largest, bad, joins, over;
} oracle_frame_scan;
int oracle_narrow_to_16(const uint32_t *in, size_t n, uint16_t *out,
                        uint32_t *offending);
void oracle_scan_u16(const uint16_t *px, size_t n, oracle_frame_scan *s);
void oracle_scan_u32(const uint32_t *px, size_t n, oracle_frame_scan *s);

#ifdef __cplusplus
}
#endif

#endif
