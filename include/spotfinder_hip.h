/*
 * spotfinder_hip: the C ABI of the MI355X (gfx950) backend for the reference's
 * spot-finding hot path -- the extended dispersion threshold, strong-pixel
 * connected-component labelling and per-spot reductions of src/spots.
 *
 * Plain C: pointers, sizes and POD records, integer status codes, no
 * exceptions and no HIP or C++ types.  The C++ gpu:: interface of the
 * reference (src/spots/dext_gpu.hh) is implemented on top of this in
 * dext_gpu.hh / dext_gpu_hip.cc; a binding from any other language calls these
 * functions directly.
 *
 * Status codes: 0 success; -1 bad argument; -2 more signal pixels than the
 * device list holds (nothing is returned, as dext_cuda.cu:821-822); -3 a
 * device or runtime failure, with the runtime's message in sf_last_error();
 * -4 out of memory; -5 unsupported request; -6 not admitted (a server has
 * no room for another sweep).  sf_last_error() describes the last failure
 * on the calling thread.
 */
#ifndef SPOTFINDER_HIP_H
#define SPOTFINDER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Records.  Layouts are the reference's, byte for byte.
 * ------------------------------------------------------------------------ */

/* SignalPixel, src/spots/signal_pixel.hh:20-28. */
typedef struct sf_signal_pixel {
  uint32_t index;      /* row-major position in the frame */
  uint32_t value;      /* the raw pixel */
  float background;    /* the masked local mean at that pixel */
  uint16_t population; /* pixels that mean was averaged over */
  uint16_t reserved;
} sf_signal_pixel;

/* dials_spots::Pixel, src/spots/dials_spots.hh:50-54. */
typedef struct sf_pixel {
  int32_t frame;
  uint32_t index;
  uint32_t value;
} sf_pixel;

/* dials_spots::Spot, src/spots/dials_spots.hh:57-79. */
typedef struct sf_spot {
  int32_t bbox[6]; /* x0, x1, y0, y1, z0, z1; upper bounds exclusive */
  double position[3];
  double variance[3];
  double intensity;
  double intensity_variance;
  uint32_t n_signal;
  uint64_t first; /* where its pixels start in the pixel list */
} sf_spot;

/* dials_spots::Counts, src/spots/dials_spots.hh:99-106. */
typedef struct sf_counts {
  uint64_t signal_pixels, groups, too_small, too_large, separated, accepted;
} sf_counts;

/* One four-connected component of one frame, reduced on the device.  New:
 * the reference has no device labelling.  All sums are exact integers; the
 * x and y moments are taken at doubled pixel centres (2x+1, 2y+1) so that
 * they stay integers, and the 128-bit sums are split into lo/hi words. */
typedef struct sf_component {
  uint32_t first;     /* lowest flat index in the component */
  uint32_t n;         /* pixels */
  uint32_t x0, x1;    /* bbox columns, x1 exclusive */
  uint32_t y0, y1;    /* bbox rows, y1 exclusive */
  /* The brightest pixel, ties to the lowest index:
   * (value << 32) | (0xFFFFFFFF - index). */
  uint64_t peak;
  uint64_t sum;        /* sum of values */
  uint64_t sum_sq[2];  /* sum of value^2, lo/hi */
  uint64_t sum_x[2];   /* sum of value*(2x+1), lo/hi */
  uint64_t sum_y[2];   /* sum of value*(2y+1), lo/hi */
  uint64_t sum_xx[2];  /* sum of value*(2x+1)^2, lo/hi */
  uint64_t sum_yy[2];  /* sum of value*(2y+1)^2, lo/hi */
} sf_component;

/* ------------------------------------------------------------------------
 * Device and host memory: gpu::available, unavailable_reason, memory_free,
 * host_alloc, host_free (src/spots/dext_gpu.hh:137-154), backend (:29).
 * ------------------------------------------------------------------------ */

const char *sf_backend(void);              /* "HIP" */
int sf_available(void);                    /* 1 if a device is usable */
const char *sf_unavailable_reason(void);   /* "" when available */
size_t sf_memory_free(void);               /* device bytes free, 0 on error */
void *sf_host_alloc(size_t bytes);         /* pinned; NULL on failure */
void sf_host_free(void *pointer);
const char *sf_last_error(void);           /* this thread's last failure */

/* Device memory for callers that hold frames in HBM, so that a binding
 * needs no second GPU runtime of its own.  Copies are synchronous; 0 or -3. */
void *sf_device_alloc(size_t bytes);       /* NULL on failure */
void sf_device_free(void *pointer);
int sf_copy_to_device(void *dst, const void *src, size_t bytes);
int sf_copy_to_host(void *dst, const void *src, size_t bytes);
int sf_copy_on_device(void *dst, const void *src, size_t bytes);
int sf_device_synchronize(void);

/* ------------------------------------------------------------------------
 * Kernel choice: replaces gpu::stage0_window/stage2_window/fused
 * (src/spots/dext_gpu.hh:50-82).  Both compute the same thing.
 *   SF_KERNEL_FUSED   one fused kernel per frame (default)
 *   SF_KERNEL_DIRECT  three direct-window kernels, the cross-check
 * ------------------------------------------------------------------------ */

enum { SF_KERNEL_FUSED = 0, SF_KERNEL_DIRECT = 1 };
void sf_set_kernel(int kernel);
int sf_get_kernel(void);

/* The fused kernel's launch shape, which the autotuner writes.  A wave sweeps
 * a strip of 64 x `lane_columns` columns (4 or 8 a lane) down `seg_rows` rows
 * (plus 20 rows of halo).  Segments are sized so that a launch makes about
 * `target_waves` waves, 40 to 256 rows each; a non-zero seg_rows caps that
 * length (so a lone frame under a shape tuned for batches still spreads over
 * the device).  `waves_per_block` is 1, 2 or 4;
 * `frames_per_launch` is the batched mode's batch.  No counterpart in the
 * reference, whose launch shape is fixed (dext_cuda.cu:728-730).  From
 * SPOTFINDER_HIP_TUNE, e.g. "seg_rows=128,target_waves=4096,
 * waves_per_block=2,frames_per_launch=32,lane_columns=8", when set.
 * sf_set_tuning returns -1 for a value out of range. */
typedef struct sf_tuning {
  int seg_rows;
  int target_waves;
  int waves_per_block;
  int frames_per_launch;
  int lane_columns;
} sf_tuning;
void sf_get_tuning(sf_tuning *t);
int sf_set_tuning(const sf_tuning *t);

/* Per-stage times of this thread's last profiled find, gpu::StageTimes
 * (src/spots/dext_gpu.hh:89-112), in milliseconds. */
typedef struct sf_stage_times {
  double upload, stage0, stage1, stage2, fused, copy, sort;
} sf_stage_times;
void sf_profile_stages(int on);
int sf_profiling_stages(void);
sf_stage_times sf_last_stage_times(void);

/* ------------------------------------------------------------------------
 * gpu::find<T> (src/spots/dext_gpu.hh:171-180; dext_cuda.cu:710-869), split
 * into run and fetch so a caller sizes its list once it knows the count.
 *
 * sf_find_run thresholds one host frame (pixel_bits 16 or 32) on this
 * thread's own stream and workspace, and leaves the signal pixels on the
 * device, ascending by index.  *found is the count.  -1 for a null frame, a
 * zero size or height*width above INT32_MAX; -2 if the count exceeds the
 * device list (frame bytes / 16 entries, the reference's bound).
 *
 * sf_find_fetch copies the first n of them to `out`.  No work is in flight
 * when either returns.
 * ------------------------------------------------------------------------ */

int sf_find_run(int pixel_bits, const void *frame, size_t height, size_t width,
                size_t *found);
int sf_find_fetch(sf_signal_pixel *out, size_t n);

/* sf_find_run with the kernel (SF_KERNEL_FUSED or SF_KERNEL_DIRECT) and the
 * stage timing chosen for this call alone, not read from the process-wide
 * settings: what gpu::find<T> calls, so that threads finding frames at once
 * under different window choices (find_spots.cc:732-835) never steer each
 * other's launches.  -1 also for an unknown kernel. */
int sf_find_run_as(int pixel_bits, const void *frame, size_t height,
                   size_t width, int kernel, int timing, size_t *found);

/* Failure injection for tests of the error path (dext_cuda.cu:46-51,
 * 445-467; find_spots.cc:410-416): the nth sf_find_run / sf_find_run_as call
 * from now, counted over all threads, fails with -3 after its frame's work is
 * enqueued, as a device failure would; 0 cancels.  No counterpart in the
 * reference.  sf_find_idle is 1 when the calling thread's find stream has no
 * work in flight (it must be, however find returns), else 0. */
void sf_inject_find_failure(uint64_t nth);
int sf_find_idle(void);

/* ------------------------------------------------------------------------
 * Allocation-free threshold.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under stream capture, tight
 * serving loops, harnesses that forbid hipMalloc).
 *
 * sf_threshold_noalloc is sf_ws_threshold + sf_ws_sync for one device frame
 * (pixel_bits 16 or 32), with the ordered SignalPixel list written to the
 * caller's device list `dev_out` (out_records entries, at least
 * sf_noalloc_output_records(height, width, pixel_bits): the frame's list
 * capacity, frame bytes / 16, and never under one record per 16 pixels) and
 * the count to *found.  It calls no allocator: every other buffer is in a
 * static device array of sf_noalloc_scratch_bytes() (80 MiB), which each
 * device holds until the process ends.  Frames whose scratch does not fit
 * in it are refused with -1 and a message (every frame up to 4148 x 4362 at
 * 32 bits fits).  Work is enqueued on `stream` (NULL: the null stream) and
 * waited for.  0; -1 bad argument; -2 more signal pixels than the list's
 * capacity (*found is the count, the list is not written); -3 device
 * failure.  All calls on one device share its scratch, which keeps the last
 * call's contents: calls that overlap in time on one device race.
 * ------------------------------------------------------------------------ */
int sf_threshold_noalloc(int pixel_bits, const void *dev_frame, uint32_t height,
                         uint32_t width, sf_signal_pixel *dev_out,
                         size_t out_records, size_t *found, void *stream);
size_t sf_noalloc_scratch_bytes(void);
size_t sf_noalloc_output_records(uint32_t height, uint32_t width, int pixel_bits);

/* ------------------------------------------------------------------------
 * Device-resident frames.  New surface, for callers that keep frames in HBM
 * (the batched and serving paths, and the benchmark).
 *
 * A workspace holds the device buffers for frames up to max_pixels and one
 * stream.  sf_ws_threshold enqueues the threshold of one device frame and
 * the ordering of its signal pixels on that stream and returns at once;
 * sf_ws_sync waits and reports the count; results stay on the device until
 * the next call.  `stream` (a hipStream_t, or NULL for the workspace's own)
 * is borrowed, never destroyed.
 * ------------------------------------------------------------------------ */

typedef struct sf_workspace sf_workspace;

int sf_ws_create(size_t max_pixels, void *stream, sf_workspace **out);
void sf_ws_destroy(sf_workspace *ws);
void *sf_ws_stream(sf_workspace *ws);
int sf_ws_threshold(sf_workspace *ws, int pixel_bits, const void *dev_frame,
                    uint32_t height, uint32_t width, int kernel);
int sf_ws_sync(sf_workspace *ws, size_t *found);
int sf_ws_fetch(sf_workspace *ws, sf_signal_pixel *out, size_t n);
/* With timing on, the last sf_ws_threshold's kernel times in milliseconds
 * from HIP events on the workspace's stream: the threshold kernel(s) alone,
 * and the ordering kernels that follow. */
void sf_ws_set_timing(sf_workspace *ws, int on);
int sf_ws_kernel_ms(sf_workspace *ws, double *threshold_ms, double *order_ms);
/* Frames this workspace's fused kernel handed back to the direct kernels
 * because a valid pixel was beyond its exact fast range. */
uint64_t sf_ws_fallbacks(sf_workspace *ws);
/* Labelling of the last thresholded frame, enqueued on the workspace's
 * stream: four-connected components numbered in order of their first pixel,
 * each pixel's component, each component's exact moments (sf_component),
 * and, with link_previous, the (component here, component on the previous
 * frame) pairs of positions set in both -- taken only when the workspace's
 * previous frame was thresholded and labelled just before and has the same
 * shape.  New surface: the reference labels on the host only
 * (dials_spots.cc:57-145). */
int sf_ws_label(sf_workspace *ws, int link_previous);
/* Waits; pixel, component and link counts.  *linked says whether links were
 * taken. */
int sf_ws_sync_label(sf_workspace *ws, size_t *n, size_t *ncomp,
                     size_t *nlinks, int *linked);
int sf_ws_fetch_labels(sf_workspace *ws, uint32_t *comp_of, size_t n,
                       sf_component *comps, size_t ncomp, uint32_t *links,
                       size_t nlinks);
/* Loads an ordered host list (as a threshold would leave it) so that it can
 * be labelled: -1 if it is not ascending or runs off an h x w frame. */
int sf_ws_load_signal(sf_workspace *ws, const sf_signal_pixel *signal,
                      size_t n, uint32_t height, uint32_t width);
/* Device pointers of the last frame's results. */
const sf_signal_pixel *sf_ws_device_signal(sf_workspace *ws);
const uint32_t *sf_ws_device_mask(sf_workspace *ws, uint32_t *words_per_row);

/* ------------------------------------------------------------------------
 * Batches.  New surface, no counterpart in the reference, which thresholds
 * one frame per call: many device frames -- any mix of sizes and of 16- and
 * 32-bit pixels -- thresholded in one persistent launch of the fused kernel,
 * each frame's signal pixels ordered, and the frames labelled, in a fixed
 * number of launches whatever the batch size.  With use_graph the sequence
 * is captured into a hipGraph the first time a batch of these shapes is seen
 * and replayed after.  A batch of n frames takes the buffers of n + 1 frames
 * at most: the previous batch's last frame is kept, so that frame 0 of a
 * batch can be linked to it; the rest of the previous batch is overwritten
 * (read it before the next sf_batch_threshold).
 * ------------------------------------------------------------------------ */

typedef struct sf_frame_desc {
  const void *frame; /* device pointer */
  uint32_t height, width;
  int32_t bits;      /* 16 or 32 */
} sf_frame_desc;

typedef struct sf_batch sf_batch;
int sf_batch_create(void *stream, sf_batch **out);
void sf_batch_destroy(sf_batch *b);
void *sf_batch_stream(sf_batch *b);
/* Device bytes the batch's frame buffers and scratch hold (they only grow). */
size_t sf_batch_device_bytes(const sf_batch *b);
/* Enqueues; returns at once.  At most 4096 frames. */
int sf_batch_threshold(sf_batch *b, const sf_frame_desc *frames, size_t n,
                       int use_graph);
/* Waits; counts[i] signal pixels and status[i] 0 or -2 per frame. */
int sf_batch_sync(sf_batch *b, size_t *counts, int *status);
int sf_batch_fetch(sf_batch *b, size_t i, sf_signal_pixel *out, size_t n);
/* Enqueues the labelling of the last batch, each frame linked to the one
 * before it (frame 0 to the previous batch's last) when link_previous is set
 * and the shapes agree. */
int sf_batch_label(sf_batch *b, int link_previous);
int sf_batch_label_counts(sf_batch *b, size_t i, size_t *n, size_t *ncomp,
                          size_t *nlinks);
int sf_batch_fetch_labels(sf_batch *b, size_t i, uint32_t *comp_of, size_t n,
                          sf_component *comps, size_t ncomp, uint32_t *links,
                          size_t nlinks);
void sf_batch_set_timing(sf_batch *b, int on);
int sf_batch_kernel_ms(sf_batch *b, double *threshold_ms, double *order_ms);
uint64_t sf_batch_fallbacks(sf_batch *b);

/* The autotuner.  No counterpart in the reference, whose launch shape is fixed
 * (dext_cuda.cu:728-730).  On the caller's own device-resident frames it
 * measures the fused kernel's shape -- waves per block {1, 2, 4} by segment
 * sizing {64, 128, 192, 256 rows, or automatic at 4096/8192/16384 target
 * waves} -- by the kernel's own duration over the first (up to 128) frames,
 * then frames per launch {16, 32, 64, 128} by the whole step per frame
 * (threshold, ordering and, with `label`, labelling) among the sizes whose
 * slots fit in 60% of free device memory, and installs the winner with
 * sf_set_tuning.  On failure the tuning is left as it was. */
typedef struct sf_autotune_report {
  sf_tuning best;
  double fused_ms_per_frame; /* the best shape's kernel time per frame */
  double step_ms_per_frame;  /* the best launch size's step time per frame */
  uint32_t shapes_tried, launch_sizes_tried;
} sf_autotune_report;
int sf_autotune(const sf_frame_desc *frames, size_t n, int label,
                sf_autotune_report *report);
/* The fused kernel's shape for batches like this one: the same 42 shapes
 * measured on these frames (all of them, at most 4096), the winner kept for
 * every later batch of about this size -- its frames' height x strips at four
 * columns a lane, within a factor of the square root of two -- leaving the
 * process-wide tuning (and frames per launch) as they were.  `best` and
 * `fused_ms_per_frame` report it.  No counterpart in the reference. */
int sf_autotune_mix(const sf_frame_desc *frames, size_t n, sf_autotune_report *report);
/* The tuning a batch of these frames would launch with; 1 when it comes from
 * sf_autotune_mix, 0 when it is the process-wide one. */
int sf_tuning_for_mix(const sf_frame_desc *frames, size_t n, sf_tuning *t);
/* Forgets every shape sf_autotune_mix kept. */
void sf_clear_mix_tunings(void);

/* ------------------------------------------------------------------------
 * Grouping into spots: dials_spots::Labeller (src/spots/dials_spots.hh:
 * 108-158) with the pixel-level labelling and reductions on the device.
 *
 * sf_labeller_add is Labeller::add: one frame's ascending signal list; frames
 * strictly ascending; a gap in numbering closes every open spot; -1 with the
 * reference's message for a frame out of order, a pixel off the frame or
 * pixels out of order.  sf_labeller_add_ws adds the frame a workspace has
 * just thresholded and labelled (sf_ws_threshold, sf_ws_label), without
 * copying its list back first.  Spots, pixels and counts are the reference's
 * records, in its order.
 * ------------------------------------------------------------------------ */

/* dials_spots::Options (src/spots/dials_spots.hh:81-95). */
typedef struct sf_grouping {
  size_t min_spot_size;   /* 3 */
  size_t max_spot_size;   /* 1000 */
  double max_separation;  /* 2.0; 0 or less turns the filter off */
  int two_d;              /* 0: six-connected across frames */
} sf_grouping;
void sf_grouping_default(sf_grouping *g);

typedef struct sf_labeller sf_labeller;
int sf_labeller_create(size_t height, size_t width, const sf_grouping *g,
                       sf_labeller **out);
/* The same grouping (dials_spots::Labeller, src/spots/dials_spots.hh:108-158)
 * with the pixels kept on the device: each added frame's signal pixels are
 * regrouped by component in device memory, only components and frame links
 * come back to the host stitch, and the spots' pixel lists are written on
 * the device and copied back as spots close.  Spots, pixels and counts are
 * identical to sf_labeller_create's.  It takes frames from workspaces,
 * batches and sf_labeller_add (which labels on the device); not
 * sf_labeller_add_labelled.  Device memory held: the staged batches whose
 * frames an open component still touches. */
int sf_labeller_create_device(size_t height, size_t width,
                              const sf_grouping *g, sf_labeller **out);
/* Device bytes the staged frames of a sf_labeller_create_device labeller
 * still hold (0 once finished, and for a host labeller).  No counterpart in
 * the reference, whose pixels are all on the host. */
size_t sf_labeller_device_bytes(const sf_labeller *l);
void sf_labeller_destroy(sf_labeller *l);
int sf_labeller_add(sf_labeller *l, int64_t frame,
                    const sf_signal_pixel *signal, size_t n);
int sf_labeller_add_ws(sf_labeller *l, int64_t frame, sf_workspace *ws);
/* Every frame of a labelled batch, numbered first_frame, first_frame+1, ... */
int sf_labeller_add_batch(sf_labeller *l, int64_t first_frame, sf_batch *b);
/* A frame labelled elsewhere (a batched launch, another device): its
 * ordered list, each pixel's component (numbered in order of first pixel),
 * the components, and (here, previous-frame) component pairs, honoured when
 * links_to_previous is set and the frame follows the last one added. */
int sf_labeller_add_labelled(sf_labeller *l, int64_t frame,
                             const sf_signal_pixel *signal, size_t n,
                             const uint32_t *comp_of,
                             const sf_component *comps, size_t ncomp,
                             const uint32_t *links, size_t nlinks,
                             int links_to_previous);
int sf_labeller_finish(sf_labeller *l);
size_t sf_labeller_spot_count(const sf_labeller *l);
const sf_spot *sf_labeller_spots(const sf_labeller *l);
size_t sf_labeller_pixel_count(const sf_labeller *l);
const sf_pixel *sf_labeller_pixels(const sf_labeller *l);
sf_counts sf_labeller_counts(const sf_labeller *l);
size_t sf_labeller_live(const sf_labeller *l);

/* ------------------------------------------------------------------------
 * Serving: frames from several concurrent sweeps, as they arrive.  No
 * counterpart in the reference, whose mxi_find runs one sweep from files
 * (find_spots.cc:732-891); this is the same per-frame work (the threshold,
 * the ordered list, 2D components linked to the sweep's previous frame)
 * behind a request loop.  A worker thread takes every frame that has
 * arrived (at most ring - 1 per sweep, counting the batch still on the
 * device) into one batch -- one fused launch for all of them -- and enqueues
 * it behind the batch before, whose results it then reads while the new one
 * runs; later arrivals wait for the next: continuous batching, two batches
 * in flight.  Each sweep keeps its own ring of frame slots, so its label
 * scratch stays resident and a frame links to its sweep's previous frame.
 * A sweep is admitted only while free HBM stays above the reserve.  Frames
 * come resident in HBM (sf_server_submit) or as compressed chunks in host
 * memory (sf_server_submit_chunk, declared after sf_chunk below).
 * ------------------------------------------------------------------------ */
typedef struct sf_server sf_server;

typedef struct sf_server_config {
  uint32_t max_batch; /* frames per launch, all sweeps together (0: 64) */
  uint32_t ring;      /* slots per sweep, >= 2 (0: 4) */
  int32_t label;      /* 1: label every frame (2D, linked to its sweep) */
  uint64_t reserve_bytes; /* HBM to leave free when admitting a sweep */
} sf_server_config;

typedef struct sf_server_result {
  uint64_t request;
  int32_t sweep;
  int32_t status;          /* 0, -2 (more signal than the list holds), -1 (a
                              chunk's compressed data malformed) or -3 */
  uint64_t signal_pixels;
  uint64_t components;     /* 0 unless labelling */
  uint64_t links;          /* frame links to the sweep's previous frame */
  double latency_ms;       /* submission to results ready */
  uint32_t batch_frames;   /* frames in the launch that served it */
  uint32_t reserved;
} sf_server_result;

typedef struct sf_server_stats {
  uint64_t frames;         /* completed since the last reset */
  uint64_t batches;
  double p50_ms, p99_ms, max_ms, mean_ms; /* latency, submission to ready */
  double frames_per_s;     /* completed frames over first submission .. last completion */
  double mean_batch;
} sf_server_stats;

void sf_server_config_default(sf_server_config *c);
int sf_server_create(const sf_server_config *c, sf_server **out);
void sf_server_destroy(sf_server *s);
/* -6 when admitting the sweep would leave less than the reserve free. */
int sf_server_open_sweep(sf_server *s, uint32_t height, uint32_t width,
                         int bits, int *sweep);
/* Queues a device-resident frame of the sweep's shape, stamped with the
 * time of the call.  With host_out, its ordered signal list (at most
 * host_capacity entries) is copied there before the request completes. */
int sf_server_submit(sf_server *s, int sweep, const void *device_frame,
                     sf_signal_pixel *host_out, size_t host_capacity,
                     uint64_t *request);
/* Waits for one request. */
int sf_server_wait(sf_server *s, uint64_t request, sf_server_result *result);
int sf_server_stats_get(sf_server *s, sf_server_stats *stats);
void sf_server_stats_reset(sf_server *s);
/* A load driver for measurement: submits frame i of `frames` for sweep
 * sweeps[i] at arrival_s[i] seconds after the call (ascending), from a
 * thread of its own, then waits for all; results in submission order. */
int sf_server_replay(sf_server *s, const int *sweeps,
                     const void *const *frames, const double *arrival_s,
                     size_t n, sf_server_result *results);

/* ------------------------------------------------------------------------
 * Decompression on the device: decompress::image (src/spots/decompress.hh:
 * 22-30, decompress.cc:76-133) with the frame left in HBM, so that only the
 * compressed chunk crosses PCIe.  The chunk is the bytes exactly as HDF5
 * stored them, filter header and all, in host memory; `algorithm` is what the
 * dataset's filter pipeline names (nxmx.cc:163-179): "bslz4" (bitshuffle +
 * LZ4), "lz4", or "" / NULL for a chunk stored uncompressed.  A chunk that
 * does not describe its frame fails with -1 and the reference's message in
 * sf_last_error().  The device decoder's limits: bslz4 elements of 1, 2 or 4
 * bytes, blocks of at most 65536 bytes (bitshuffle's default is 8192), frames
 * and chunks below 4 GiB, device frames 8-byte aligned.
 * ------------------------------------------------------------------------ */

/* decompress::frame_bytes (decompress.hh:18-20). */
size_t sf_frame_bytes(size_t height, size_t width, unsigned bit_depth);

typedef struct sf_chunk {
  const uint8_t *data;   /* host memory, `size` bytes */
  size_t size;
  const char *algorithm; /* "bslz4", "lz4", "" or NULL */
  uint32_t bit_depth;    /* 8, 16 or 32 */
  uint32_t height, width;
  void *frame;           /* device memory, sf_frame_bytes() long */
} sf_chunk;

/* decompress::image for one chunk, on the calling thread's decoder; returns
 * once the frame is in device memory (device_frame_bytes long).  0, -1 (the
 * chunk does not describe the frame, or its data are malformed), -3. */
int sf_decompress_image(const uint8_t *data, size_t size, const char *algorithm,
                        unsigned bit_depth, uint32_t height, uint32_t width,
                        void *device_frame, size_t device_frame_bytes);

/* Serving from compressed chunks: queues one chunk of a frame of the sweep's
 * shape, in host memory exactly as HDF5 stored it, stamped with the time of
 * the call.  The worker uploads a batch's chunks in one copy and decodes them
 * on its stream, into frame buffers of the sweep's own (one per slot, made on
 * its first chunk), ahead of the threshold: only the compressed bytes cross
 * PCIe.  The bytes must stay valid until the request completes.  A chunk
 * that does not describe its frame is refused here: -1 with the reference's
 * message (decompress.cc:76-133); one whose compressed data are malformed
 * completes with status -1, and sf_server_wait returns -1 with the message. */
int sf_server_submit_chunk(sf_server *s, int sweep, const uint8_t *data,
                           size_t size, const char *algorithm,
                           sf_signal_pixel *host_out, size_t host_capacity,
                           uint64_t *request);
/* sf_server_replay for compressed chunks: chunk i (its data, size and
 * algorithm; the other fields are the sweep's) for sweep sweeps[i]. */
int sf_server_replay_chunks(sf_server *s, const int *sweeps, const sf_chunk *chunks,
                            const double *arrival_s, size_t n,
                            sf_server_result *results);

/* A decoder for many chunks at once: its stream (the caller's, or NULL for
 * its own), pinned staging and device buffers, grow-only.  No counterpart in
 * the reference, which decodes a chunk at a time on a host thread
 * (find_spots.cc:766-780).  load() checks every chunk against its frame
 * (all or nothing), walks the bslz4 block lengths on the host and enqueues
 * one upload; run() enqueues the decoding of what the last load holds (it may
 * be run again: the chunks stay resident); sync() waits and reports each
 * chunk (status[i] 0 or -1; the first failure's message in sf_last_error()). */
typedef struct sf_decoder sf_decoder;
sf_decoder *sf_decoder_create(void *stream);
void sf_decoder_destroy(sf_decoder *d);
void *sf_decoder_stream(sf_decoder *d);
int sf_decoder_load(sf_decoder *d, const sf_chunk *chunks, size_t n);
int sf_decoder_run(sf_decoder *d);
int sf_decoder_sync(sf_decoder *d, int *status);
/* With timing on, the device time of the last run (its kernels and copies). */
void sf_decoder_set_timing(sf_decoder *d, int on);
int sf_decoder_kernel_ms(sf_decoder *d, double *ms);

/* ------------------------------------------------------------------------
 * The spots as a DIALS reflection table: refl::write (src/spots/refl.hh:20-78,
 * refl.cc:223-351), byte for byte.  One msgpack document of the columns
 * dials.find_spots writes (bbox, flags, id, intensity.sum.value and
 * .variance, n_signal, panel, optionally shoebox, xyzobs.px.value and
 * .variance).  `spots` and `pixels` are the grouping's output and `width` the
 * frame width it was given.  -1 with sf_last_error() when the file cannot be
 * written or the spots are inconsistent; sf_refl_encode returns -2 when the
 * table (*size bytes) does not fit in `capacity`.
 * ------------------------------------------------------------------------ */
typedef struct sf_refl_options {
  uint64_t panel;          /* one panel (refl.hh: Options::panel) */
  int32_t id;              /* the experiment every spot belongs to */
  int32_t shoeboxes;       /* 1: write the shoebox column */
  const char *identifier;  /* the experiment identifier, or NULL / "" */
} sf_refl_options;

void sf_refl_options_default(sf_refl_options *o); /* 0, 0, shoeboxes, none */
int sf_refl_write(const char *path, const sf_spot *spots, size_t nspots,
                  const sf_pixel *pixels, size_t npixels, size_t width,
                  const sf_refl_options *options);
int sf_refl_encode(const sf_spot *spots, size_t nspots, const sf_pixel *pixels,
                   size_t npixels, size_t width, const sf_refl_options *options,
                   uint8_t *out, size_t capacity, size_t *size);
/* A finished grouping's spots, written as above with its own frame width. */
int sf_labeller_write_refl(const sf_labeller *l, const char *path,
                           const sf_refl_options *options);

/* ------------------------------------------------------------------------
 * --gpu-force and mxi_max on frames in HBM.  narrow_to_16
 * (src/spots/find_spots.cc:162-182): a 32-bit frame's pixels as their low 16
 * bits, so that the markers 0xfffffffe (bad pixel) and 0xffffffff (tile join)
 * become 0xfffe and 0xffff; it fits when no other pixel reaches 0xFFFD.
 * scan<T> (apps/mxi_max.cc:45-68): the largest count that is not a marker,
 * and how many pixels are each marker or a count at or above 0xFFFD.
 * ------------------------------------------------------------------------ */
typedef struct sf_narrow_job {
  const void *frame; /* device memory, `pixels` pixels of `bits`, 16-byte aligned */
  void *out16;       /* device memory for the 16-bit frame (32-bit frames), or
                        NULL to scan only (the only choice for 16-bit frames) */
  uint64_t pixels;
  uint32_t bits;     /* 16 or 32 */
  uint32_t reserved;
} sf_narrow_job;

typedef struct sf_frame_scan {
  uint64_t largest;  /* the largest count that is not a marker */
  uint64_t bad, joins, over; /* pixels at max() - 1, at max(), counts >= 0xFFFD */
  int32_t fits;      /* largest < 0xFFFD: narrow_to_16's answer */
  int32_t reserved;
} sf_frame_scan;

/* All jobs in one launch on the calling thread's stream; returns when the
 * results are in (device_ms, if not NULL: the launch's device time).
 * 0, -1 (bad argument), -3. */
int sf_narrow_frames(const sf_narrow_job *jobs, size_t n, sf_frame_scan *results,
                     double *device_ms);

#ifdef __cplusplus
}
#endif

#endif
