// The device engine behind every entry point: frame slots with their
// buffers, one stream, and the enqueue of a batch of frames -- the fused
// threshold over all of them in one persistent launch, the ordering of each
// frame's signal pixels, and the labelling -- in a fixed number of launches
// whatever the batch size, optionally captured once into a hipGraph and
// replayed.  sf_workspace is an engine of two slots used one frame at a
// time; sf_batch is an engine of two sets of slots, one batch in flight while
// the previous one is read.
#pragma once

#include <map>
#include <memory>
#include <vector>

#include "sf_internal.hh"

namespace sfi {

// One frame's buffers, kept between calls and grown to the largest frame.
struct FrameSlot {
  DeviceBuffer mask, prefix, totals, unordered, ordered;
  DeviceBuffer parent, roots, root_prefix, comp_of, comps, links;
  DeviceBuffer row_runs; // runs per bitmask row, then each row's first run
  DeviceBuffer comp_runs; // runs per component, then each component's first
  uint32_t height = 0, width = 0, wpr = 0, capacity = 0;
  int bits = 0;
  size_t words = 0, root_words = 0;
  bool thresholded = false, labelled = false;
  uint64_t serial = 0;  // which frame of the engine it holds
  uint64_t linked_serial = 0; // the frame its links were taken against, or 0
  int linked_slot = -1;       // and that frame's slot
  const void *frame = nullptr;
  int kernel = 0;
  int lane_columns = 4;       // of the fused plan it was last enqueued under
  uint32_t link_capacity = 0;
};

// A frame to threshold: a device pointer and its shape.
struct FrameIn {
  const void *frame;
  uint32_t height, width;
  int bits;
};

class Engine {
public:
  explicit Engine(void *borrowed_stream);
  ~Engine();
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;

  hipStream_t stream() const { return stream_; }
  FrameSlot &slot(int i) { return *slots_[static_cast<size_t>(i)]; }
  size_t slot_count() const { return slots_.size(); }
  uint64_t fallbacks() const { return fallbacks_; }

  // Enqueues the threshold of frames[i] into slot ids[i] and the ordering of
  // each; returns at once.  `graph` captures the enqueue into a hipGraph the
  // first time a batch of these shapes and slots is seen, and replays it
  // after, with the new frame pointers.
  void threshold(const FrameIn *frames, const int *ids, size_t n, int kernel,
                 bool graph);
  // Waits; the count of each frame's signal pixels and 0 or -2.  A frame the
  // fused kernel handed back (a valid pixel beyond its exact range) is
  // thresholded again by the direct kernels here.
  void sync(const int *ids, size_t n, size_t *counts, int *status);
  void fetch(int id, sf_signal_pixel *out, size_t n);

  // Batches in flight two at a time: mark() records on the stream the end of
  // the work enqueued so far; sync_after() waits for such a mark and reads
  // the counters of slots ids[i] on a read stream of the engine's own, so a
  // batch enqueued behind the mark keeps running.  A frame the fused kernel
  // handed back is re-run as in sync() (which then waits for everything);
  // the frames of the batch behind (`later`, linked to `later_prev`) are
  // then labelled again, since their links may read a re-run frame.
  void mark(hipEvent_t e);
  void sync_after(hipEvent_t done, const int *ids, size_t n, size_t *counts,
                  int *status, const int *later, const int *later_prev,
                  size_t nlater);
  // A slot's ordered list to host memory once `done` has fired, on the read
  // stream; returns when it is there.
  void fetch_after(hipEvent_t done, int id, sf_signal_pixel *out, size_t n);
  hipStream_t read_stream();

  // Loads an ascending host list into a slot as if thresholded.
  void load_signal(int id, const sf_signal_pixel *list, size_t n,
                   uint32_t height, uint32_t width);

  // Enqueues the labelling of slots ids[i], each linked to slot prev[i]
  // (-1 for none; the caller checks that it holds the frame before, of the
  // same shape, labelled).
  void label(const int *ids, const int *prev, size_t n);
  // Waits; pixel, component and link counts of one slot.
  // `fresh` false reuses the counters the last sync() read (after a
  // labelling enqueued before it).
  void label_counts(int id, size_t *n, size_t *ncomp, size_t *nlinks,
                    bool fresh = true);
  // Device bytes a slot of this shape holds once labelled.
  static size_t slot_bytes(uint32_t height, uint32_t width, int bits);
  // Device bytes the slots hold now (their buffers only grow).
  size_t device_bytes() const;
  // Copies a labelled slot's labels, components, frame links and, when
  // `list` is given, its signal pixels (n of each), and waits unless `wait`
  // is false (the caller then waits on stream()).
  void fetch_labels(int id, uint32_t *comp_of, size_t n, sf_component *comps,
                    size_t ncomp, uint32_t *links, size_t nlinks,
                    sf_signal_pixel *list = nullptr, bool wait = true);

  // Events around the threshold kernel(s) and the ordering of the last batch.
  void set_timing(bool on) { timing_on_ = on; }
  bool timed() const { return timed_; }
  void kernel_ms(double *threshold_ms, double *order_ms);

private:
  // Where a batch's threshold descriptors sit in desc_, and its totals.
  struct Prepared {
    size_t b_fused = 0, b_frames = 0, b_segs = 0;
    uint32_t items = 0;
    size_t max_words = 0;
    int lane_columns = 4, waves_per_block = 4; // the launch shape it was planned for
    uint32_t seg_rows = 0;
  };
  void ensure_slots(size_t n);
  void grow(DeviceBuffer &b, size_t bytes);
  void shape(FrameSlot &s, int bits, uint32_t h, uint32_t w);
  void shape_label(FrameSlot &s);
  FrameDev dev(int id, int prev);
  ScanSeg mask_seg(int id);
  FrameArgs args(int id, const void *frame);
  // Pinned host staging for descriptors, two buffers used in turn, each
  // reused only once the copy that last read it has run.
  char *stage(size_t bytes);
  void staged(); // records that the current staging buffer's copy is queued
  Prepared prepare(const FrameIn *frames, const int *ids, size_t n, int kernel);
  void enqueue_kernels(const Prepared &p, const FrameIn *frames,
                       const int *ids, size_t n, int kernel, bool timing);
  void read_counters();

  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  hipStream_t read_stream_ = nullptr; // made on first use
  std::vector<std::unique_ptr<FrameSlot>> slots_;
  DeviceBuffer counters_, table_, desc_;
  DeviceBuffer queue_; // the fused kernel's item queue, one word
  uint32_t *host_counters_ = nullptr;
  size_t host_counter_words_ = 0;
  char *staging_[2] = {nullptr, nullptr};
  size_t staging_bytes_[2] = {0, 0};
  hipEvent_t staging_done_[2] = {nullptr, nullptr};
  int turn_ = 0;
  DeviceBuffer scr0_, scr1_;
  hipEvent_t marks_[4] = {};
  uint64_t serial_ = 0;
  uint64_t fallbacks_ = 0;
  uint64_t epoch_ = 0; // bumped whenever a buffer moves
  bool timing_on_ = false, timed_ = false;

  // The captured batch, and what it was captured for.
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<uint64_t> graph_key_;

  // The labelling sequence, captured once per frame count: its kernels take
  // only the descriptor block, at a fixed address, and the frame count.
  struct LabelGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::map<uint32_t, LabelGraph> label_graphs_;
  void drop_label_graphs();
};

} // namespace sfi
