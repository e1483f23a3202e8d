// The three-dimensional grouping on the device: the reference's
// dials_spots::Labeller add/flush/emit (src/spots/dials_spots.cc:57-360)
// over components that the device labelling has already reduced, with the
// union-find, the closing order, the spot records and the spot pixel list
// all in HBM.  The host sees counts once per batch and copies the spots and
// pixels when asked for them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "engine.hh"
#include "sf_internal.hh"
#include "sf_labeller.hh"

namespace sfi {

// Records and their count.
template <typename T> struct Records {
  const T *data;
  size_t size;
};

class DeviceStitch {
public:
  DeviceStitch(size_t height, size_t width, const sfl::Options &options);
  ~DeviceStitch();
  DeviceStitch(const DeviceStitch &) = delete;
  DeviceStitch &operator=(const DeviceStitch &) = delete;

  // Frames first_frame .. first_frame + count - 1 are engine slots ids[i],
  // labelled, on the engine's stream; linked[i] says whether frame i's links
  // were taken against frame i - 1 (for i = 0: the previous frame added).
  // Frames must ascend; a gap closes every open spot first.
  void add(Engine &e, const int *ids, size_t count, int64_t first_frame,
           const bool *linked);
  // Closes every open spot (dials_spots.cc:147-151).
  void finish();

  bool started() const { return started_; }
  int64_t last_frame() const { return last_frame_; }

  // The records emitted so far, in host memory: each batch's are copied back
  // while the next batch runs, and these wait for the copies.  Valid until
  // the next add or finish.
  Records<sf_spot> spots();
  Records<sf_pixel> pixels();
  sf_counts counts() const { return counts_; }
  size_t live_pixels() const { return live_; }
  // Device bytes of the frames' staged pixels an open spot still needs.
  size_t held_bytes() const;

private:
  class Drain;
  struct Staged {
    std::unique_ptr<DeviceBuffer> buf;
    uint32_t id = 0;
  };
  void close_all();
  // Claims the roots of the open components (the previous last frame's
  // nodes) as candidates of the next process(); before any union.
  void collect(size_t open_begin, size_t open_end);
  // The batch's kernels after its nodes are in the table: union, merge,
  // close, emit, gather; `flush_all` closes every root among the candidates.
  void process(int first_rel, int nframes, bool flush_all, size_t new_begin,
               size_t new_end, size_t open_begin, size_t open_end,
               size_t last_begin, size_t last_end);
  void grow_table(size_t need);

  hipStream_t stream_ = nullptr;
  size_t height_, width_;
  sfl::Options options_;
  bool started_ = false;
  int64_t last_frame_ = 0;
  int64_t base_ = 0;        // frame that relative frame 0 stands for
  size_t table_end_ = 0;    // nodes in the table
  size_t table_cap_ = 0;
  size_t pl_begin_ = 0, pl_end_ = 0; // the last frame's nodes
  size_t dead_ = 0;         // nodes of closed components still in the table
  uint32_t epoch_ = 0;
  uint32_t next_buffer_ = 0;

  DeviceBuffer nodes_, acc_, stamp_open_, stamp_claim_, nodes_tmp_, acc_tmp_;
  DeviceBuffer keybits_, keyprefix_, keytotals_, emit_list_, cand_, spot_tmp_;
  DeviceBuffer decision_, accidx_, pixoff_, scan_tmp_, summary_, counters_;
  DeviceBuffer membits_, memprefix_, memtotals_, mkeys_, mvals_, mkeys2_, mvals2_;
  DeviceBuffer sort_hist_, spot_members_, spot_pix_, member_pix_, live_counts_;
  DeviceBuffer frame_desc_, comp_prefix_, alive_bits_, alive_prefix_, alive_totals_;
  DeviceBuffer spots_out_[2], pixels_out_[2], error_, scan_desc_;
  Pinned host_desc_, host_summary_, host_scan_;
  std::vector<Staged> staged_;
  std::vector<std::unique_ptr<DeviceBuffer>> spare_;

  size_t spots_dev_ = 0, pixels_dev_ = 0; // records emitted (and copies queued)
  uint64_t batches_ = 0;                   // process() calls: the output set's parity
  uint64_t out_ticket_[2] = {0, 0};        // the drain's job on each output set
  hipEvent_t out_ready_[2] = {nullptr, nullptr};
  std::unique_ptr<Drain> drain_;
  sf_counts counts_{};
  size_t live_ = 0;
};

} // namespace sfi
