// The host side of the grouping: components labelled on the device, joined
// across frames and reduced to spots as dials_spots::Labeller does
// (src/spots/dials_spots.hh:108-158).  Not part of the public ABI; the C
// functions sf_labeller_* in include/spotfinder_hip.h wrap it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <vector>

#include "spotfinder_hip.h"

namespace sfl {

// A growable array of plain records in an anonymous mapping advised for
// huge pages, grown by remapping (no copy).  A sweep's spot pixels reach
// hundreds of megabytes: grown in 4 KB pages, and copied at every doubling,
// that memory cost more than the grouping itself.
template <typename T> class BigArray {
public:
  BigArray() = default;
  BigArray(const BigArray &) = delete;
  BigArray &operator=(const BigArray &) = delete;
  ~BigArray();
  void push_back(const T &v) {
    if (size_ == cap_)
      grow(size_ + 1);
    data_[size_++] = v;
  }
  void reserve(size_t n) {
    if (n > cap_)
      grow(n);
  }
  // Grows by `count` records left for the caller to write; their address.
  T *extend(size_t count) {
    reserve(size_ + count);
    size_ += count;
    return data_ + size_ - count;
  }
  const T *data() const { return data_; }
  size_t size() const { return size_; }
  size_t capacity() const { return cap_; }

private:
  void grow(size_t need);
  T *data_ = nullptr;
  size_t size_ = 0, cap_ = 0, bytes_ = 0;
};

// dials_spots::Options (src/spots/dials_spots.hh:81-95).
struct Options {
  size_t min_spot_size = 3;
  size_t max_spot_size = 1000;
  double max_separation = 2.0;
  bool two_d = false;
};

class Labeller {
public:
  Labeller(size_t height, size_t width, const Options &options);

  // Validates a frame's list as Labeller::add does before grouping it
  // (dials_spots.cc:57-98): throws std::runtime_error.
  void check_frame(int64_t frame, const sf_signal_pixel *signal,
                   size_t n) const;

  // One device-labelled frame: the ordered signal list, each pixel's
  // component, the components, and (frame-1 components) links.  The links
  // are honoured only when `links_to_previous` says they were taken against
  // the frame added last.
  void add(int64_t frame, const sf_signal_pixel *signal, size_t n,
           const uint32_t *comp_of, const sf_component *comps, size_t ncomp,
           const uint32_t *links, size_t nlinks, bool links_to_previous);
  void finish();

  bool started() const { return started_; }
  int64_t last_frame() const { return last_frame_; }
  const BigArray<sf_spot> &spots() const { return spots_; }
  const BigArray<sf_pixel> &pixels() const { return kept_; }
  const sf_counts &counts() const { return counts_; }
  size_t live() const { return live_pixels_; }
  size_t height() const { return height_; }
  size_t width() const { return width_; }
  const Options &options() const { return options_; }

private:
  // A node's component record is comps_[i]: the passes over nodes read
  // these small records, and only emit reads the moments.
  struct Node {
    int64_t frame;
    uint32_t begin, end;   // pixel range in the frame's arena
    bool retained = false; // the range is in retained_ instead
    bool dead = false;     // its component was emitted or dropped
  };
  struct Arena {
    std::vector<uint32_t> index, value;
    size_t live = 0;             // pixels of live nodes still held here
    size_t first = 0, last = 0;  // the frame's nodes, [first, last)
  };

  uint32_t find(uint32_t i);
  void merge(uint32_t a, uint32_t b);
  void flush(int64_t keep_frame);
  void emit(const uint32_t *members, size_t count);
  static void move_pixels(const Arena &from, Node &node, Arena &to);
  void release(std::map<int64_t, Arena>::iterator it);
  void begin_frame(int64_t frame);
  // Throws for a frame link that names a component outside this frame or
  // the previous one; changes nothing.
  void check_links(int64_t frame, size_t ncomp, const uint32_t *links,
                   size_t nlinks, bool links_to_previous) const;
  void add_nodes(int64_t frame, Arena &arena, const std::vector<uint32_t> &start,
                 size_t n, const sf_component *comps, size_t ncomp,
                 const uint32_t *links, size_t nlinks, bool links_to_previous);
  void compact();

  size_t height_, width_;
  Options options_;
  // Nodes in (frame, first index) order; closed ones stay, marked dead, until
  // they outnumber the live ones.  Union-find keeps each component's lowest
  // node as its root; next_ chains a component's nodes from its root, and
  // tail_ (at the root) ends the chain.
  std::vector<Node> nodes_;
  std::vector<sf_component> comps_;
  std::vector<uint32_t> parent_, next_, tail_;
  size_t live_nodes_ = 0;
  size_t previous_ = 0;          // where the last frame's nodes start
  size_t retained_live_ = 0;     // live pixels in retained_
  std::map<int64_t, Arena> arenas_;
  int64_t last_frame_ = 0;
  bool started_ = false;
  size_t live_pixels_ = 0;
  BigArray<sf_spot> spots_;
  BigArray<sf_pixel> kept_;
  std::vector<Arena> spare_arenas_; // released frames' buffers, for reuse
  // Pixels of long-lived components whose frames' arenas were released.
  Arena retained_;
  // Scratch reused from frame to frame: add's counting sort, flush's
  // labels, members and marks, emit's merge cursors.
  std::vector<uint32_t> start_, at_, roots_, member_, remap_;
  std::vector<uint8_t> mark_;
  std::vector<size_t> pos_;
  sf_counts counts_{};
};

} // namespace sfl
