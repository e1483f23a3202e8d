// Grouping of device-labelled frames into spots: the host side of the
// reference's dials_spots::Labeller (src/spots/dials_spots.cc:36-360).
//
// The device hands over, per frame, the ordered signal pixels, each pixel's
// four-connected component (numbered in order of first pixel) and each
// component's exact integer moments, plus the pairs of components that share
// a pixel position with the previous frame.  What is left is the part of
// Labeller that works above the pixel level:
//   * components of consecutive frames joined by a shared position are one
//     three-dimensional component (six-connected, dials_spots.cc:128-137);
//     a gap in frame numbers closes everything (:68-69);
//   * after each frame, every component with nothing on the newest frame is
//     emitted, in order of its first (frame, index) pixel (:153-219) -- in two
//     dimensions every component of the frame at once;
//   * emit applies the size filters, sums the moments into the bounding box,
//     intensity, centroid at pixel centres, variance and peak, and applies the
//     peak-centroid separation filter (:221-360).
// Node-level union-find mirrors the reference's pixel-level one: nodes are
// held in (frame, first index) order and merged toward the earlier, so labels
// in order of first appearance are the reference's labels.
//
// Exactness.  Intensity, centroid and the filters come out identical to the
// reference wherever its double sums are exact (every partial sum below 2^53,
// true for counts below 2^32 on spots of a few thousand pixels): the moments
// here are exact integers and each quantity is formed with the reference's
// own final operations.  The variance's sum of value * delta^2 is formed from
// the moments (N = S Q - W^2, exact in 128 bits) rather than summed pixel by
// pixel, so it can differ from the reference's rounding in the last bits.
#include <sys/mman.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "sf_labeller.hh"

namespace sfl {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr double kPixelVariance = 1.0 / 12.0;
// An earlier frame's arena is released, its kept pixels moved, once fewer
// than one pixel in kRetainRatio is still read; the retained store is
// rewritten once it holds more than twice its live pixels plus kRetainSlack.
constexpr size_t kRetainRatio = 8;
constexpr size_t kRetainSlack = 1024;
// Closed components' nodes are dropped once they outnumber the live ones by
// kDeadSlack.
constexpr size_t kDeadSlack = 1024;
using u128 = unsigned __int128;

u128 pair128(const uint64_t p[2]) {
  return (static_cast<u128>(p[1]) << 64) | p[0];
}

// a * b, exactly, or false when it overflows 128 bits.  Operands below 2^64,
// the usual case, multiply without the overflow test.
bool mul_exact(u128 a, u128 b, u128 *out) {
  if ((a >> 64) == 0 && (b >> 64) == 0) {
    *out = static_cast<u128>(static_cast<uint64_t>(a)) * static_cast<uint64_t>(b);
    return true;
  }
  return !__builtin_mul_overflow(a, b, out);
}

double to_double(u128 v) {
  // Correctly rounded: split so that each part converts exactly, then one
  // rounding in the addition when the value is below 2^106.
  const uint64_t hi = static_cast<uint64_t>(v >> 64);
  const uint64_t lo = static_cast<uint64_t>(v);
  if (hi == 0)
    return static_cast<double>(lo);
  // Scaling by 2^64 is exact (a power of two, far from overflow).
  return static_cast<double>(hi) * 0x1p64 + static_cast<double>(lo);
}

} // namespace

template <typename T> void BigArray<T>::grow(size_t need) {
  constexpr size_t kHuge = 2u << 20;
  size_t cap = std::max<size_t>({need, 2 * cap_, kHuge / sizeof(T)});
  const size_t bytes = (cap * sizeof(T) + kHuge - 1) / kHuge * kHuge;
  void *next;
  if (data_) // the pages move with their contents; nothing is copied
    next = mremap(data_, bytes_, bytes, MREMAP_MAYMOVE);
  else
    next = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS,
                -1, 0);
  if (next == MAP_FAILED)
    throw std::bad_alloc();
  (void)madvise(next, bytes, MADV_HUGEPAGE); // a hint: small pages otherwise
  data_ = static_cast<T *>(next);
  bytes_ = bytes;
  cap_ = bytes / sizeof(T);
}

template <typename T> BigArray<T>::~BigArray() {
  if (data_)
    (void)munmap(data_, bytes_);
}

template class BigArray<sf_pixel>;
template class BigArray<sf_spot>;

Labeller::Labeller(size_t height, size_t width, const Options &options)
    : height_(height), width_(width), options_(options) {
  if (height == 0 || width == 0)
    throw std::runtime_error("dials_spots: a frame with no pixels in it");
  if (options_.max_spot_size < options_.min_spot_size)
    throw std::runtime_error("dials_spots: max_spot_size is below "
                             "min_spot_size, so nothing can be accepted");
  if (options_.min_spot_size == 0)
    throw std::runtime_error("dials_spots: min_spot_size must be at least 1");
}

void Labeller::check_frame(int64_t frame, const sf_signal_pixel *signal,
                           size_t n) const {
  if (started_ && frame <= last_frame_)
    throw std::runtime_error("dials_spots: frame " + std::to_string(frame) +
                             " arrived after frame " +
                             std::to_string(last_frame_) +
                             "; frames must be added in ascending order");
  if (n == 0)
    return;
  // One pass without branches; the error, if any, is found afterwards. An
  // ascending list is in range when its last index is.
  const size_t pixels = height_ * width_;
  bool descends = false;
  for (size_t i = 1; i < n; i++)
    descends |= signal[i].index <= signal[i - 1].index;
  if (descends) {
    for (size_t i = 0; i < n; i++)
      if (signal[i].index >= pixels)
        throw std::runtime_error("dials_spots: signal pixel " +
                                 std::to_string(signal[i].index) +
                                 " is off the end of the frame");
    throw std::runtime_error(
        "dials_spots: signal pixels are not ascending by index");
  }
  if (signal[n - 1].index >= pixels)
    throw std::runtime_error("dials_spots: signal pixel " +
                             std::to_string(signal[n - 1].index) +
                             " is off the end of the frame");
}

uint32_t Labeller::find(uint32_t i) {
  while (parent_[i] != i) {
    parent_[i] = parent_[parent_[i]];
    i = parent_[i];
  }
  return i;
}

void Labeller::merge(uint32_t a, uint32_t b) {
  const uint32_t ra = find(a), rb = find(b);
  if (ra == rb)
    return;
  // The lower node stays the root; its chain takes the other's.
  const uint32_t keep = ra < rb ? ra : rb, join = ra < rb ? rb : ra;
  parent_[join] = keep;
  next_[tail_[keep]] = join;
  tail_[keep] = tail_[join];
}

void Labeller::check_links(int64_t frame, size_t ncomp, const uint32_t *links,
                           size_t nlinks, bool links_to_previous) const {
  if (options_.two_d || !started_ || last_frame_ + 1 != frame ||
      !links_to_previous)
    return;
  const size_t prev_count = nodes_.size() - previous_;
  for (size_t l = 0; l < nlinks; l++)
    if (links[2 * l] >= ncomp || links[2 * l + 1] >= prev_count)
      throw std::runtime_error("device labelling: a frame link out of range");
}

void Labeller::add(int64_t frame, const sf_signal_pixel *signal, size_t n,
                   const uint32_t *comp_of, const sf_component *comps,
                   size_t ncomp, const uint32_t *links, size_t nlinks,
                   bool links_to_previous) {
  // Every check before any state changes, so that a refused frame leaves
  // the spots, pixels and counts as they were and the next frame groups as
  // if it had never been offered.
  check_frame(frame, signal, n);
  std::vector<uint32_t> &start = start_;
  start.assign(ncomp + 1, 0);
  uint32_t top = 0;
  for (size_t i = 0; i < n; i++)
    top = std::max(top, comp_of[i]);
  if (n && top >= ncomp)
    throw std::runtime_error("device labelling: component out of range");
  for (size_t i = 0; i < n; i++)
    start[comp_of[i] + 1]++;
  for (size_t c = 0; c < ncomp; c++)
    if (start[c + 1] != comps[c].n)
      throw std::runtime_error("device labelling: component sizes disagree");
  check_links(frame, ncomp, links, nlinks, links_to_previous);

  begin_frame(frame);
  // Pixels of this frame grouped by component, each group still ascending.
  Arena &arena = arenas_[frame];
  if (!spare_arenas_.empty()) { // a released frame's buffers, capacity kept
    arena = std::move(spare_arenas_.back());
    spare_arenas_.pop_back();
  }
  arena.index.resize(n);
  arena.value.resize(n);
  for (size_t c = 0; c < ncomp; c++)
    start[c + 1] += start[c];
  std::vector<uint32_t> &at = at_;
  at.assign(start.begin(), start.end() - 1);
  for (size_t i = 0; i < n; i++) {
    const uint32_t k = at[comp_of[i]]++;
    arena.index[k] = signal[i].index;
    arena.value[k] = signal[i].value;
  }
  add_nodes(frame, arena, start, n, comps, ncomp, links, nlinks,
            links_to_previous);
}
void Labeller::begin_frame(int64_t frame) {
  if (started_ && frame != last_frame_ + 1)
    flush(frame);
}
void Labeller::add_nodes(int64_t frame, Arena &arena,
                         const std::vector<uint32_t> &start, size_t n,
                         const sf_component *comps, size_t ncomp,
                         const uint32_t *links, size_t nlinks,
                         bool links_to_previous) {
  const size_t base = nodes_.size();
  const size_t previous = previous_;
  for (size_t c = 0; c < ncomp; c++) {
    Node node;
    node.frame = frame;
    node.begin = start[c];
    node.end = start[c + 1];
    const uint32_t self = static_cast<uint32_t>(nodes_.size());
    nodes_.push_back(node);
    comps_.push_back(comps[c]);
    parent_.push_back(self);
    next_.push_back(kNone);
    tail_.push_back(self);
  }
  arena.live = n;
  arena.first = base;
  arena.last = nodes_.size();
  live_nodes_ += ncomp;
  counts_.signal_pixels += n;
  live_pixels_ += n;

  // Six-connected: components sharing a position with the previous frame
  // (the links were checked against both frames before anything changed).
  if (!options_.two_d && started_ && last_frame_ + 1 == frame && links_to_previous) {
    for (size_t l = 0; l < nlinks; l++) {
      const uint32_t here = links[2 * l], there = links[2 * l + 1];
      merge(static_cast<uint32_t>(base + here),
            static_cast<uint32_t>(previous + there));
    }
  }
  started_ = true;
  last_frame_ = frame;
  flush(options_.two_d ? frame + 1 : frame);
}

void Labeller::finish() {
  if (!started_)
    return;
  flush(last_frame_ + 1);
}

void Labeller::flush(int64_t keep_frame) {
  // Every open component has a node on the last frame, so the nodes from
  // there on name every component that can close now: those with no node on
  // keep_frame.  Marks: 1 seen, 2 open.
  const size_t n = nodes_.size();
  mark_.resize(n, 0);
  std::vector<uint32_t> &roots = roots_;
  roots.clear();
  for (size_t i = previous_; i < n; i++) {
    if (nodes_[i].dead)
      continue;
    const uint32_t r = find(static_cast<uint32_t>(i));
    if (!mark_[r]) {
      mark_[r] = 1;
      roots.push_back(r);
    }
    if (nodes_[i].frame == keep_frame)
      mark_[r] = 2;
  }
  // Closing components in order of their first node, as the reference numbers
  // them (dials_spots.cc:158-170); each one's nodes in (frame, first) order.
  size_t closing = 0;
  for (const uint32_t r : roots) {
    if (mark_[r] != 2)
      roots[closing++] = r;
    mark_[r] = 0;
  }
  roots.resize(closing);
  std::sort(roots.begin(), roots.end());
  std::vector<uint32_t> &member = member_;
  for (const uint32_t r : roots) {
    member.clear();
    for (uint32_t k = r; k != kNone; k = next_[k])
      member.push_back(k);
    if (!std::is_sorted(member.begin(), member.end()))
      std::sort(member.begin(), member.end());
    emit(member.data(), member.size());
    for (const uint32_t k : member) {
      Node &node = nodes_[k];
      node.dead = true;
      live_pixels_ -= comps_[k].n;
      const size_t held = node.end - node.begin;
      if (node.retained) {
        retained_live_ -= held;
      } else {
        auto it = arenas_.find(node.frame);
        it->second.live -= held;
        if (it->second.live == 0)
          release(it);
      }
    }
    live_nodes_ -= member.size();
  }

  if (live_nodes_ == 0) {
    nodes_.clear();
    comps_.clear();
    parent_.clear();
    next_.clear();
    tail_.clear();
    mark_.clear();
    previous_ = 0;
    while (!arenas_.empty())
      release(arenas_.begin());
    retained_.index.clear();
    retained_.value.clear();
    retained_live_ = 0;
    return;
  }
  if (n - live_nodes_ > live_nodes_ + kDeadSlack)
    compact();
  previous_ = nodes_.size();
  while (previous_ > 0 && nodes_[previous_ - 1].frame == keep_frame)
    previous_--;

  // An earlier frame's arena that only a few long-lived components still
  // read (hot pixels and powder rings stay open for the whole sweep) hands
  // their pixels to the retained store and is released, so memory follows
  // the live pixels and not the sweep's length.  The retained store is
  // rewritten once it is mostly closed components'.
  for (auto it = arenas_.begin(); it != arenas_.end() && it->first < keep_frame;) {
    Arena &arena = it->second;
    if (arena.live > 0 && arena.live * kRetainRatio >= arena.index.size()) {
      ++it;
      continue;
    }
    for (size_t i = arena.first; i < arena.last; i++)
      if (!nodes_[i].dead) {
        move_pixels(arena, nodes_[i], retained_);
        retained_live_ += nodes_[i].end - nodes_[i].begin;
      }
    auto next = std::next(it);
    release(it);
    it = next;
  }
  if (retained_.index.size() > 2 * retained_live_ + kRetainSlack) {
    Arena next;
    next.index.reserve(2 * retained_live_ + kRetainSlack);
    next.value.reserve(2 * retained_live_ + kRetainSlack);
    for (Node &node : nodes_)
      if (!node.dead && node.retained)
        move_pixels(retained_, node, next);
    std::swap(retained_.index, next.index);
    std::swap(retained_.value, next.value);
  }
}
void Labeller::release(std::map<int64_t, Arena>::iterator it) {
  it->second.live = 0;
  spare_arenas_.push_back(std::move(it->second));
  arenas_.erase(it);
}
// Drops dead nodes, keeping order: roots stay each component's lowest node,
// and the chains are rebuilt in node order.
void Labeller::compact() {
  const size_t n = nodes_.size();
  std::vector<uint32_t> &remap = remap_, &root_of = roots_;
  // Every root before any entry moves: find() reads the old parents.
  root_of.assign(n, kNone);
  for (size_t i = 0; i < n; i++)
    if (!nodes_[i].dead)
      root_of[i] = find(static_cast<uint32_t>(i));
  remap.assign(n, kNone);
  size_t write = 0;
  for (size_t i = 0; i < n; i++) {
    if (nodes_[i].dead)
      continue;
    remap[i] = static_cast<uint32_t>(write);
    const uint32_t root = remap[root_of[i]]; // a root is its lowest node
    nodes_[write] = nodes_[i];
    comps_[write] = comps_[i];
    parent_[write] = root;
    next_[write] = kNone;
    if (root == write)
      tail_[write] = root;
    else
      next_[tail_[root]] = static_cast<uint32_t>(write), tail_[root] = static_cast<uint32_t>(write);
    write++;
  }
  nodes_.resize(write);
  comps_.resize(write);
  parent_.resize(write);
  next_.resize(write);
  tail_.resize(write);
  mark_.assign(write, 0);
  for (auto &a : arenas_) {
    Arena &arena = a.second;
    size_t first = kNone, last = 0;
    for (size_t i = arena.first; i < arena.last; i++)
      if (remap[i] != kNone) {
        first = std::min<size_t>(first, remap[i]);
        last = remap[i] + 1;
      }
    arena.first = first == kNone ? 0 : first;
    arena.last = last;
  }
}
void Labeller::move_pixels(const Arena &from, Node &node, Arena &to) {
  const size_t begin = to.index.size(), count = node.end - node.begin;
  if (begin + count > 0xFFFFFFFFu)
    throw std::runtime_error("dials_spots: more than 2^32 pixels held open");
  to.index.insert(to.index.end(), from.index.begin() + node.begin,
                  from.index.begin() + node.end);
  to.value.insert(to.value.end(), from.value.begin() + node.begin,
                  from.value.begin() + node.end);
  node.begin = static_cast<uint32_t>(begin);
  node.end = static_cast<uint32_t>(begin + count);
  node.retained = true;
}

void Labeller::emit(const uint32_t *members, size_t count) {
  counts_.groups++;
  uint64_t n = 0;
  for (size_t k = 0; k < count; k++)
    n += comps_[members[k]].n;
  if (n < options_.min_spot_size) {
    counts_.too_small++;
    return;
  }
  if (n > options_.max_spot_size) {
    counts_.too_large++;
    return;
  }

  sf_spot spot;
  std::memset(&spot, 0, sizeof(spot));
  spot.n_signal = static_cast<uint32_t>(n);
  u128 sum = 0, sum_sq = 0, wx = 0, wy = 0, wz = 0, qx = 0, qy = 0, qz = 0;
  int64_t x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
  uint32_t peak_value = 0;
  int64_t peak_frame = 0;
  uint32_t peak_index = 0;
  bool have_peak = false;
  for (size_t k = 0; k < count; k++) {
    const Node &node = nodes_[members[k]];
    const sf_component &c = comps_[members[k]];
    const int64_t z = node.frame;
    if (k == 0) {
      x0 = c.x0, x1 = c.x1, y0 = c.y0, y1 = c.y1, z0 = z, z1 = z + 1;
    } else {
      x0 = std::min<int64_t>(x0, c.x0);
      x1 = std::max<int64_t>(x1, c.x1);
      y0 = std::min<int64_t>(y0, c.y0);
      y1 = std::max<int64_t>(y1, c.y1);
      z0 = std::min(z0, z);
      z1 = std::max(z1, z + 1);
    }
    const u128 s = c.sum;
    const u128 az = static_cast<u128>(2 * z + 1);
    sum += s;
    sum_sq += pair128(c.sum_sq);
    wx += pair128(c.sum_x);
    wy += pair128(c.sum_y);
    wz += az * s;
    qx += pair128(c.sum_xx);
    qy += pair128(c.sum_yy);
    qz += az * az * s;
    // The brightest pixel, ties to the lowest (frame, index): members come in
    // (frame, first index) order, so a strictly-greater test across them
    // matches the reference's (:279-284) once each component's own peak is
    // its lowest-index brightest pixel.
    const uint32_t value = static_cast<uint32_t>(c.peak >> 32);
    const uint32_t index = 0xFFFFFFFFu - static_cast<uint32_t>(c.peak);
    if (!have_peak || value > peak_value ||
        (value == peak_value && (z < peak_frame || (z == peak_frame && index < peak_index)))) {
      have_peak = true;
      peak_value = value;
      peak_frame = z;
      peak_index = index;
    }
  }
  spot.bbox[0] = static_cast<int32_t>(x0);
  spot.bbox[1] = static_cast<int32_t>(x1);
  spot.bbox[2] = static_cast<int32_t>(y0);
  spot.bbox[3] = static_cast<int32_t>(y1);
  spot.bbox[4] = static_cast<int32_t>(z0);
  spot.bbox[5] = static_cast<int32_t>(z1);

  const double sum_d = to_double(sum);
  const double sum_sq_d = to_double(sum_sq);
  spot.intensity = sum_d;
  spot.intensity_variance = sum_d;
  const double peak_pos[3] = {
      static_cast<double>(peak_index % width_) + 0.5,
      static_cast<double>(peak_index / width_) + 0.5,
      static_cast<double>(peak_frame) + 0.5};
  if (sum_d > 0.0) {
    const u128 w[3] = {wx, wy, wz};
    const u128 q[3] = {qx, qy, qz};
    for (int a = 0; a < 3; a++)
      spot.position[a] = (to_double(w[a]) * 0.5) / sum_d;
    if (sum_d * sum_d > sum_sq_d) {
      for (int a = 0; a < 3; a++) {
        // sum of value * (pos - centroid)^2 = (S Q - W^2) / (4 S), with pos
        // and W at doubled centres.
        double delta_sq;
        u128 sq, ww;
        if (mul_exact(sum, q[a], &sq) && mul_exact(w[a], w[a], &ww)) {
          delta_sq = to_double(sq - ww) / (4.0 * sum_d);
        } else {
          const long double S = static_cast<long double>(sum);
          const long double W = static_cast<long double>(w[a]);
          const long double Q = static_cast<long double>(q[a]);
          delta_sq = static_cast<double>((S * Q - W * W) / (4.0L * S));
        }
        const double v = delta_sq * sum_d / (sum_d * sum_d - sum_sq_d);
        spot.variance[a] = v / sum_d + kPixelVariance;
      }
    } else {
      for (int a = 0; a < 3; a++)
        spot.variance[a] = kPixelVariance;
    }
    if (spot.bbox[5] == spot.bbox[4] + 1)
      spot.position[2] = static_cast<double>(spot.bbox[4]) + 0.5;
  } else {
    spot.position[0] = (spot.bbox[1] + spot.bbox[0]) / 2.0;
    spot.position[1] = (spot.bbox[3] + spot.bbox[2]) / 2.0;
    spot.position[2] = (spot.bbox[5] + spot.bbox[4]) / 2.0;
  }
  if (options_.max_separation > 0.0) {
    const double dx = peak_pos[0] - spot.position[0];
    const double dy = peak_pos[1] - spot.position[1];
    const double dz = peak_pos[2] - spot.position[2];
    if (std::sqrt(dx * dx + dy * dy + dz * dz) > options_.max_separation) {
      counts_.separated++;
      return;
    }
  }

  // The spot's pixels in (frame, index) order: its components come in
  // (frame, first) order, and those sharing a frame are merged by index.  A
  // frame with one component of the spot (nearly always) is a straight copy.
  spot.first = kept_.size();
  kept_.reserve(kept_.size() + n);
  size_t k = 0;
  while (k < count) {
    const int64_t z = nodes_[members[k]].frame;
    size_t e = k;
    while (e < count && nodes_[members[e]].frame == z)
      e++;
    // A frame's kept nodes move to the retained store together.
    const Arena &arena =
        nodes_[members[k]].retained ? retained_ : arenas_.at(z);
    if (e == k + 1) {
      const Node &node = nodes_[members[k]];
      sf_pixel *dst = kept_.extend(node.end - node.begin);
      for (size_t i = node.begin; i < node.end; i++)
        *dst++ = sf_pixel{static_cast<int32_t>(z), arena.index[i], arena.value[i]};
      k = e;
      continue;
    }
    pos_.assign(e - k, 0);
    std::vector<size_t> &pos = pos_;
    for (size_t j = k; j < e; j++)
      pos[j - k] = nodes_[members[j]].begin;
    for (;;) {
      size_t best = kNone;
      uint32_t best_index = 0;
      for (size_t j = k; j < e; j++) {
        const Node &node = nodes_[members[j]];
        if (pos[j - k] < node.end) {
          const uint32_t idx = arena.index[pos[j - k]];
          if (best == kNone || idx < best_index) {
            best = j;
            best_index = idx;
          }
        }
      }
      if (best == kNone)
        break;
      sf_pixel p;
      p.frame = static_cast<int32_t>(z);
      p.index = best_index;
      p.value = arena.value[pos[best - k]];
      kept_.push_back(p);
      pos[best - k]++;
    }
    k = e;
  }
  spots_.push_back(spot);
  counts_.accepted++;
}

} // namespace sfl
