// The three-dimensional grouping on the device (see stitch.hh).
//
// The reference's Labeller (src/spots/dials_spots.cc:57-360) keeps every live
// pixel and, frame by frame, unions each new pixel with its left and upper
// neighbours and with the same pixel of the previous frame, then emits the
// components that have nothing on the newest frame, in order of their first
// pixel.  The device labelling has already reduced each frame to components
// (numbered in order of first pixel) with exact integer moments, and listed
// the pairs of components that share a position with the previous frame.
// What is left works on those components ("nodes"), all of it here:
//
//   * pixels: each frame's signal pixels regrouped by component, each
//     component ascending by index, by a scatter driven by the labelling's
//     runs in component order (a component's few runs put in rank order
//     first) -- the layout every later gather reads;
//   * a node table in (frame, component) order, which is the reference's
//     order of first appearance; open components' nodes stay in it across
//     batches, closed ones are compacted away once they dominate;
//   * union-find over the frame links, lock-free, the larger root hooked
//     under the smaller, so a root is its component's first node;
//   * exact moments per root, merged with 64-bit atomics (128-bit sums as a
//     low word whose wrap carries into the high word), bounding box by
//     min/max, peak in two passes (the largest value, then the lowest
//     (frame, index) holding it);
//   * the closing order: a component with nothing on frame c closes when c is
//     added (dials_spots.cc:153-219), so closed roots are emitted by
//     (c, root), which is the order of set bits in a bitmap over
//     (c - first frame) x table + root, ranked by a popcount scan;
//   * emit (size filters, centroid, variance, separation filter), exactly as
//     the host stitch computes it from the same integers (labeller.cc);
//   * the spot pixel list: member nodes of accepted spots listed in table
//     order, sorted stably by spot (a radix sort of our own), so each spot's
//     nodes come in (frame, component) order; nodes of one spot on one frame
//     are merged by index, each pixel placed by binary searches in the
//     others (dials_spots.cc:286-327 sorts the spot's pixels by frame, index).
// The host reads a handful of counts per batch, grows the output buffers, and
// copies spots and pixels back only when they are asked for.
#include "stitch.hh"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <condition_variable>
#include <deque>
#include <exception>
#include <mutex>
#include <string>
#include <thread>

#include "sf_device.hh"

namespace sfi {

namespace {

using sfd::gmem;

constexpr uint32_t kDecisionSmall = 1, kDecisionLarge = 2,
                   kDecisionSeparated = 3, kDecisionAccepted = 4;
constexpr size_t kDeadSlack = 4096;

// A node: one component of one frame.
struct Node {
  const uint64_t *pix; // its grouped pixels, {index | value << 32}
  uint32_t n;          // pixels
  int32_t frame;       // relative to the run's base frame
  uint32_t parent;
  uint32_t buffer;     // id of the staged buffer holding its pixels
  uint32_t dead;       // its component has closed
  uint32_t pad;
};
static_assert(sizeof(Node) == 32, "node records are 32 bytes");

// Exact moments of a component (valid at roots); x and y moments at doubled
// pixel centres (2x + 1, 2y + 1) as the labelling gives them, z moments at
// doubled absolute frame centres, signed (two's complement in lo/hi).
struct Acc {
  uint64_t n, sum;
  uint64_t sum_sq[2], sum_x[2], sum_y[2], sum_xx[2], sum_yy[2], wz[2], qz[2];
  int32_t x0, x1, y0, y1, z0, z1, last;
  uint32_t peak_value;
  uint64_t peak_pos; // frame << 32 | index of the brightest pixel
  uint32_t maxv;     // scratch of the peak merge
  uint32_t pad;
  uint64_t minpos;
};

// One frame of a batch.
struct Frame {
  const sf_signal_pixel *ordered;
  const uint4 *runs;         // {first rank, x0, y, length}
  uint32_t *order;           // the runs in component order
  const uint32_t *comp_runs; // each component's first run in that order
  const sf_component *comps;
  const uint32_t *links;     // pairs (component, previous frame's component)
  uint32_t ncomp, nlinks, linked, buffer;
  uint32_t node_base, prev_base, prev_ncomp;
  int32_t frame;
  uint32_t *comp_prefix;     // scratch: exclusive prefix of component sizes
  uint64_t *pixels;          // the frame's grouped pixels
};

// What the host reads after the emission.
struct Summary {
  uint64_t closed, accepted, pixels, members, alive, dead, error;
};

// Device counters of the whole grouping.
struct Counters {
  unsigned long long groups, too_small, too_large, separated, accepted,
      closed_pixels;
};

__device__ __forceinline__ uint32_t load_parent(const Node *nodes, uint32_t x) {
  return __hip_atomic_load(&nodes[x].parent, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t root_of(const Node *nodes, uint32_t x) {
  for (;;) {
    const uint32_t p = load_parent(nodes, x);
    if (p == x)
      return x;
    x = p;
  }
}

// Lock-free union: the larger root is hooked under the smaller.
__device__ void unite(Node *nodes, uint32_t a, uint32_t b) {
  for (;;) {
    a = root_of(nodes, a);
    b = root_of(nodes, b);
    if (a == b)
      return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    uint32_t expected = a;
    if (__hip_atomic_compare_exchange_strong(&nodes[a].parent, &expected, b,
                                             __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}

__device__ __forceinline__ void add128(uint64_t *p, uint64_t lo, uint64_t hi) {
  if (lo) {
    const uint64_t old = atomicAdd(reinterpret_cast<unsigned long long *>(p),
                                   static_cast<unsigned long long>(lo));
    if (old + lo < old)
      hi += 1;
  }
  if (hi)
    atomicAdd(reinterpret_cast<unsigned long long *>(p + 1),
              static_cast<unsigned long long>(hi));
}

// ---- pixels regrouped by component ----------------------------------------

// Each frame's exclusive prefix of its components' sizes, one block a frame.
__global__ __launch_bounds__(1024) void component_prefix(const Frame *__restrict__ frames) {
  __shared__ uint64_t wave_sums[16];
  const Frame f = frames[blockIdx.x];
  const uint32_t nc = f.ncomp;
  const uint32_t per = (nc + blockDim.x - 1) / blockDim.x;
  const uint32_t c0 = threadIdx.x * per, c1 = min(nc, c0 + per);
  uint64_t mine = 0;
  for (uint32_t c = c0; c < c1; c++)
    mine += f.comps[c].n;
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  uint64_t incl = mine;
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t v = __shfl_up(incl, d);
    if (lane >= static_cast<uint32_t>(d))
      incl += v;
  }
  if (lane == 63)
    wave_sums[wave] = incl;
  __syncthreads();
  uint64_t before = incl - mine;
  for (uint32_t w = 0; w < wave; w++)
    before += wave_sums[w];
  for (uint32_t c = c0; c < c1; c++) {
    f.comp_prefix[c] = static_cast<uint32_t>(before);
    before += f.comps[c].n;
  }
  if (c1 == nc) // every thread ending there holds the same total
    f.comp_prefix[nc] = static_cast<uint32_t>(before);
}

// Each component's runs put in rank order, then its pixels copied run by run
// to their place: a thread per component.
__global__ __launch_bounds__(256) void group_pixels(const Frame *__restrict__ frames,
                                                    unsigned long long *__restrict__ error) {
  const Frame f = frames[blockIdx.y];
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= f.ncomp)
    return;
  const uint32_t lo = f.comp_runs[c], hi = f.comp_runs[c + 1];
  if (hi < lo) {
    atomicOr(error, 1ull);
    return;
  }
  uint32_t *order = f.order;
  for (uint32_t i = lo + 1; i < hi; i++) { // insertion sort by first rank
    const uint32_t k = order[i];
    const uint32_t key = f.runs[k].x;
    uint32_t j = i;
    while (j > lo && f.runs[order[j - 1]].x > key) {
      order[j] = order[j - 1];
      j--;
    }
    order[j] = k;
  }
  uint64_t *out = f.pixels + f.comp_prefix[c];
  const uint32_t n = f.comps[c].n;
  uint32_t written = 0;
  for (uint32_t i = lo; i < hi; i++) {
    const uint4 r = f.runs[order[i]];
    for (uint32_t j = 0; j < r.w && written < n; j++, written++) {
      const sf_signal_pixel p = f.ordered[r.x + j];
      out[written] = static_cast<uint64_t>(p.index) |
                     (static_cast<uint64_t>(p.value) << 32);
    }
  }
  if (written != n)
    atomicOr(error, 1ull);
}

// ---- nodes -----------------------------------------------------------------

__global__ __launch_bounds__(256) void make_nodes(const Frame *__restrict__ frames,
                                                  Node *__restrict__ nodes,
                                                  Acc *__restrict__ acc,
                                                  int64_t base) {
  const Frame f = frames[blockIdx.y];
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= f.ncomp)
    return;
  const uint32_t x = f.node_base + c;
  const sf_component cp = f.comps[c];
  Node nd;
  nd.pix = f.pixels + f.comp_prefix[c];
  nd.n = cp.n;
  nd.frame = f.frame;
  nd.parent = x;
  nd.buffer = f.buffer;
  nd.dead = 0;
  nd.pad = 0;
  nodes[x] = nd;
  Acc a;
  a.n = cp.n;
  a.sum = cp.sum;
  for (int k = 0; k < 2; k++) {
    a.sum_sq[k] = cp.sum_sq[k];
    a.sum_x[k] = cp.sum_x[k];
    a.sum_y[k] = cp.sum_y[k];
    a.sum_xx[k] = cp.sum_xx[k];
    a.sum_yy[k] = cp.sum_yy[k];
  }
  const int64_t az = 2 * (base + f.frame) + 1;
  const __int128 wz = static_cast<__int128>(az) * static_cast<__int128>(cp.sum);
  a.wz[0] = static_cast<uint64_t>(wz);
  a.wz[1] = static_cast<uint64_t>(static_cast<unsigned __int128>(wz) >> 64);
  const uint64_t az2 = static_cast<uint64_t>(az) * static_cast<uint64_t>(az);
  const unsigned __int128 qz = static_cast<unsigned __int128>(az2) * cp.sum;
  a.qz[0] = static_cast<uint64_t>(qz);
  a.qz[1] = static_cast<uint64_t>(qz >> 64);
  a.x0 = static_cast<int32_t>(cp.x0);
  a.x1 = static_cast<int32_t>(cp.x1);
  a.y0 = static_cast<int32_t>(cp.y0);
  a.y1 = static_cast<int32_t>(cp.y1);
  a.z0 = f.frame;
  a.z1 = f.frame + 1;
  a.last = f.frame;
  a.peak_value = static_cast<uint32_t>(cp.peak >> 32);
  a.peak_pos = (static_cast<uint64_t>(static_cast<uint32_t>(f.frame)) << 32) |
               (0xFFFFFFFFu - static_cast<uint32_t>(cp.peak));
  a.maxv = a.peak_value;
  a.pad = 0;
  a.minpos = ~0ull;
  acc[x] = a;
}

// The previous last frame's nodes name the open components: their roots
// (their parents, flattened) become candidates, each once.
__global__ __launch_bounds__(256) void collect_open(const Node *__restrict__ nodes,
                                                    uint32_t begin, uint32_t end,
                                                    uint32_t *__restrict__ claim,
                                                    uint32_t epoch,
                                                    uint32_t *__restrict__ cand,
                                                    uint32_t *__restrict__ ncand) {
  const uint32_t x = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= end)
    return;
  const uint32_t r = root_of(nodes, x);
  if (atomicExch(&claim[r], epoch) != epoch)
    cand[atomicAdd(ncand, 1u)] = r;
}

__global__ __launch_bounds__(256) void unite_links(const Frame *__restrict__ frames,
                                                   Node *__restrict__ nodes,
                                                   unsigned long long *__restrict__ error) {
  const Frame f = frames[blockIdx.y];
  if (!f.linked)
    return;
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= f.nlinks)
    return;
  const uint32_t here = f.links[2 * l], there = f.links[2 * l + 1];
  if (here >= f.ncomp || there >= f.prev_ncomp) {
    atomicOr(error, 2ull);
    return;
  }
  unite(nodes, f.node_base + here, f.prev_base + there);
}

// Candidate t: the old open roots first, then the batch's new nodes.
__device__ __forceinline__ bool candidate(uint32_t t, const uint32_t *cand,
                                          const uint32_t *ncand, uint32_t new_begin,
                                          uint32_t new_end, uint32_t *x) {
  const uint32_t nold = *ncand;
  if (t < nold) {
    *x = cand[t];
    return true;
  }
  const uint32_t k = new_begin + (t - nold);
  if (k >= new_end)
    return false;
  *x = k;
  return true;
}

// Every candidate that is no longer a root points at its root and adds its
// component's moments there.
__global__ __launch_bounds__(256) void merge_sums(Node *__restrict__ nodes,
                                                  Acc *__restrict__ acc,
                                                  const uint32_t *__restrict__ cand,
                                                  const uint32_t *__restrict__ ncand,
                                                  uint32_t new_begin, uint32_t new_end) {
  uint32_t x;
  if (!candidate(blockIdx.x * blockDim.x + threadIdx.x, cand, ncand, new_begin,
                 new_end, &x))
    return;
  const uint32_t r = root_of(nodes, x);
  if (r == x)
    return;
  nodes[x].parent = r;
  const Acc a = acc[x];
  Acc *t = &acc[r];
  atomicAdd(reinterpret_cast<unsigned long long *>(&t->n), a.n);
  atomicAdd(reinterpret_cast<unsigned long long *>(&t->sum), a.sum);
  add128(t->sum_sq, a.sum_sq[0], a.sum_sq[1]);
  add128(t->sum_x, a.sum_x[0], a.sum_x[1]);
  add128(t->sum_y, a.sum_y[0], a.sum_y[1]);
  add128(t->sum_xx, a.sum_xx[0], a.sum_xx[1]);
  add128(t->sum_yy, a.sum_yy[0], a.sum_yy[1]);
  add128(t->wz, a.wz[0], a.wz[1]);
  add128(t->qz, a.qz[0], a.qz[1]);
  atomicMin(&t->x0, a.x0);
  atomicMax(&t->x1, a.x1);
  atomicMin(&t->y0, a.y0);
  atomicMax(&t->y1, a.y1);
  atomicMin(&t->z0, a.z0);
  atomicMax(&t->z1, a.z1);
  atomicMax(&t->last, a.last);
  atomicMax(&t->maxv, a.peak_value);
}

// The lowest (frame, index) among the candidates holding the largest value.
__global__ __launch_bounds__(256) void merge_peaks(const Node *__restrict__ nodes,
                                                   Acc *__restrict__ acc,
                                                   const uint32_t *__restrict__ cand,
                                                   const uint32_t *__restrict__ ncand,
                                                   uint32_t new_begin, uint32_t new_end) {
  uint32_t x;
  if (!candidate(blockIdx.x * blockDim.x + threadIdx.x, cand, ncand, new_begin,
                 new_end, &x))
    return;
  const uint32_t r = nodes[x].parent; // flattened by merge_sums
  if (acc[x].peak_value == acc[r].maxv)
    atomicMin(reinterpret_cast<unsigned long long *>(&acc[r].minpos),
              static_cast<unsigned long long>(acc[x].peak_pos));
}

__global__ __launch_bounds__(256) void mark_open(const Node *__restrict__ nodes,
                                                 uint32_t begin, uint32_t end,
                                                 uint32_t *__restrict__ open,
                                                 uint32_t epoch) {
  const uint32_t x = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (x < end)
    open[nodes[x].parent] = epoch;
}

// Each root among the candidates takes its merged peak; a closed one sets its
// bit at (closing frame - first frame) x T + root.
__global__ __launch_bounds__(256) void close_roots(
    const Node *__restrict__ nodes, Acc *__restrict__ acc,
    const uint32_t *__restrict__ cand, const uint32_t *__restrict__ ncand,
    uint32_t new_begin, uint32_t new_end, const uint32_t *__restrict__ open,
    uint32_t epoch, int flush_all, int32_t first, uint64_t T,
    uint32_t *__restrict__ bits) {
  uint32_t x;
  if (!candidate(blockIdx.x * blockDim.x + threadIdx.x, cand, ncand, new_begin,
                 new_end, &x))
    return;
  if (nodes[x].parent != x)
    return;
  Acc *a = &acc[x];
  a->peak_pos = a->minpos;
  a->peak_value = a->maxv;
  a->minpos = ~0ull;
  if (!flush_all && open[x] == epoch)
    return;
  const uint64_t bucket = flush_all ? 0 : static_cast<uint64_t>(a->last + 1 - first);
  const uint64_t bit = bucket * T + x;
  atomicOr(&bits[bit / 32], 1u << (bit % 32));
}

__device__ __forceinline__ uint32_t rank_of(const uint32_t *bits,
                                            const uint32_t *prefix, uint64_t bit) {
  const uint64_t w = bit / 32;
  return prefix[w] + static_cast<uint32_t>(
                         __popc(bits[w] & ((1u << (bit % 32)) - 1u)));
}

__device__ __forceinline__ bool closed_bit(const Node *nodes, const Acc *acc,
                                           const uint32_t *bits, uint32_t r,
                                           int flush_all, int32_t first, uint64_t T,
                                           uint64_t *bit) {
  const uint64_t bucket =
      flush_all ? 0 : static_cast<uint64_t>(static_cast<int64_t>(acc[r].last) + 1 - first);
  if (!flush_all && (acc[r].last + 1 - first < 0))
    return false;
  *bit = bucket * T + r;
  return (bits[*bit / 32] >> (*bit % 32)) & 1u;
}

__global__ __launch_bounds__(256) void list_closed(
    const Node *__restrict__ nodes, const Acc *__restrict__ acc,
    const uint32_t *__restrict__ cand, const uint32_t *__restrict__ ncand,
    uint32_t new_begin, uint32_t new_end, int flush_all, int32_t first,
    uint64_t T, const uint32_t *__restrict__ bits,
    const uint32_t *__restrict__ prefix, uint32_t *__restrict__ emit_list) {
  uint32_t x;
  if (!candidate(blockIdx.x * blockDim.x + threadIdx.x, cand, ncand, new_begin,
                 new_end, &x))
    return;
  if (nodes[x].parent != x)
    return;
  uint64_t bit;
  if (closed_bit(nodes, acc, bits, x, flush_all, first, T, &bit))
    emit_list[rank_of(bits, prefix, bit)] = x;
}

// ---- emit ------------------------------------------------------------------

__device__ __forceinline__ unsigned __int128 u128(const uint64_t p[2]) {
  return (static_cast<unsigned __int128>(p[1]) << 64) | p[0];
}

// The host stitch's conversion, operation for operation.
__device__ __forceinline__ double to_double(unsigned __int128 v) {
  const uint64_t hi = static_cast<uint64_t>(v >> 64);
  const uint64_t lo = static_cast<uint64_t>(v);
  if (hi == 0)
    return static_cast<double>(lo);
  return __dadd_rn(__dmul_rn(static_cast<double>(hi), 0x1p64), static_cast<double>(lo));
}
__device__ __forceinline__ double to_double_signed(__int128 v) {
  return v < 0 ? -to_double(static_cast<unsigned __int128>(-v))
               : to_double(static_cast<unsigned __int128>(v));
}

// a * b when it fits 128 bits.
__device__ __forceinline__ bool mul128(unsigned __int128 a, unsigned __int128 b,
                                       unsigned __int128 *out) {
  const uint64_t ah = static_cast<uint64_t>(a >> 64), bh = static_cast<uint64_t>(b >> 64);
  const uint64_t al = static_cast<uint64_t>(a), bl = static_cast<uint64_t>(b);
  if (ah && bh)
    return false;
  const unsigned __int128 low = static_cast<unsigned __int128>(al) * bl;
  const unsigned __int128 cross =
      static_cast<unsigned __int128>(ah) * bl + static_cast<unsigned __int128>(al) * bh;
  if ((cross >> 64) != 0)
    return false;
  const unsigned __int128 shifted = cross << 64;
  *out = low + shifted;
  return *out >= low;
}

struct EmitParams {
  uint32_t width;
  uint64_t min_size, max_size;
  double max_separation;
  int64_t base;
};

__global__ __launch_bounds__(256) void emit_spots(
    const Acc *__restrict__ acc, const uint32_t *__restrict__ emit_list,
    const uint32_t *__restrict__ bits_prefix_total, uint32_t kmax, EmitParams p,
    sf_spot *__restrict__ spot_tmp, uint32_t *__restrict__ decision,
    uint64_t *__restrict__ accidx, uint64_t *__restrict__ pixoff,
    Counters *__restrict__ counters) {
  // Every lane reaches the counting at the end; entries 0 .. kmax exist (one
  // past the largest count, for the scans).
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t K = *bits_prefix_total;
  const bool valid = k < K;
  if (!valid && k <= kmax) {
    accidx[k] = 0;
    pixoff[k] = 0;
  }
  Acc a{};
  if (valid)
    a = acc[emit_list[k]];
  uint32_t d = 0;
  sf_spot s{};
  if (!valid) {
  } else if (a.n < p.min_size) {
    d = kDecisionSmall;
  } else if (a.n > p.max_size) {
    d = kDecisionLarge;
  } else {
    s.n_signal = static_cast<uint32_t>(a.n);
    s.bbox[0] = a.x0;
    s.bbox[1] = a.x1;
    s.bbox[2] = a.y0;
    s.bbox[3] = a.y1;
    s.bbox[4] = static_cast<int32_t>(p.base + a.z0);
    s.bbox[5] = static_cast<int32_t>(p.base + a.z1);
    const unsigned __int128 S = a.sum;
    const double sum_d = to_double(S);
    const double sum_sq_d = to_double(u128(a.sum_sq));
    s.intensity = sum_d;
    s.intensity_variance = sum_d;
    const uint32_t peak_index = static_cast<uint32_t>(a.peak_pos);
    const int64_t peak_frame = p.base + static_cast<int32_t>(a.peak_pos >> 32);
    const double peak_pos[3] = {
        __dadd_rn(static_cast<double>(peak_index % p.width), 0.5),
        __dadd_rn(static_cast<double>(peak_index / p.width), 0.5),
        __dadd_rn(static_cast<double>(peak_frame), 0.5)};
    if (sum_d > 0.0) {
      const __int128 wz = static_cast<__int128>(u128(a.wz));
      const double w_d[3] = {to_double(u128(a.sum_x)), to_double(u128(a.sum_y)),
                             to_double_signed(wz)};
      for (int ax = 0; ax < 3; ax++)
        s.position[ax] = __ddiv_rn(__dmul_rn(w_d[ax], 0.5), sum_d);
      if (__dmul_rn(sum_d, sum_d) > sum_sq_d) {
        const unsigned __int128 W[3] = {u128(a.sum_x), u128(a.sum_y),
                                        static_cast<unsigned __int128>(wz < 0 ? -wz : wz)};
        const unsigned __int128 Q[3] = {u128(a.sum_xx), u128(a.sum_yy), u128(a.qz)};
        const double denom = __dsub_rn(__dmul_rn(sum_d, sum_d), sum_sq_d);
        for (int ax = 0; ax < 3; ax++) {
          double delta_sq;
          unsigned __int128 sq, ww;
          if (mul128(S, Q[ax], &sq) && mul128(W[ax], W[ax], &ww)) {
            delta_sq = __ddiv_rn(to_double(sq - ww), __dmul_rn(4.0, sum_d));
          } else {
            // Beyond 128 bits (never at detector counts): in double.
            const double Sd = sum_d, Wd = to_double(W[ax]), Qd = to_double(Q[ax]);
            delta_sq = __ddiv_rn(__dsub_rn(__dmul_rn(Sd, Qd), __dmul_rn(Wd, Wd)),
                                 __dmul_rn(4.0, Sd));
          }
          const double v = __ddiv_rn(__dmul_rn(delta_sq, sum_d), denom);
          s.variance[ax] = __dadd_rn(__ddiv_rn(v, sum_d), 1.0 / 12.0);
        }
      } else {
        for (int ax = 0; ax < 3; ax++)
          s.variance[ax] = 1.0 / 12.0;
      }
      if (s.bbox[5] == s.bbox[4] + 1)
        s.position[2] = __dadd_rn(static_cast<double>(s.bbox[4]), 0.5);
    } else {
      s.position[0] = (s.bbox[1] + s.bbox[0]) / 2.0;
      s.position[1] = (s.bbox[3] + s.bbox[2]) / 2.0;
      s.position[2] = (s.bbox[5] + s.bbox[4]) / 2.0;
    }
    d = kDecisionAccepted;
    if (p.max_separation > 0.0) {
      const double dx = __dsub_rn(peak_pos[0], s.position[0]);
      const double dy = __dsub_rn(peak_pos[1], s.position[1]);
      const double dz = __dsub_rn(peak_pos[2], s.position[2]);
      const double dist = __dsqrt_rn(
          __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)));
      if (dist > p.max_separation)
        d = kDecisionSeparated;
    }
  }
  if (valid) {
    decision[k] = d;
    spot_tmp[k] = s;
    accidx[k] = d == kDecisionAccepted ? 1 : 0;
    pixoff[k] = d == kDecisionAccepted ? a.n : 0;
  }
  // One atomic per counter per wave.
  uint64_t pix = valid ? a.n : 0;
  for (int o = 32; o > 0; o >>= 1)
    pix += __shfl_xor(pix, o);
  const unsigned long long groups = __popcll(__ballot(valid));
  const unsigned long long small = __popcll(__ballot(d == kDecisionSmall));
  const unsigned long long large = __popcll(__ballot(d == kDecisionLarge));
  const unsigned long long sep = __popcll(__ballot(d == kDecisionSeparated));
  const unsigned long long acc_n = __popcll(__ballot(d == kDecisionAccepted));
  if ((threadIdx.x & 63) == 0 && groups) {
    atomicAdd(&counters->groups, groups);
    atomicAdd(&counters->closed_pixels, static_cast<unsigned long long>(pix));
    if (small)
      atomicAdd(&counters->too_small, small);
    if (large)
      atomicAdd(&counters->too_large, large);
    if (sep)
      atomicAdd(&counters->separated, sep);
    if (acc_n)
      atomicAdd(&counters->accepted, acc_n);
  }
}

// Per table node: dead already, dead now (its component closed in this
// batch; marked), or alive (counted, and against its staged buffer); a node
// dying now in an accepted spot is a member, counted against its spot.  The
// grid's lanes cover the nodes in order, so each wave owns two whole words
// of the member and alive bitmaps and writes them without atomics; counts
// go in one atomic per wave where the wave agrees.
__global__ __launch_bounds__(256) void classify_nodes(
    Node *__restrict__ nodes, const Acc *__restrict__ acc, uint32_t end,
    int flush_all, int32_t first, uint64_t T, const uint32_t *__restrict__ bits,
    const uint32_t *__restrict__ prefix, const uint32_t *__restrict__ decision,
    const uint64_t *__restrict__ accidx, uint32_t *__restrict__ membits,
    uint32_t *__restrict__ alive_bits, unsigned long long *__restrict__ spot_members,
    uint32_t *__restrict__ live_counts, uint32_t buffer_base, uint32_t nbuffers,
    Summary *__restrict__ summary) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63;
  bool alive = false, dies = false, member = false;
  uint32_t b = 0;
  if (x < end && !nodes[x].dead) {
    const uint32_t r = root_of(nodes, x);
    uint64_t bit;
    if (!closed_bit(nodes, acc, bits, r, flush_all, first, T, &bit)) {
      alive = true;
      b = nodes[x].buffer - buffer_base;
    } else {
      dies = true;
      nodes[x].dead = 1;
      const uint32_t k = rank_of(bits, prefix, bit);
      if (decision[k] == kDecisionAccepted) {
        member = true;
        atomicAdd(&spot_members[accidx[k]], 1ull);
      }
    }
  }
  const uint64_t alive_mask = __ballot(alive);
  const uint64_t member_mask = __ballot(member);
  const uint32_t w0 = (x - lane) / 32; // the wave's first word (x - lane is 64-aligned)
  if (lane == 0 && w0 * 32 < end) {
    alive_bits[w0] = static_cast<uint32_t>(alive_mask);
    membits[w0] = static_cast<uint32_t>(member_mask);
  }
  if (lane == 32 && w0 * 32 + 32 < end) {
    alive_bits[w0 + 1] = static_cast<uint32_t>(alive_mask >> 32);
    membits[w0 + 1] = static_cast<uint32_t>(member_mask >> 32);
  }
  const unsigned long long died = __popcll(__ballot(dies));
  if (lane == 0 && died)
    atomicAdd(reinterpret_cast<unsigned long long *>(&summary->dead), died);
  if (alive_mask) {
    // Nodes of one batch share a buffer: usually one atomic for the wave.
    const int leader = __ffsll(static_cast<unsigned long long>(alive_mask)) - 1;
    const uint32_t b0 = __shfl(b, leader);
    const uint64_t same = __ballot(alive && b == b0);
    if (static_cast<int>(lane) == leader && b0 < nbuffers)
      atomicAdd(&live_counts[b0], static_cast<uint32_t>(__popcll(same)));
    if (alive && b != b0 && b < nbuffers)
      atomicAdd(&live_counts[b], 1u);
  }
}

__device__ __forceinline__ uint64_t bits_total(const uint32_t *bits,
                                               const uint32_t *prefix, uint64_t words) {
  return prefix[words - 1] + static_cast<uint32_t>(__popc(bits[words - 1]));
}

__global__ void key_total(const uint32_t *__restrict__ bits,
                          const uint32_t *__restrict__ prefix, uint64_t words,
                          uint32_t *__restrict__ out) {
  *out = static_cast<uint32_t>(bits_total(bits, prefix, words));
}

__global__ void write_summary(const uint32_t *__restrict__ ktotal,
                              const uint64_t *__restrict__ accidx,
                              const uint64_t *__restrict__ pixoff,
                              const uint32_t *__restrict__ membits,
                              const uint32_t *__restrict__ memprefix,
                              const uint32_t *__restrict__ alive_bits,
                              const uint32_t *__restrict__ alive_prefix,
                              uint64_t twords,
                              const unsigned long long *__restrict__ error,
                              Summary *__restrict__ s) {
  const uint32_t K = *ktotal;
  s->closed = K;
  s->accepted = accidx[K]; // exclusive scans over K + 1 entries
  s->pixels = pixoff[K];
  s->members = bits_total(membits, memprefix, twords);
  s->alive = bits_total(alive_bits, alive_prefix, twords);
  s->error = *error;
}

__global__ __launch_bounds__(256) void write_spots(
    const uint32_t *__restrict__ key_total, const uint32_t *__restrict__ decision,
    const sf_spot *__restrict__ spot_tmp, const uint64_t *__restrict__ accidx,
    const uint64_t *__restrict__ pixoff, uint64_t pixels_before,
    sf_spot *__restrict__ out,
    uint64_t *__restrict__ spot_pix) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= *key_total || decision[k] != kDecisionAccepted)
    return;
  sf_spot s = spot_tmp[k];
  s.first = pixels_before + pixoff[k];
  out[accidx[k]] = s; // this batch's spots only
  spot_pix[accidx[k]] = pixoff[k];
}

__global__ __launch_bounds__(256) void list_members(
    const Node *__restrict__ nodes, const Acc *__restrict__ acc, uint32_t end,
    int flush_all, int32_t first, uint64_t T, const uint32_t *__restrict__ bits,
    const uint32_t *__restrict__ prefix, const uint64_t *__restrict__ accidx,
    const uint32_t *__restrict__ membits, const uint32_t *__restrict__ memprefix,
    uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= end || !((membits[x / 32] >> (x % 32)) & 1u))
    return;
  const uint32_t r = root_of(nodes, x);
  uint64_t bit;
  closed_bit(nodes, acc, bits, r, flush_all, first, T, &bit);
  const uint32_t j = rank_of(membits, memprefix, x);
  keys[j] = static_cast<uint32_t>(accidx[rank_of(bits, prefix, bit)]);
  vals[j] = x;
}

__global__ __launch_bounds__(256) void member_sizes(const Node *__restrict__ nodes,
                                                    const uint32_t *__restrict__ vals,
                                                    uint32_t m,
                                                    uint64_t *__restrict__ npix) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < m)
    npix[p] = nodes[vals[p]].n;
}

// One wave per member node (in spot, frame, component order): its pixels to
// their place in the spot pixel list, merged by index with the spot's other
// nodes on the same frame.
__global__ __launch_bounds__(256) void gather_members(
    const Node *__restrict__ nodes, const uint32_t *__restrict__ keys,
    const uint32_t *__restrict__ vals, uint32_t m,
    const uint64_t *__restrict__ npix_prefix, const uint64_t *__restrict__ spot_first,
    const uint64_t *__restrict__ spot_pix, int64_t base, sf_pixel *__restrict__ out) {
  const uint32_t p = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / 64);
  if (p >= m)
    return;
  const uint32_t s = keys[p];
  const uint32_t sp = static_cast<uint32_t>(spot_first[s]);
  const uint32_t se = static_cast<uint32_t>(spot_first[s + 1]);
  const Node nd = nodes[vals[p]];
  uint32_t gs = p, ge = p + 1;
  while (gs > sp && nodes[vals[gs - 1]].frame == nd.frame)
    gs--;
  while (ge < se && nodes[vals[ge]].frame == nd.frame)
    ge++;
  const uint64_t at = spot_pix[s] + (npix_prefix[gs] - npix_prefix[sp]);
  const int32_t frame = static_cast<int32_t>(base + nd.frame);
  // `out` holds this batch's pixels; `at` is relative to them.
  SF_GLOBAL uint32_t *dst = reinterpret_cast<SF_GLOBAL uint32_t *>(gmem(out) + at);
  const SF_GLOBAL uint64_t *px = gmem(nd.pix);
  for (uint32_t k = threadIdx.x & 63; k < nd.n; k += 64) {
    const uint64_t v = px[k];
    const uint32_t index = static_cast<uint32_t>(v);
    uint64_t rank = k;
    for (uint32_t o = gs; o < ge; o++) {
      if (o == p)
        continue;
      const Node other = nodes[vals[o]];
      const SF_GLOBAL uint64_t *q = gmem(other.pix);
      uint32_t lo = 0, hi = other.n;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (static_cast<uint32_t>(q[mid]) < index)
          lo = mid + 1;
        else
          hi = mid;
      }
      rank += lo;
    }
    dst[3 * rank] = static_cast<uint32_t>(frame);
    dst[3 * rank + 1] = index;
    dst[3 * rank + 2] = static_cast<uint32_t>(v >> 32);
  }
}

// ---- compaction ------------------------------------------------------------

__global__ __launch_bounds__(256) void compact_nodes(
    const Node *__restrict__ nodes, const Acc *__restrict__ acc, uint32_t end,
    const uint32_t *__restrict__ alive, const uint32_t *__restrict__ prefix,
    Node *__restrict__ nodes_out, Acc *__restrict__ acc_out) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= end || !((alive[x / 32] >> (x % 32)) & 1u))
    return;
  const uint32_t r = root_of(nodes, x);
  Node nd = nodes[x];
  const uint32_t to = rank_of(alive, prefix, x);
  nd.parent = rank_of(alive, prefix, r); // a live node's root is live
  nd.dead = 0;
  nodes_out[to] = nd;
  if (r == x)
    acc_out[to] = acc[x];
}

// ---- scans and the sort ----------------------------------------------------

constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = 256 * kScanItems;

__device__ uint64_t block_exclusive(uint64_t v, uint64_t *total) {
  __shared__ uint64_t sums[4];
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  uint64_t incl = v;
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t t = __shfl_up(incl, d);
    if (lane >= static_cast<uint32_t>(d))
      incl += t;
  }
  if (lane == 63)
    sums[wave] = incl;
  __syncthreads();
  uint64_t before = incl - v, all = 0;
  for (uint32_t w = 0; w < 4; w++) {
    if (w < wave)
      before += sums[w];
    all += sums[w];
  }
  __syncthreads();
  *total = all;
  return before;
}

__global__ __launch_bounds__(256) void scan_tiles(uint64_t *__restrict__ data, uint64_t n,
                                                  uint64_t *__restrict__ tiles) {
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  uint64_t v[kScanItems], mine = 0;
  for (uint32_t i = 0; i < kScanItems; i++) {
    v[i] = t0 + i < n ? data[t0 + i] : 0;
    mine += v[i];
  }
  uint64_t total;
  uint64_t run = block_exclusive(mine, &total);
  for (uint32_t i = 0; i < kScanItems; i++) {
    if (t0 + i < n)
      data[t0 + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0)
    tiles[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void scan_tile_sums(uint64_t *__restrict__ tiles,
                                                      uint32_t ntiles) {
  uint64_t carry = 0;
  for (uint32_t t0 = 0; t0 < ntiles; t0 += 256) {
    const uint32_t t = t0 + threadIdx.x;
    const uint64_t v = t < ntiles ? tiles[t] : 0;
    uint64_t total;
    const uint64_t before = block_exclusive(v, &total);
    if (t < ntiles)
      tiles[t] = carry + before;
    carry += total;
  }
}

__global__ __launch_bounds__(256) void scan_add(uint64_t *__restrict__ data, uint64_t n,
                                                const uint64_t *__restrict__ tiles) {
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  const uint64_t add = tiles[blockIdx.x];
  for (uint32_t i = 0; i < kScanItems; i++)
    if (t0 + i < n)
      data[t0 + i] += add;
}

constexpr uint32_t kSortItems = 16;
constexpr uint32_t kSortTile = 256 * kSortItems;

__global__ __launch_bounds__(256) void sort_histogram(const uint32_t *__restrict__ keys,
                                                      uint32_t n, uint32_t shift,
                                                      uint64_t *__restrict__ hist,
                                                      uint32_t nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t t0 = blockIdx.x * kSortTile;
  for (uint32_t i = 0; i < kSortItems; i++) {
    const uint32_t e = t0 + i * 256 + threadIdx.x;
    if (e < n)
      atomicAdd(&h[(keys[e] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[static_cast<uint64_t>(threadIdx.x) * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable: a tile's elements are visited in index order (round by round,
// lanes in order within a round), and each takes the next place of its digit.
__global__ __launch_bounds__(256) void sort_scatter(
    const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals, uint32_t n,
    uint32_t shift, const uint64_t *__restrict__ hist, uint32_t nblocks,
    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wave_count[4][256];
  base[threadIdx.x] =
      static_cast<uint32_t>(hist[static_cast<uint64_t>(threadIdx.x) * nblocks + blockIdx.x]);
  for (uint32_t w = 0; w < 4; w++)
    wave_count[w][threadIdx.x] = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t t0 = blockIdx.x * kSortTile;
  for (uint32_t i = 0; i < kSortItems; i++) {
    const uint32_t e = t0 + i * 256 + threadIdx.x;
    const bool valid = e < n;
    const uint32_t key = valid ? keys[e] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t same = __ballot(valid);
    for (uint32_t b = 0; b < 8; b++) {
      const uint64_t set = __ballot(valid && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? set : ~set;
    }
    const uint32_t rank = static_cast<uint32_t>(__popcll(same & below));
    if (valid && rank == 0)
      wave_count[wave][d] = static_cast<uint32_t>(__popcll(same));
    __syncthreads();
    if (valid) {
      uint32_t at = base[d] + rank;
      for (uint32_t w = 0; w < wave; w++)
        at += wave_count[w][d];
      keys_out[at] = key;
      vals_out[at] = vals[e];
    }
    __syncthreads();
    base[threadIdx.x] += wave_count[0][threadIdx.x] + wave_count[1][threadIdx.x] +
                         wave_count[2][threadIdx.x] + wave_count[3][threadIdx.x];
    for (uint32_t w = 0; w < 4; w++)
      wave_count[w][threadIdx.x] = 0;
    __syncthreads();
  }
}

unsigned blocks_for(uint64_t n, uint32_t per) {
  return static_cast<unsigned>(std::max<uint64_t>(1, (n + per - 1) / per));
}

int bits_for(uint64_t values) {
  int b = 0;
  while (b < 32 && values > (uint64_t(1) << b))
    b++;
  return b;
}

} // namespace

// Exclusive scan of n u64 values in place; `tiles` holds one word per 2048.
static void scan_u64(uint64_t *data, uint64_t n, DeviceBuffer &tiles, hipStream_t s) {
  if (n == 0)
    return;
  const uint32_t nt = blocks_for(n, kScanTile);
  tiles.ensure(static_cast<size_t>(nt) * 8, s);
  scan_tiles<<<nt, 256, 0, s>>>(data, n, tiles.as<uint64_t>());
  scan_tile_sums<<<1, 256, 0, s>>>(tiles.as<uint64_t>(), nt);
  scan_add<<<nt, 256, 0, s>>>(data, n, tiles.as<uint64_t>());
  check(hipGetLastError(), "launching a scan");
}

// Stable sort of n (key, value) pairs by keys below 2^key_bits; the result
// ends in (keys, vals).
static void sort_pairs(DeviceBuffer &keys, DeviceBuffer &vals, DeviceBuffer &keys2,
                       DeviceBuffer &vals2, uint32_t n, int key_bits,
                       DeviceBuffer &hist, DeviceBuffer &tiles, hipStream_t s) {
  if (n <= 1 || key_bits == 0)
    return;
  const uint32_t nb = blocks_for(n, kSortTile);
  hist.ensure(static_cast<size_t>(nb) * 256 * 8, s);
  uint32_t *k_in = keys.as<uint32_t>(), *v_in = vals.as<uint32_t>();
  uint32_t *k_out = keys2.as<uint32_t>(), *v_out = vals2.as<uint32_t>();
  for (int shift = 0; shift < key_bits; shift += 8) {
    sort_histogram<<<nb, 256, 0, s>>>(k_in, n, static_cast<uint32_t>(shift),
                                      hist.as<uint64_t>(), nb);
    scan_u64(hist.as<uint64_t>(), static_cast<uint64_t>(nb) * 256, tiles, s);
    sort_scatter<<<nb, 256, 0, s>>>(k_in, v_in, n, static_cast<uint32_t>(shift),
                                    hist.as<uint64_t>(), nb, k_out, v_out);
    check(hipGetLastError(), "launching the sort");
    std::swap(k_in, k_out);
    std::swap(v_in, v_out);
  }
  if (k_in != keys.as<uint32_t>()) {
    check(hipMemcpyAsync(keys.ptr, k_in, static_cast<size_t>(n) * 4,
                         hipMemcpyDeviceToDevice, s),
          "placing the sorted keys");
    check(hipMemcpyAsync(vals.ptr, v_in, static_cast<size_t>(n) * 4,
                         hipMemcpyDeviceToDevice, s),
          "placing the sorted values");
  }
}

// ---- copies back to host memory --------------------------------------------

namespace {

// Host memory copy of a block, on the process's copy threads when large: the
// destination's pages are touched for the first time here, and faulting them
// in on one thread is the slow part.
void parallel_copy(void *dst, const void *src, size_t bytes) {
  const CopyPiece piece{dst, src, bytes};
  host_copy(&piece, 1);
}

} // namespace

// Each batch's new records go back to host memory on a thread of its own
// while the next batch runs: the device copies them, once their kernels are
// done, into two pinned pieces on a stream of the drain's own, and the thread
// moves each piece on to the end of the host arrays.  The host arrays are
// only touched on that thread until wait() says every copy is done.
class DeviceStitch::Drain {
public:
  Drain() {
    check(hipGetDevice(&device_), "hipGetDevice");
    check(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking), "hipStreamCreate");
    for (int i = 0; i < 2; i++) {
      check(hipHostMalloc(&buf_[i], kPiece, hipHostMallocDefault),
            "hipHostMalloc (the grouping's copy-back staging)");
      check(hipEventCreateWithFlags(&piece_[i], hipEventDisableTiming), "hipEventCreate");
    }
    thread_ = std::thread([this] { run(); });
  }
  ~Drain() {
    {
      std::lock_guard<std::mutex> lock(m_);
      stop_ = true;
    }
    cv_.notify_all();
    thread_.join();
    (void)hipStreamSynchronize(copy_);
    for (int i = 0; i < 2; i++) {
      (void)hipEventDestroy(piece_[i]);
      (void)hipHostFree(buf_[i]);
    }
    (void)hipStreamDestroy(copy_);
  }
  Drain(const Drain &) = delete;
  Drain &operator=(const Drain &) = delete;

  // A spots and P pixels, at `spots` and `pixels` on the device, onto the
  // host arrays after `ready` has fired; the job's ticket.
  uint64_t submit(const sf_spot *spots, size_t A, const sf_pixel *pixels, size_t P,
                  hipEvent_t ready) {
    std::lock_guard<std::mutex> lock(m_);
    jobs_.push_back(Job{spots, A, pixels, P, ready, ++submitted_});
    cv_.notify_all();
    return submitted_;
  }
  // Waits for the job `ticket` and every one before it; rethrows the first
  // failure of a copy.
  void wait(uint64_t ticket) {
    std::unique_lock<std::mutex> lock(m_);
    done_cv_.wait(lock, [&] { return done_ >= ticket; });
    if (error_)
      std::rethrow_exception(error_);
  }
  void wait_all() { wait(submitted_); }
  const sf_spot *spots() const { return spots_.data(); }
  const sf_pixel *pixels() const { return pixels_.data(); }

private:
  struct Job {
    const sf_spot *spots;
    size_t A;
    const sf_pixel *pixels;
    size_t P;
    hipEvent_t ready;
    uint64_t ticket;
  };
  static constexpr size_t kPiece = size_t(8) << 20;

  void run() {
    (void)hipSetDevice(device_);
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lock(m_);
        cv_.wait(lock, [&] { return stop_ || !jobs_.empty(); });
        if (jobs_.empty())
          return; // stopped with nothing left
        j = jobs_.front();
      }
      try {
        if (!error_) {
          check(hipStreamWaitEvent(copy_, j.ready, 0), "hipStreamWaitEvent");
          copy(spots_.extend(j.A), j.spots, j.A * sizeof(sf_spot));
          copy(pixels_.extend(j.P), j.pixels, j.P * sizeof(sf_pixel));
        }
      } catch (...) {
        error_ = std::current_exception();
      }
      {
        std::lock_guard<std::mutex> lock(m_);
        jobs_.pop_front();
        done_ = j.ticket;
      }
      done_cv_.notify_all();
    }
  }

  // A block through the two pieces: the next crosses PCIe while the last is
  // copied on.
  void copy(void *dst, const void *src, size_t bytes) {
    char *to = static_cast<char *>(dst);
    const char *from = static_cast<const char *>(src);
    size_t off_[2] = {0, 0}, n_[2] = {0, 0};
    auto finish = [&](int i) {
      if (n_[i] == 0)
        return;
      check(hipEventSynchronize(piece_[i]), "waiting for a copy back");
      parallel_copy(to + off_[i], buf_[i], n_[i]);
      n_[i] = 0;
    };
    int i = 0;
    for (size_t off = 0; off < bytes; off += kPiece, i ^= 1) {
      finish(i);
      const size_t n = std::min(kPiece, bytes - off);
      check(hipMemcpyAsync(buf_[i], from + off, n, hipMemcpyDeviceToHost, copy_),
            "copying records back");
      check(hipEventRecord(piece_[i], copy_), "hipEventRecord");
      off_[i] = off;
      n_[i] = n;
    }
    finish(i); // the older piece first
    finish(i ^ 1);
  }

  int device_ = 0;
  hipStream_t copy_ = nullptr;
  void *buf_[2] = {nullptr, nullptr};
  hipEvent_t piece_[2] = {nullptr, nullptr};
  sfl::BigArray<sf_spot> spots_;
  sfl::BigArray<sf_pixel> pixels_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  std::deque<Job> jobs_;
  uint64_t submitted_ = 0, done_ = 0;
  bool stop_ = false;
  std::exception_ptr error_;
  std::thread thread_;
};

// ---- the host side ---------------------------------------------------------

DeviceStitch::DeviceStitch(size_t height, size_t width, const sfl::Options &options)
    : height_(height), width_(width), options_(options) {
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  counters_.ensure(sizeof(Counters), stream_);
  summary_.ensure(sizeof(Summary), stream_);
  error_.ensure(8, stream_);
  launch_fill(counters_.as<uint32_t>(), sizeof(Counters) / 4, 0, stream_);
  launch_fill(error_.as<uint32_t>(), 2, 0, stream_);
  for (hipEvent_t &e : out_ready_)
    check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  drain_ = std::make_unique<Drain>();
}

DeviceStitch::~DeviceStitch() {
  drain_.reset(); // its copies read the output buffers
  for (hipEvent_t e : out_ready_)
    if (e)
      (void)hipEventDestroy(e);
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
  }
}

size_t DeviceStitch::held_bytes() const {
  size_t b = 0;
  for (const Staged &s : staged_)
    b += s.buf ? s.buf->bytes : 0;
  return b;
}

void DeviceStitch::grow_table(size_t need) {
  need = std::max<size_t>(need, 1);
  if (need <= table_cap_)
    return;
  const size_t cap = std::max(need + need / 2, size_t(1) << 16);
  nodes_.ensure_keep(cap * sizeof(Node), stream_);
  acc_.ensure_keep(cap * sizeof(Acc), stream_);
  // The stamps start over with the epoch.
  stamp_open_.ensure(cap * 4, stream_);
  stamp_claim_.ensure(cap * 4, stream_);
  launch_fill(stamp_open_.as<uint32_t>(), cap, 0, stream_);
  launch_fill(stamp_claim_.as<uint32_t>(), cap, 0, stream_);
  epoch_ = 0;
  table_cap_ = cap;
}

void DeviceStitch::add(Engine &e, const int *ids, size_t count, int64_t first_frame,
                       const bool *linked) {
  if (count == 0)
    return;
  if (started_ && first_frame <= last_frame_)
    throw std::runtime_error("dials_spots: frame " + std::to_string(first_frame) +
                             " arrived after frame " + std::to_string(last_frame_) +
                             "; frames must be added in ascending order");
  // The engine's labelling must be complete before this stream reads it.
  check(hipStreamSynchronize(e.stream()), "waiting for the labelling");
  if (started_ && first_frame != last_frame_ + 1)
    close_all(); // a gap closes everything (dials_spots.cc:68-69)
  if (!started_ || table_end_ == 0)
    base_ = first_frame;
  if (first_frame - base_ > (int64_t(1) << 30))
    throw std::runtime_error("dials_spots: a run of more than 2^30 frames is "
                             "beyond the device grouping");

  std::vector<size_t> n(count), nc(count), nl(count);
  size_t total_pix = 0, total_comp = 0, most_comp = 0, most_links = 0;
  for (size_t i = 0; i < count; i++) {
    e.label_counts(ids[i], &n[i], &nc[i], &nl[i], i == 0);
    total_pix += n[i];
    total_comp += nc[i];
    most_comp = std::max(most_comp, nc[i]);
    most_links = std::max(most_links, nl[i]);
  }
  if (table_end_ + total_comp > 0xF0000000u)
    throw std::runtime_error("dials_spots: too many components open at once");
  counts_.signal_pixels += total_pix;
  live_ += total_pix;

  // A pooled buffer for the batch's grouped pixels.
  Staged st;
  const size_t bytes = std::max<size_t>(total_pix * 8, 256);
  for (auto it = spare_.begin(); it != spare_.end(); ++it)
    if ((*it)->bytes >= bytes) {
      st.buf = std::move(*it);
      spare_.erase(it);
      break;
    }
  if (!st.buf) {
    st.buf = std::make_unique<DeviceBuffer>();
    st.buf->ensure(bytes + bytes / 4, stream_);
  }
  st.id = next_buffer_++;

  const size_t new_begin = table_end_;
  grow_table(table_end_ + total_comp);
  comp_prefix_.ensure((total_comp + count) * 4 + 256, stream_);
  Frame *fr = host_desc_.get<Frame>(count);
  size_t node_at = new_begin, pix_at = 0, cp_at = 0;
  for (size_t i = 0; i < count; i++) {
    const FrameSlot &s = e.slot(ids[i]);
    Frame f;
    std::memset(&f, 0, sizeof f);
    f.ordered = s.ordered.as<sf_signal_pixel>();
    f.runs = s.unordered.as<uint4>();
    f.order = s.parent.as<uint32_t>();
    f.comp_runs = s.comp_runs.as<uint32_t>();
    f.comps = s.comps.as<sf_component>();
    f.links = s.links.as<uint32_t>();
    f.ncomp = static_cast<uint32_t>(nc[i]);
    f.nlinks = static_cast<uint32_t>(nl[i]);
    f.buffer = st.id;
    f.node_base = static_cast<uint32_t>(node_at);
    f.frame = static_cast<int32_t>(first_frame + static_cast<int64_t>(i) - base_);
    if (i == 0) {
      f.prev_base = static_cast<uint32_t>(pl_begin_);
      f.prev_ncomp = static_cast<uint32_t>(pl_end_ - pl_begin_);
      f.linked = (linked[0] && started_ && !options_.two_d && table_end_ > 0) ? 1u : 0u;
    } else {
      f.prev_base = static_cast<uint32_t>(node_at - nc[i - 1]);
      f.prev_ncomp = static_cast<uint32_t>(nc[i - 1]);
      f.linked = (linked[i] && !options_.two_d) ? 1u : 0u;
    }
    f.comp_prefix = comp_prefix_.as<uint32_t>() + cp_at;
    f.pixels = st.buf->as<uint64_t>() + pix_at;
    fr[i] = f;
    node_at += nc[i];
    pix_at += n[i];
    cp_at += nc[i] + 1;
  }
  frame_desc_.ensure(count * sizeof(Frame), stream_);
  check(hipMemcpyAsync(frame_desc_.ptr, fr, count * sizeof(Frame),
                       hipMemcpyHostToDevice, stream_),
        "uploading the grouping's frames");
  const Frame *fd = frame_desc_.as<Frame>();
  const unsigned nf = static_cast<unsigned>(count);
  unsigned long long *err = reinterpret_cast<unsigned long long *>(error_.ptr);
  component_prefix<<<nf, 1024, 0, stream_>>>(fd);
  if (most_comp) {
    group_pixels<<<dim3(blocks_for(most_comp, 256), nf), 256, 0, stream_>>>(fd, err);
    make_nodes<<<dim3(blocks_for(most_comp, 256), nf), 256, 0, stream_>>>(
        fd, nodes_.as<Node>(), acc_.as<Acc>(), base_);
  }
  check(hipGetLastError(), "launching the grouping's node build");
  table_end_ = node_at;
  staged_.push_back(std::move(st));

  const int32_t first_rel = static_cast<int32_t>(first_frame - base_);
  const size_t last_begin = node_at - nc[count - 1];
  // The open components' roots, taken before the union can hook one under
  // another.
  collect(pl_begin_, pl_end_);
  if (most_links)
    unite_links<<<dim3(blocks_for(most_links, 256), nf), 256, 0, stream_>>>(
        fd, nodes_.as<Node>(), err);
  check(hipGetLastError(), "launching the grouping's union");
  process(first_rel, static_cast<int>(count), false, new_begin, node_at, pl_begin_,
          pl_end_, last_begin, node_at);
  started_ = true;
  last_frame_ = first_frame + static_cast<int64_t>(count) - 1;
  if (options_.two_d) {
    // Every component closed with its own frame.
    pl_begin_ = pl_end_ = table_end_;
  }
}


void DeviceStitch::finish() {
  if (!started_)
    return;
  close_all();
}

void DeviceStitch::collect(size_t open_begin, size_t open_end) {
  hipStream_t s = stream_;
  epoch_++;
  if (epoch_ == 0) { // wrapped: the stamps start over
    launch_fill(stamp_open_.as<uint32_t>(), table_cap_, 0, s);
    launch_fill(stamp_claim_.as<uint32_t>(), table_cap_, 0, s);
    epoch_ = 1;
  }
  const size_t nopen_max = open_end - open_begin;
  cand_.ensure((nopen_max + 2) * 4, s);
  uint32_t *cand = cand_.as<uint32_t>();
  launch_fill(cand + nopen_max, 2, 0, s); // the count and the closed total
  if (nopen_max)
    collect_open<<<blocks_for(nopen_max, 256), 256, 0, s>>>(
        nodes_.as<Node>(), static_cast<uint32_t>(open_begin),
        static_cast<uint32_t>(open_end), stamp_claim_.as<uint32_t>(), epoch_, cand,
        cand + nopen_max);
  check(hipGetLastError(), "launching the open roots' collection");
}

void DeviceStitch::close_all() {
  if (table_end_ > 0 && pl_end_ > pl_begin_) {
    const int32_t first_rel = static_cast<int32_t>(last_frame_ + 1 - base_);
    collect(pl_begin_, pl_end_);
    process(first_rel, 0, true, table_end_, table_end_, pl_begin_, pl_end_,
            table_end_, table_end_);
  }
  // Nothing is open: the table and every staged buffer start over (the
  // stream orders any later reuse of a buffer after the gathers reading it).
  table_end_ = pl_begin_ = pl_end_ = 0;
  dead_ = 0;
  for (Staged &st : staged_)
    spare_.push_back(std::move(st.buf));
  staged_.clear();
}

void DeviceStitch::process(int first_rel, int nframes, bool flush_all,
                           size_t new_begin, size_t new_end, size_t open_begin,
                           size_t open_end, size_t last_begin, size_t last_end) {
  hipStream_t s = stream_;
  Node *nodes = nodes_.as<Node>();
  Acc *acc = acc_.as<Acc>();
  const size_t nopen_max = open_end - open_begin;
  const size_t ncand_max = nopen_max + (new_end - new_begin);
  const uint64_t T = (table_end_ + 31) / 32 * 32;
  const uint64_t buckets = flush_all ? 1 : static_cast<uint64_t>(nframes) + 1;
  const uint64_t kwords = std::max<uint64_t>(1, buckets * T / 32);
  const uint64_t twords = std::max<uint64_t>(1, T / 32);
  // Live nodes per staged buffer, by id from the oldest held.
  const uint32_t buffer_base = staged_.empty() ? 0 : staged_.front().id;
  const size_t nbuf = staged_.empty() ? 0 : staged_.back().id - buffer_base + 1;

  keybits_.ensure(kwords * 4, s);
  keyprefix_.ensure(kwords * 4, s);
  keytotals_.ensure(scan_tmp_words(kwords) * 4 + 4, s);
  membits_.ensure(twords * 4, s);
  memprefix_.ensure(twords * 4, s);
  memtotals_.ensure(scan_tmp_words(twords) * 4, s);
  alive_bits_.ensure(twords * 4, s);
  alive_prefix_.ensure(twords * 4, s);
  alive_totals_.ensure(scan_tmp_words(twords) * 4, s);
  emit_list_.ensure((ncand_max + 1) * 4, s);
  spot_tmp_.ensure((ncand_max + 1) * sizeof(sf_spot), s);
  decision_.ensure((ncand_max + 1) * 4, s);
  accidx_.ensure((ncand_max + 1) * 8, s);
  pixoff_.ensure((ncand_max + 1) * 8, s);
  spot_members_.ensure((ncand_max + 2) * 8, s);
  live_counts_.ensure((nbuf + 1) * 4, s);

  uint32_t *cand = cand_.as<uint32_t>();
  uint32_t *ncand = cand + nopen_max;    // the list's count, after the list
  uint32_t *ktotal = ncand + 1;          // the closed roots' count
  uint32_t *kb = keybits_.as<uint32_t>();
  launch_fill(kb, kwords, 0, s);
  launch_fill(membits_.as<uint32_t>(), twords, 0, s);
  launch_fill(alive_bits_.as<uint32_t>(), twords, 0, s);
  launch_fill(spot_members_.as<uint32_t>(), 2 * (ncand_max + 2), 0, s);
  launch_fill(live_counts_.as<uint32_t>(), nbuf + 1, 0, s);
  launch_fill(summary_.as<uint32_t>(), sizeof(Summary) / 4, 0, s);

  // The three bitmaps' scan descriptors: closing keys; members and alive.
  ScanSeg *hs = host_scan_.get<ScanSeg>(3);
  hs[0] = ScanSeg{kb, keyprefix_.as<uint32_t>(), keytotals_.as<uint32_t>(),
                  static_cast<size_t>(kwords), nullptr, 0, 0};
  hs[1] = ScanSeg{membits_.as<uint32_t>(), memprefix_.as<uint32_t>(),
                  memtotals_.as<uint32_t>(), static_cast<size_t>(twords), nullptr, 0, 0};
  hs[2] = ScanSeg{alive_bits_.as<uint32_t>(), alive_prefix_.as<uint32_t>(),
                  alive_totals_.as<uint32_t>(), static_cast<size_t>(twords), nullptr, 0, 0};
  scan_desc_.ensure(3 * sizeof(ScanSeg), s);
  check(hipMemcpyAsync(scan_desc_.ptr, hs, 3 * sizeof(ScanSeg), hipMemcpyHostToDevice, s),
        "uploading the grouping's scan descriptors");
  const ScanSeg *sd = scan_desc_.as<ScanSeg>();

  // Union-find results: the candidates (collected before the union) merged.
  const uint32_t nb = static_cast<uint32_t>(new_begin), ne = static_cast<uint32_t>(new_end);
  const unsigned gc = blocks_for(ncand_max, 256);
  const int fa = flush_all ? 1 : 0;
  if (ncand_max) {
    merge_sums<<<gc, 256, 0, s>>>(nodes, acc, cand, ncand, nb, ne);
    merge_peaks<<<gc, 256, 0, s>>>(nodes, acc, cand, ncand, nb, ne);
    if (!flush_all && !options_.two_d && last_end > last_begin)
      mark_open<<<blocks_for(last_end - last_begin, 256), 256, 0, s>>>(
          nodes, static_cast<uint32_t>(last_begin), static_cast<uint32_t>(last_end),
          stamp_open_.as<uint32_t>(), epoch_);
    close_roots<<<gc, 256, 0, s>>>(nodes, acc, cand, ncand, nb, ne,
                                   stamp_open_.as<uint32_t>(), epoch_, fa, first_rel,
                                   T, kb);
  }
  check(hipGetLastError(), "launching the grouping's merge");
  launch_popcount_scan(sd, 1, static_cast<size_t>(kwords), s, false);
  key_total<<<1, 1, 0, s>>>(kb, keyprefix_.as<uint32_t>(), kwords, ktotal);
  if (ncand_max)
    list_closed<<<gc, 256, 0, s>>>(nodes, acc, cand, ncand, nb, ne, fa, first_rel, T,
                                   kb, keyprefix_.as<uint32_t>(),
                                   emit_list_.as<uint32_t>());
  EmitParams ep;
  ep.width = static_cast<uint32_t>(width_);
  ep.min_size = options_.min_spot_size;
  ep.max_size = options_.max_spot_size;
  ep.max_separation = options_.max_separation;
  ep.base = base_;
  Counters *ctr = counters_.as<Counters>();
  emit_spots<<<blocks_for(ncand_max + 1, 256), 256, 0, s>>>(
      acc, emit_list_.as<uint32_t>(), ktotal, static_cast<uint32_t>(ncand_max), ep,
      spot_tmp_.as<sf_spot>(), decision_.as<uint32_t>(), accidx_.as<uint64_t>(),
      pixoff_.as<uint64_t>(), ctr);
  check(hipGetLastError(), "launching the grouping's emit");
  scan_u64(accidx_.as<uint64_t>(), ncand_max + 1, scan_tmp_, s);
  scan_u64(pixoff_.as<uint64_t>(), ncand_max + 1, scan_tmp_, s);
  Summary *sum_d = summary_.as<Summary>();
  if (table_end_)
    classify_nodes<<<blocks_for(table_end_, 256), 256, 0, s>>>(
        nodes, acc, static_cast<uint32_t>(table_end_), fa, first_rel, T, kb,
        keyprefix_.as<uint32_t>(), decision_.as<uint32_t>(), accidx_.as<uint64_t>(),
        membits_.as<uint32_t>(), alive_bits_.as<uint32_t>(),
        reinterpret_cast<unsigned long long *>(spot_members_.ptr),
        live_counts_.as<uint32_t>(), buffer_base, static_cast<uint32_t>(nbuf), sum_d);
  check(hipGetLastError(), "launching the grouping's classification");
  launch_popcount_scan(sd + 1, 2, static_cast<size_t>(twords), s, false);
  write_summary<<<1, 1, 0, s>>>(ktotal, accidx_.as<uint64_t>(), pixoff_.as<uint64_t>(),
                                membits_.as<uint32_t>(), memprefix_.as<uint32_t>(),
                                alive_bits_.as<uint32_t>(), alive_prefix_.as<uint32_t>(),
                                twords, reinterpret_cast<unsigned long long *>(error_.ptr),
                                sum_d);
  check(hipGetLastError(), "launching the grouping's summary");
  char *h = host_summary_.get<char>(sizeof(Summary) + sizeof(Counters) + 4 * (nbuf + 1));
  check(hipMemcpyAsync(h, sum_d, sizeof(Summary), hipMemcpyDeviceToHost, s),
        "reading the grouping's summary");
  check(hipMemcpyAsync(h + sizeof(Summary), ctr, sizeof(Counters), hipMemcpyDeviceToHost, s),
        "reading the grouping's counts");
  check(hipMemcpyAsync(h + sizeof(Summary) + sizeof(Counters), live_counts_.ptr,
                       4 * (nbuf + 1), hipMemcpyDeviceToHost, s),
        "reading the staged frames' live nodes");
  check(hipStreamSynchronize(s), "hipStreamSynchronize");
  Summary sm;
  Counters c;
  std::memcpy(&sm, h, sizeof sm);
  std::memcpy(&c, h + sizeof(Summary), sizeof c);
  std::vector<uint32_t> live(nbuf);
  if (nbuf)
    std::memcpy(live.data(), h + sizeof(Summary) + sizeof(Counters), 4 * nbuf);
  if (sm.error)
    throw std::runtime_error("device grouping: the labelling's components, runs "
                             "or links disagree (code " + std::to_string(sm.error) + ")");
  counts_.groups = c.groups;
  counts_.too_small = c.too_small;
  counts_.too_large = c.too_large;
  counts_.separated = c.separated;
  counts_.accepted = c.accepted;
  live_ = counts_.signal_pixels - c.closed_pixels;

  // The spots, and their pixels, into the output lists.
  const uint64_t A = sm.accepted, P = sm.pixels, M = sm.members;
  // This batch's records only, in the output set of its parity: the drain
  // copies them back while the next batch runs, and the batch after that
  // waits for the copy before it writes the set again.
  const int par = static_cast<int>(batches_ & 1);
  drain_->wait(out_ticket_[par]);
  DeviceBuffer &spots_out = spots_out_[par], &pixels_out = pixels_out_[par];
  spots_out.ensure(std::max<uint64_t>(A, 1) * sizeof(sf_spot), s);
  pixels_out.ensure(std::max<uint64_t>(P, 1) * sizeof(sf_pixel), s);
  spot_pix_.ensure((A + 1) * 8, s);
  if (A)
    write_spots<<<blocks_for(ncand_max + 1, 256), 256, 0, s>>>(
        ktotal, decision_.as<uint32_t>(), spot_tmp_.as<sf_spot>(), accidx_.as<uint64_t>(),
        pixoff_.as<uint64_t>(), pixels_dev_, spots_out.as<sf_spot>(),
        spot_pix_.as<uint64_t>());
  if (M) {
    mkeys_.ensure(M * 4, s);
    mvals_.ensure(M * 4, s);
    mkeys2_.ensure(M * 4, s);
    mvals2_.ensure(M * 4, s);
    member_pix_.ensure((M + 1) * 8, s);
    list_members<<<blocks_for(table_end_, 256), 256, 0, s>>>(
        nodes, acc, static_cast<uint32_t>(table_end_), fa, first_rel, T, kb,
        keyprefix_.as<uint32_t>(), accidx_.as<uint64_t>(), membits_.as<uint32_t>(),
        memprefix_.as<uint32_t>(), mkeys_.as<uint32_t>(), mvals_.as<uint32_t>());
    sort_pairs(mkeys_, mvals_, mkeys2_, mvals2_, static_cast<uint32_t>(M),
               bits_for(A), sort_hist_, scan_tmp_, s);
    // Each spot's first member (counts scanned over A + 1 entries).
    scan_u64(spot_members_.as<uint64_t>(), A + 1, scan_tmp_, s);
    launch_fill(member_pix_.as<uint32_t>() + 2 * M, 2, 0, s);
    member_sizes<<<blocks_for(M, 256), 256, 0, s>>>(nodes, mvals_.as<uint32_t>(),
                                                    static_cast<uint32_t>(M),
                                                    member_pix_.as<uint64_t>());
    scan_u64(member_pix_.as<uint64_t>(), M + 1, scan_tmp_, s);
    gather_members<<<blocks_for(M, 4), 256, 0, s>>>(
        nodes, mkeys_.as<uint32_t>(), mvals_.as<uint32_t>(), static_cast<uint32_t>(M),
        member_pix_.as<uint64_t>(), spot_members_.as<uint64_t>(), spot_pix_.as<uint64_t>(),
        base_, pixels_out.as<sf_pixel>());
    check(hipGetLastError(), "launching the spot pixel gather");
  }
  check(hipEventRecord(out_ready_[par], s), "hipEventRecord");
  out_ticket_[par] = drain_->submit(spots_out.as<sf_spot>(), A, pixels_out.as<sf_pixel>(),
                                    P, out_ready_[par]);
  batches_++;
  spots_dev_ += A;
  pixels_dev_ += P;
  dead_ += sm.dead;

  if (flush_all)
    return;
  // The open components' nodes: the batch's last frame, unless compacted.
  const size_t nlast = last_end - last_begin;
  if (dead_ > sm.alive + kDeadSlack && table_end_ > 0) {
    nodes_tmp_.ensure(table_cap_ * sizeof(Node), s);
    acc_tmp_.ensure(table_cap_ * sizeof(Acc), s);
    compact_nodes<<<blocks_for(table_end_, 256), 256, 0, s>>>(
        nodes, acc, static_cast<uint32_t>(table_end_), alive_bits_.as<uint32_t>(),
        alive_prefix_.as<uint32_t>(), nodes_tmp_.as<Node>(), acc_tmp_.as<Acc>());
    check(hipGetLastError(), "launching the node compaction");
    std::swap(nodes_.ptr, nodes_tmp_.ptr);
    std::swap(nodes_.bytes, nodes_tmp_.bytes);
    std::swap(acc_.ptr, acc_tmp_.ptr);
    std::swap(acc_.bytes, acc_tmp_.bytes);
    table_end_ = sm.alive;
    dead_ = 0;
    // Stamps are indexed by node: the old ones name other nodes now.
    launch_fill(stamp_open_.as<uint32_t>(), table_cap_, 0, s);
    launch_fill(stamp_claim_.as<uint32_t>(), table_cap_, 0, s);
    epoch_ = 0;
    pl_end_ = table_end_;
    pl_begin_ = table_end_ - std::min(table_end_, nlast);
  } else {
    pl_begin_ = last_begin;
    pl_end_ = last_end;
  }
  // Staged frames no live node reads go back to the pool.
  for (size_t i = 0; i < staged_.size();) {
    const size_t k = staged_[i].id - buffer_base;
    if (k < live.size() && live[k] == 0) {
      spare_.push_back(std::move(staged_[i].buf));
      staged_.erase(staged_.begin() + static_cast<std::ptrdiff_t>(i));
    } else {
      i++;
    }
  }
}

Records<sf_spot> DeviceStitch::spots() {
  drain_->wait_all();
  return {drain_->spots(), spots_dev_};
}

Records<sf_pixel> DeviceStitch::pixels() {
  drain_->wait_all();
  return {drain_->pixels(), pixels_dev_};
}

} // namespace sfi
