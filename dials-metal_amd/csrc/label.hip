// Strong-pixel connected components of each frame of a batch, and their
// exact moments, on the device.
//
// The reference groups on the host (dials_spots::Labeller::add,
// src/spots/dials_spots.cc:57-145): pixels are unioned with their left
// neighbour (same row) and the one above, toward the lower position, and a
// component's rows are emitted in order of its first pixel.  Here the same
// forest is built in the rank space of each frame's ordered signal list:
//   * the left neighbour of rank i, when set, is rank i - 1; the one above is
//     found from the strong-pixel bitmask and its word prefix;
//   * a lock-free union-find hooks the larger root under the smaller, so the
//     root of every component is its lowest rank, i.e. its first pixel;
//   * roots are numbered by a popcount scan of a root bitmask, which numbers
//     components in order of first appearance -- the reference's order;
//   * every pixel adds its exact integer moments to its component with
//     atomics (128-bit where a product can exceed 64 bits).
// Links to the previous frame (same index set in both frames) are listed for
// the host's three-dimensional stitch (dials_spots.cc:128-137).
//
// Every kernel works on frame blockIdx.y of a descriptor array, so a batch
// is labelled in the same few launches as one frame.
#include "sf_device.hh"
#include "sf_internal.hh"

namespace {

using sfi::FrameDev;

__device__ __forceinline__ uint32_t load_relaxed(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The root of i, shortening the path as it goes.  Only non-roots are written,
// and only with one of their ancestors, so a concurrent hook of a root is
// never undone.
__device__ uint32_t find(uint32_t *parent, uint32_t i) {
  uint32_t cur = load_relaxed(&parent[i]);
  if (cur == i)
    return i;
  uint32_t prev = i;
  for (;;) {
    const uint32_t next = load_relaxed(&parent[cur]);
    if (next == cur)
      return cur;
    store_relaxed(&parent[prev], next);
    prev = cur;
    cur = next;
  }
}

__device__ void unite(uint32_t *parent, uint32_t a, uint32_t b) {
  uint32_t ra = find(parent, a), rb = find(parent, b);
  while (ra != rb) {
    if (ra < rb) {
      const uint32_t t = ra;
      ra = rb;
      rb = t;
    }
    const uint32_t old = atomicCAS(&parent[ra], ra, rb);
    if (old == ra)
      return;
    ra = find(parent, old);
    rb = find(parent, rb);
  }
}

__device__ __forceinline__ bool bit_set(const uint32_t *mask, uint32_t wpr,
                                        uint32_t x, uint32_t y) {
  return (mask[static_cast<size_t>(y) * wpr + x / 32] >> (x & 31u)) & 1u;
}

__device__ __forceinline__ uint32_t rank_of(const uint32_t *mask,
                                            const uint32_t *prefix, uint32_t wpr,
                                            uint32_t x, uint32_t y) {
  const size_t w = static_cast<size_t>(y) * wpr + x / 32;
  return prefix[w] + static_cast<uint32_t>(__popc(mask[w] & ((1u << (x & 31u)) - 1u)));
}

// An overfull frame is refused, not labelled.
__device__ __forceinline__ uint32_t live_count(const FrameDev &f) {
  const uint32_t n = f.counter[0];
  return n > f.capacity ? 0u : n;
}

#define FOR_EACH_PIXEL(i, n)                                                   \
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (n);            \
       i += gridDim.x * blockDim.x)

// parent[i] = i - 1 when the left neighbour is set, else i.
__global__ __launch_bounds__(256) void link_left(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t n = live_count(f);
  FOR_EACH_PIXEL(i, n) {
    const uint32_t idx = f.ordered[i].index;
    const bool left =
        i > 0 && idx % f.width != 0 && f.ordered[i - 1].index == idx - 1;
    f.parent[i] = left ? i - 1 : i;
  }
}

// Union with the pixel above.
__global__ __launch_bounds__(256) void link_up(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t n = live_count(f);
  FOR_EACH_PIXEL(i, n) {
    const uint32_t idx = f.ordered[i].index;
    const uint32_t y = idx / f.width, x = idx - y * f.width;
    if (y > 0 && bit_set(f.mask, f.wpr, x, y - 1))
      unite(f.parent, i, rank_of(f.mask, f.prefix, f.wpr, x, y - 1));
  }
}

// Flatten every tree to its root; mark the roots in a bitmask of ranks.  A
// wave covers 64 consecutive ranks from a multiple of 64, so its ballot is
// two whole words, and the words up to the live count are all written.
__global__ __launch_bounds__(256) void flatten(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t n = live_count(f);
  const uint32_t span = (n + 63u) & ~63u;
  FOR_EACH_PIXEL(i, span) {
    bool root = false;
    if (i < n) {
      uint32_t r = load_relaxed(&f.parent[i]);
      for (;;) {
        const uint32_t up = load_relaxed(&f.parent[r]);
        if (up == r)
          break;
        r = up;
      }
      store_relaxed(&f.parent[i], r);
      root = r == i;
    }
    const uint64_t b = __ballot(root);
    const uint32_t lane = sfd::lane_id();
    if (lane == 0)
      f.roots[i / 32] = static_cast<uint32_t>(b);
    else if (lane == 32)
      f.roots[i / 32] = static_cast<uint32_t>(b >> 32);
  }
}

__device__ __forceinline__ uint32_t component_of(const FrameDev &f, uint32_t r) {
  return f.root_prefix[r / 32] +
         static_cast<uint32_t>(__popc(f.roots[r / 32] & ((1u << (r & 31u)) - 1u)));
}

// Component of every pixel; each root opens its component's record, and the
// last pixel, which knows how many roots precede it, writes the count.
__global__ __launch_bounds__(256) void number_components(
    const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t n = live_count(f);
  FOR_EACH_PIXEL(i, n) {
    const uint32_t r = f.parent[i];
    const uint32_t c = component_of(f, r);
    f.comp_of[i] = c;
    if (r == i) {
      sf_component z{};
      z.first = f.ordered[i].index;
      z.x0 = z.y0 = 0xFFFFFFFFu;
      f.comps[c] = z;
    }
    if (i == n - 1)
      f.counter[2] = component_of(f, i) + (r == i ? 1u : 0u);
  }
}

__device__ __forceinline__ void add128(uint64_t *pair, uint64_t lo, uint64_t hi) {
  const unsigned long long old =
      atomicAdd(reinterpret_cast<unsigned long long *>(&pair[0]),
                static_cast<unsigned long long>(lo));
  const uint64_t carry = (static_cast<uint64_t>(old) + lo) < lo ? 1u : 0u;
  if (hi + carry)
    atomicAdd(reinterpret_cast<unsigned long long *>(&pair[1]),
              static_cast<unsigned long long>(hi + carry));
}

// Adds a 128-bit value held as (lo, hi) to a running (lo, hi) in registers.
__device__ __forceinline__ void acc128(uint64_t &lo, uint64_t &hi, uint64_t l,
                                       uint64_t h) {
  const uint64_t s = lo + l;
  hi += h + (s < lo ? 1u : 0u);
  lo = s;
}

// Exact moments into the components, one row-run at a time: the pixel that
// starts a run (no set pixel to its left) sums the run in registers, then
// adds the run to its component with one set of atomics.  A run lies in one
// row and one component, and runs are short, so the component records see a
// few atomics per row of a spot rather than one set per pixel.
__global__ __launch_bounds__(256) void accumulate(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t n = live_count(f);
  FOR_EACH_PIXEL(i, n) {
    const uint32_t first = f.ordered[i].index;
    const uint32_t y = first / f.width, x0 = first - y * f.width;
    if (x0 > 0 && i > 0 && f.ordered[i - 1].index == first - 1)
      continue; // not the start of a run
    const uint64_t ay = 2ull * y + 1ull;
    uint32_t count = 0, x = x0;
    uint64_t peak = 0, sum = 0;
    uint64_t sq_lo = 0, sq_hi = 0, sx_lo = 0, sx_hi = 0, xx_lo = 0, xx_hi = 0;
    uint32_t j = i;
    for (;;) {
      const uint64_t v = f.ordered[j].value;
      const uint64_t ax = 2ull * x + 1ull;
      const uint64_t key = (v << 32) | (0xFFFFFFFFu - (first + count));
      peak = key > peak ? key : peak;
      sum += v;
      acc128(sq_lo, sq_hi, v * v, 0);
      acc128(sx_lo, sx_hi, v * ax, 0);
      const uint64_t xx = ax * ax;
      acc128(xx_lo, xx_hi, v * xx, __umul64hi(v, xx));
      count++;
      j++;
      x++;
      if (j >= n || x >= f.width || f.ordered[j].index != first + count)
        break;
    }
    sf_component *c = &f.comps[f.comp_of[i]];
    atomicAdd(&c->n, count);
    atomicMin(&c->x0, x0);
    atomicMax(&c->x1, x0 + count);
    atomicMin(&c->y0, y);
    atomicMax(&c->y1, y + 1);
    atomicMax(reinterpret_cast<unsigned long long *>(&c->peak),
              static_cast<unsigned long long>(peak));
    atomicAdd(reinterpret_cast<unsigned long long *>(&c->sum),
              static_cast<unsigned long long>(sum));
    add128(c->sum_sq, sq_lo, sq_hi);
    add128(c->sum_x, sx_lo, sx_hi);
    // sum of value * ay and value * ay^2 over the run: ay is the run's
    // constant, so these are products of the run's sum.
    add128(c->sum_y, sum * ay, __umul64hi(sum, ay));
    const uint64_t yy = ay * ay;
    add128(c->sum_yy, sum * yy, __umul64hi(sum, yy));
    add128(c->sum_xx, xx_lo, xx_hi);
  }
}

// Pairs (component here, component on the previous frame) for every pixel
// set in both frames; a pixel repeating its left neighbour's pair is skipped.
__global__ __launch_bounds__(256) void link_frames(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  if (f.prev_mask == nullptr)
    return;
  const uint32_t n = live_count(f);
  FOR_EACH_PIXEL(i, n) {
    const uint32_t idx = f.ordered[i].index;
    const uint32_t y = idx / f.width, x = idx - y * f.width;
    bool emit = false;
    uint32_t here = 0, there = 0;
    if (bit_set(f.prev_mask, f.wpr, x, y)) {
      here = f.comp_of[i];
      there = f.prev_comp_of[rank_of(f.prev_mask, f.prev_prefix, f.wpr, x, y)];
      emit = !(x > 0 && i > 0 && f.ordered[i - 1].index == idx - 1 &&
               f.comp_of[i - 1] == here && bit_set(f.prev_mask, f.wpr, x - 1, y) &&
               f.prev_comp_of[rank_of(f.prev_mask, f.prev_prefix, f.wpr, x - 1,
                                      y)] == there);
    }
    // One atomic on the frame's link count per wave, not per link.
    const uint64_t b = __ballot(emit);
    if (b == 0)
      continue;
    const uint32_t lane = sfd::lane_id();
    const int leader = __ffsll(static_cast<unsigned long long>(b)) - 1;
    uint32_t base = 0;
    if (static_cast<int>(lane) == leader)
      base = atomicAdd(&f.counter[3], static_cast<uint32_t>(__popcll(b)));
    base = __shfl(base, leader);
    if (emit) {
      const uint32_t at =
          base + static_cast<uint32_t>(__popcll(b & ((1ull << lane) - 1ull)));
      if (at < f.link_capacity) {
        f.links[2 * at] = here;
        f.links[2 * at + 1] = there;
      }
    }
  }
}

} // namespace

namespace sfi {

// frames_device: the frames' descriptors; root_segs_device: their root-bit
// scans; both in device memory, nframes long.  counter[2] and counter[3] of
// every frame are cleared by the caller.
void launch_label(const FrameDev *frames_device, const ScanSeg *root_segs_device,
                  uint32_t nframes, size_t max_root_words, hipStream_t stream) {
  if (nframes == 0)
    return;
  const dim3 grid(256, nframes);
  link_left<<<grid, 256, 0, stream>>>(frames_device);
  link_up<<<grid, 256, 0, stream>>>(frames_device);
  flatten<<<grid, 256, 0, stream>>>(frames_device);
  launch_popcount_scan(root_segs_device, nframes, max_root_words, stream);
  number_components<<<grid, 256, 0, stream>>>(frames_device);
  accumulate<<<grid, 256, 0, stream>>>(frames_device);
  link_frames<<<grid, 256, 0, stream>>>(frames_device);
  check(hipGetLastError(), "launching the labelling");
}

} // namespace sfi
