// Strong-pixel connected components of each frame of a batch, and their
// exact moments, on the device.
//
// The reference groups on the host (dials_spots::Labeller::add,
// src/spots/dials_spots.cc:57-145): pixels are unioned with their left
// neighbour (same row) and the one above, toward the lower position, and a
// component's rows are emitted in order of its first pixel.  Here the same
// components are built from row runs -- maximal horizontal segments of set
// bits in the strong-pixel bitmask -- in the rank space of each frame's
// ordered signal list:
//   * one wave per bitmask row counts the row's runs (a run starts at a set
//     bit whose left neighbour is clear), a scan over the rows places each
//     row's runs in the frame's run list, and a second pass lists each run
//     {first rank, x0, y, length} (no atomic: a returning atomic per row on
//     one counter serialises the rows behind it);
//   * each run, named by its first rank, is unioned with the runs above it
//     that it touches, found as segments of the bitmask row above over its
//     columns and traced left in that row to their runs' first pixels; a
//     lock-free union-find hooks the larger root under the smaller, so the
//     root of a component is its lowest rank, i.e. its first pixel;
//   * roots are numbered by a popcount scan of a root bitmask, which numbers
//     components in order of first appearance -- the reference's order;
//   * the runs are sorted by component (one atomic per run takes its place,
//     a scan of the counts gives each component's first), and a group of
//     eight lanes per component sums its runs' pixels' exact integer moments
//     in registers (128-bit where a product can exceed 64 bits), adds the
//     lanes' sums by lane shuffles and writes the record whole -- no atomics
//     on the records -- and its pixels' component.
// Links to the previous frame (pixels at the same position in both frames)
// are listed per run, as the components of the previous frame's bitmask
// segments under the run, for the host's three-dimensional stitch
// (dials_spots.cc:128-137).  No step divides a pixel index by the width: rows
// and columns come from the bitmask's layout.
//
// Every kernel works on frame blockIdx.y of a descriptor array, so a batch
// is labelled in the same few launches as one frame.
#include "sf_device.hh"
#include "sf_internal.hh"

namespace {

using sfd::gmem;
using sfi::FrameDev;

constexpr int kLanes = 64;

// A run record {first rank, x0, y, length} as the kernels move it.
typedef uint32_t run4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ SF_GLOBAL run4 *run_list(const FrameDev &f) {
  return reinterpret_cast<SF_GLOBAL run4 *>(gmem(f.runs));
}

__device__ __forceinline__ uint32_t load_relaxed(SF_GLOBAL uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(SF_GLOBAL uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The root of i, shortening the path as it goes.  Only non-roots are written,
// and only with one of their ancestors, so a concurrent hook of a root is
// never undone.
__device__ uint32_t find(SF_GLOBAL uint32_t *parent, uint32_t i) {
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

__device__ void unite(SF_GLOBAL uint32_t *parent, uint32_t a, uint32_t b) {
  uint32_t ra = find(parent, a), rb = find(parent, b);
  while (ra != rb) {
    if (ra < rb) {
      const uint32_t t = ra;
      ra = rb;
      rb = t;
    }
    uint32_t expected = ra;
    if (__hip_atomic_compare_exchange_strong(&parent[ra], &expected, rb,
                                             __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
    ra = find(parent, expected);
    rb = find(parent, rb);
  }
}

// An overfull frame is refused, not labelled.
__device__ __forceinline__ uint32_t live_count(const FrameDev &f) {
  const uint32_t n = f.counter[0];
  return n > f.capacity ? 0u : n;
}

// Bits of bitmask word j that fall in columns [x0, x1).
__device__ __forceinline__ uint32_t columns(uint32_t j, uint32_t x0, uint32_t x1) {
  const uint32_t base = 32u * j;
  const uint32_t lo = x0 > base ? x0 - base : 0u;
  const uint32_t hi = x1 < base + 32u ? x1 - base : 32u;
  const uint32_t upto = hi == 32u ? 0xFFFFFFFFu : (1u << hi) - 1u;
  return upto & ~((1u << lo) - 1u);
}

// Wave-wide exclusive prefix of v and the wave's total (every lane active).
__device__ __forceinline__ uint32_t wave_prefix(uint32_t v, uint32_t lane,
                                                uint32_t *total) {
  (void)lane;
  const uint32_t incl = sfd::wave_inclusive(v);
  *total = sfd::last_lane(incl);
  return incl - v;
}

// Bitmask words a lane holds at once in count_runs and emit_runs: a row of
// up to 64 x kRowChunks words (6144 columns) in one round of loads, wider
// rows in several.
constexpr uint32_t kRowChunks = 3;

// Lane shifts across the whole wave (DPP wave_shr:1 / wave_shl:1 with
// bound_ctrl): lane i gets lane i-1's value (lane 0 gets 0), or lane i+1's
// (lane 63 gets 0).
__device__ __forceinline__ uint32_t from_left(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ uint32_t from_right(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ uint32_t lane_value(uint32_t v, uint32_t l) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(l)));
}

// Run starts in a round of a bitmask row held one word a lane per chunk
// (lane l, chunk c: word j0 + 64 c + l): a run starts at a set bit whose left
// neighbour in the row is clear.  `carry_in` is bit 31 of the word before
// the round's first, `left_of` gives it for the next round.
struct RowRound {
  uint32_t m[kRowChunks];
  __device__ __forceinline__ uint32_t starts(uint32_t c, uint32_t lane,
                                             uint32_t carry_in) const {
    uint32_t left = from_left(m[c]);
    const uint32_t first_left = c == 0 ? carry_in : lane_value(m[c - 1], kLanes - 1);
    if (lane == 0)
      left = first_left;
    return m[c] & ~((m[c] << 1) | (left >> 31));
  }
  __device__ __forceinline__ uint32_t left_of() const {
    return lane_value(m[kRowChunks - 1], kLanes - 1);
  }
};

// Runs per bitmask row of every frame, one wave per row, the row's words
// loaded in one round.
__global__ __launch_bounds__(256) void count_runs(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  if (live_count(f) == 0)
    return;
  const uint32_t lane = threadIdx.x % kLanes;
  const uint32_t per_block = blockDim.x / kLanes;
  const uint32_t wpr = f.wpr;
  for (uint32_t y = blockIdx.x * per_block + threadIdx.x / kLanes; y < f.height;
       y += gridDim.x * per_block) {
    const SF_GLOBAL uint32_t *row = gmem(f.mask) + static_cast<size_t>(y) * wpr;
    uint32_t count = 0, carry_in = 0;
    for (uint32_t j0 = 0; j0 < wpr; j0 += kLanes * kRowChunks) {
      RowRound r;
#pragma unroll
      for (uint32_t c = 0; c < kRowChunks; c++) {
        const uint32_t j = j0 + c * kLanes + lane;
        r.m[c] = j < wpr ? row[j] : 0u;
      }
#pragma unroll
      for (uint32_t c = 0; c < kRowChunks; c++)
        count += static_cast<uint32_t>(__popc(r.starts(c, lane, carry_in)));
      carry_in = r.left_of();
    }
    uint32_t total;
    (void)wave_prefix(count, lane, &total);
    if (lane == 0)
      gmem(f.row_runs)[y] = total;
  }
}

// Each row's first run (an exclusive scan of the runs per row, in place) and
// the frame's run count; one block of 1024 threads per frame, each thread a
// contiguous share of the rows.
__global__ __launch_bounds__(1024) void scan_runs(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.x];
  __shared__ uint32_t wave_sums[1024 / kLanes];
  // The components are counted later; the links accumulate from zero.
  if (threadIdx.x == 0) {
    f.counter[2] = 0;
    f.counter[3] = 0;
  }
  if (live_count(f) == 0) {
    if (threadIdx.x == 0)
      f.counter[4] = 0;
    return;
  }
  SF_GLOBAL uint32_t *rows = gmem(f.row_runs);
  const uint32_t per = (f.height + blockDim.x - 1) / blockDim.x;
  const uint32_t y0 = threadIdx.x * per;
  const uint32_t y1 = min(f.height, y0 + per);
  uint32_t mine = 0;
  for (uint32_t y = y0; y < y1; y++)
    mine += rows[y];
  const uint32_t lane = threadIdx.x % kLanes, wave = threadIdx.x / kLanes;
  uint32_t total;
  uint32_t before = wave_prefix(mine, lane, &total);
  if (lane == 0)
    wave_sums[wave] = total;
  __syncthreads();
  uint32_t all = 0;
  for (uint32_t w = 0; w < blockDim.x / kLanes; w++) {
    if (w < wave)
      before += wave_sums[w];
    all += wave_sums[w];
  }
  for (uint32_t y = y0; y < y1; y++) {
    const uint32_t c = rows[y];
    rows[y] = before;
    before += c;
  }
  if (threadIdx.x == 0)
    f.counter[4] = all;
  // The root bitmask's live words start clear: flatten sets the roots' bits.
  const uint32_t root_words = (live_count(f) + 31u) / 32u;
  for (uint32_t w = threadIdx.x; w < root_words; w += blockDim.x)
    gmem(f.roots)[w] = 0;
}

// The runs of every frame, one wave per bitmask row, each row's at its place
// in the frame's run list; each run's first rank is a root of the
// union-find's initial forest (the union-find works on runs alone, each
// named by its first rank).  A row's base, words and rank prefixes are
// loaded in one round (lane l holds words l, l + 64, ...), neighbouring
// words come from the neighbouring lanes, and only a run that crosses a whole
// word reads the row again.
__global__ __launch_bounds__(256) void emit_runs(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  if (live_count(f) == 0)
    return;
  const uint32_t lane = threadIdx.x % kLanes;
  const uint32_t per_block = blockDim.x / kLanes;
  const uint32_t wpr = f.wpr;
  const SF_GLOBAL uint32_t *bases = gmem(f.row_runs);
  SF_GLOBAL uint32_t *parent = gmem(f.parent);
  SF_GLOBAL run4 *runs = run_list(f);
  for (uint32_t y = blockIdx.x * per_block + threadIdx.x / kLanes; y < f.height;
       y += gridDim.x * per_block) {
    // (The row's base is read with its words, in the same round: a row
    // without runs only costs that round.)
    uint32_t k = bases[y];
    const SF_GLOBAL uint32_t *row = gmem(f.mask) + static_cast<size_t>(y) * wpr;
    const SF_GLOBAL uint32_t *pre = gmem(f.prefix) + static_cast<size_t>(y) * wpr;
    uint32_t carry_in = 0; // bit 31 of the word before this round's first
    for (uint32_t j0 = 0; j0 < wpr; j0 += kLanes * kRowChunks) {
      RowRound rr;
      uint32_t p[kRowChunks];
#pragma unroll
      for (uint32_t c = 0; c < kRowChunks; c++) {
        const uint32_t j = j0 + c * kLanes + lane;
        rr.m[c] = j < wpr ? row[j] : 0u;
        p[c] = j < wpr ? pre[j] : 0u;
      }
      const uint32_t(&m)[kRowChunks] = rr.m;
#pragma unroll
      for (uint32_t c = 0; c < kRowChunks; c++) {
        const uint32_t j = j0 + c * kLanes + lane;
        uint32_t right = from_right(m[c]);
        if (c + 1 < kRowChunks) {
          const uint32_t next_first = lane_value(m[c + 1], 0);
          if (lane == kLanes - 1)
            right = next_first;
        }
        uint32_t starts = rr.starts(c, lane, carry_in);
        uint32_t total;
        uint32_t at = k + wave_prefix(static_cast<uint32_t>(__popc(starts)), lane, &total);
        while (starts) {
          const uint32_t b = static_cast<uint32_t>(__ffs(starts)) - 1u;
          starts &= starts - 1u;
          const uint32_t r0 = p[c] + static_cast<uint32_t>(__popc(m[c] & ((1u << b) - 1u)));
          // The run's length: set bits from b on, into the row's next words.
          const uint32_t w = ~(m[c] >> b);
          uint32_t len = w == 0u ? 32u : static_cast<uint32_t>(__ffs(w)) - 1u;
          if (len >= 32u - b) {
            len = 32u - b;
            // The next word is the right neighbour's, except past this
            // round's last chunk, where it is read again.
            uint32_t jj = j + 1;
            uint32_t nw = (c + 1 < kRowChunks || lane + 1 < kLanes)
                              ? right
                              : (jj < wpr ? row[jj] : 0u);
            for (;;) {
              const uint32_t t = ~nw == 0u ? 32u : static_cast<uint32_t>(__ffs(~nw)) - 1u;
              len += t;
              if (t < 32u || ++jj >= wpr)
                break;
              nw = row[jj];
            }
          }
          if (at < f.capacity) // always: runs never outnumber pixels
            runs[at] = run4{r0, 32u * j + b, y, len};
          at++;
          parent[r0] = r0;
        }
        k += total;
      }
      carry_in = rr.left_of();
    }
  }
}

// Rank of the first pixel of the run that holds bit b of word j of bitmask
// row `row` (a set bit; `full` is that word), `pre` the row's rank prefixes:
// after the last clear bit before it, found in this word or, for a run that
// reaches the word's bit 0, in the words to its left.
__device__ uint32_t run_start_rank(const SF_GLOBAL uint32_t *row,
                                   const SF_GLOBAL uint32_t *pre, uint32_t j,
                                   uint32_t b, uint32_t full) {
  const uint32_t clear = ~full & ((1u << b) - 1u);
  if (clear != 0) {
    const uint32_t s = 32u - static_cast<uint32_t>(__clz(clear));
    return pre[j] + static_cast<uint32_t>(__popc(full & ((1u << s) - 1u)));
  }
  for (; j > 0; j--) {
    const uint32_t w = row[j - 1];
    if (w == 0xFFFFFFFFu)
      continue;
    if ((w >> 31) == 0) // the run starts at bit 0 of word j
      return pre[j];
    const uint32_t s = 32u - static_cast<uint32_t>(__clz(~w));
    return pre[j - 1] + static_cast<uint32_t>(__popc(w & ((1u << s) - 1u)));
  }
  return pre[0]; // the run starts the row
}

// Unites each run with the runs above it that share a column with it: one
// union per segment of the row above within the run's columns, with the run
// that holds the segment.
__global__ __launch_bounds__(256) void unite_runs(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t nr = f.counter[4];
  SF_GLOBAL uint32_t *parent = gmem(f.parent);
  const SF_GLOBAL run4 *runs = run_list(f);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nr;
       k += gridDim.x * blockDim.x) {
    const run4 r = runs[k]; // r0, x0, y, length
    if (r.z == 0)
      continue;
    const uint32_t x1 = r.y + r.w;
    const size_t above = static_cast<size_t>(r.z - 1) * f.wpr;
    const SF_GLOBAL uint32_t *row = gmem(f.mask) + above;
    const SF_GLOBAL uint32_t *pre = gmem(f.prefix) + above;
    uint32_t carry = 0;
    for (uint32_t j = r.y >> 5; j <= (x1 - 1u) >> 5; j++) {
      const uint32_t full = row[j];
      const uint32_t a = full & columns(j, r.y, x1);
      uint32_t segs = a & ~((a << 1) | carry);
      carry = a >> 31;
      while (segs) {
        const uint32_t b = static_cast<uint32_t>(__ffs(segs)) - 1u;
        segs &= segs - 1u;
        unite(parent, r.x, run_start_rank(row, pre, j, b, full));
      }
    }
  }
}

// Flatten every run's tree to its root and mark the roots (runs that are
// their own root, each the first pixel of its component) in the bitmask of
// ranks, one atomic per root.  Run k also clears the kth count of runs per
// component: components never outnumber runs.
__global__ __launch_bounds__(256) void flatten(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t nr = f.counter[4];
  SF_GLOBAL uint32_t *parent = gmem(f.parent);
  const SF_GLOBAL run4 *runs = run_list(f);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nr;
       k += gridDim.x * blockDim.x) {
    const uint32_t i = runs[k].x;
    gmem(f.comp_runs)[k] = 0;
    uint32_t r = load_relaxed(&parent[i]);
    for (;;) {
      const uint32_t up = load_relaxed(&parent[r]);
      if (up == r)
        break;
      r = up;
    }
    if (r == i)
      __hip_atomic_fetch_or(gmem(f.roots) + i / 32, 1u << (i & 31u),
                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
      store_relaxed(&parent[i], r);
  }
}

// Component numbers: an exclusive popcount scan of each frame's root bitmask
// over its live words, one block of 1024 threads a frame, each thread four
// consecutive words a round.
__global__ __launch_bounds__(1024) void number_roots(const FrameDev *__restrict__ frames) {
  __shared__ uint32_t wave_sums[1024 / kLanes];
  const FrameDev f = frames[blockIdx.x];
  const uint32_t words = (live_count(f) + 31u) / 32u;
  const SF_GLOBAL uint32_t *roots = gmem(f.roots);
  SF_GLOBAL uint32_t *prefix = gmem(f.root_prefix);
  const uint32_t lane = threadIdx.x % kLanes, wave = threadIdx.x / kLanes;
  uint32_t carry = 0;
  for (uint32_t w0 = 0; w0 < words; w0 += 4u * blockDim.x) {
    const uint32_t w = w0 + 4u * threadIdx.x;
    uint32_t c[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++)
      c[k] = w + k < words ? static_cast<uint32_t>(__popc(roots[w + k])) : 0u;
    const uint32_t mine = (c[0] + c[1]) + (c[2] + c[3]);
    uint32_t total;
    uint32_t run = wave_prefix(mine, lane, &total);
    if (lane == 0)
      wave_sums[wave] = total;
    __syncthreads();
    uint32_t all = 0;
    for (uint32_t v = 0; v < blockDim.x / kLanes; v++) {
      run += v < wave ? wave_sums[v] : 0u;
      all += wave_sums[v];
    }
    __syncthreads();
    run += carry;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      if (w + k < words)
        prefix[w + k] = run;
      run += c[k];
    }
    carry += all;
  }
}

// Roots strictly before rank r: the number of the component rooted at r.
__device__ __forceinline__ uint32_t component_of(const FrameDev &f, uint32_t r) {
  return gmem(f.root_prefix)[r / 32] +
         static_cast<uint32_t>(__popc(gmem(f.roots)[r / 32] & ((1u << (r & 31u)) - 1u)));
}

// Each run takes a place among its component's runs, one atomic per run (a
// component's few runs contend, nothing else does); its component and place
// wait in the link list's memory, which is free until the links are listed.
// The run holding the last rank writes the component count.
__global__ __launch_bounds__(256) void count_component_runs(
    const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t nr = f.counter[4];
  const uint32_t n = live_count(f);
  const SF_GLOBAL uint32_t *parent = gmem(f.parent);
  const SF_GLOBAL run4 *runs = run_list(f);
  SF_GLOBAL uint32_t *scratch = gmem(f.links);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nr;
       k += gridDim.x * blockDim.x) {
    const run4 r = runs[k];
    if (r.x + r.w == n)
      f.counter[2] = component_of(f, n - 1) +
                     ((gmem(f.roots)[(n - 1) / 32] >> ((n - 1) & 31u)) & 1u);
    const uint32_t c = component_of(f, parent[r.x]);
    scratch[2 * k] = c;
    scratch[2 * k + 1] = __hip_atomic_fetch_add(gmem(f.comp_runs) + c, 1u,
                                                __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Each component's first run in component order -- an exclusive scan of the
// runs per component, in place -- and the run count after the last; one block
// of 1024 threads per frame, each thread a contiguous share.
__global__ __launch_bounds__(1024) void scan_components(
    const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.x];
  __shared__ uint32_t wave_sums[1024 / kLanes];
  const uint32_t nc = f.counter[2];
  SF_GLOBAL uint32_t *counts = gmem(f.comp_runs);
  const uint32_t per = (nc + blockDim.x - 1) / blockDim.x;
  const uint32_t c0 = threadIdx.x * per;
  const uint32_t c1 = min(nc, c0 + per);
  uint32_t mine = 0;
  for (uint32_t c = c0; c < c1; c++)
    mine += counts[c];
  const uint32_t lane = threadIdx.x % kLanes, wave = threadIdx.x / kLanes;
  uint32_t total;
  uint32_t before = wave_prefix(mine, lane, &total);
  if (lane == 0)
    wave_sums[wave] = total;
  __syncthreads();
  uint32_t all = 0;
  for (uint32_t w = 0; w < blockDim.x / kLanes; w++) {
    if (w < wave)
      before += wave_sums[w];
    all += wave_sums[w];
  }
  for (uint32_t c = c0; c < c1; c++) {
    const uint32_t k = counts[c];
    counts[c] = before;
    before += k;
  }
  if (threadIdx.x == 0)
    counts[nc] = all;
}

// The runs in component order, in the parent array's memory (the forest is
// not needed once every run knows its component).
__global__ __launch_bounds__(256) void place_runs(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t nr = f.counter[4];
  const SF_GLOBAL uint32_t *scratch = gmem(f.links);
  const SF_GLOBAL uint32_t *first = gmem(f.comp_runs);
  SF_GLOBAL uint32_t *order = gmem(f.parent);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nr;
       k += gridDim.x * blockDim.x) {
    const uint32_t at = first[scratch[2 * k]] + scratch[2 * k + 1];
    if (at < nr) // always: the places are a permutation of the runs
      order[at] = k;
  }
}

// Adds a 128-bit value held as (lo, hi) to a running (lo, hi) in registers.
__device__ __forceinline__ void acc128(uint64_t &lo, uint64_t &hi, uint64_t l,
                                       uint64_t h) {
  const uint64_t s = lo + l;
  hi += h + (s < lo ? 1u : 0u);
  lo = s;
}

// Lanes that sum one component together: a component has up to a few tens
// of runs of a few pixels each.
constexpr uint32_t kGroup = 8;
// Pixel values a lane loads at once.
constexpr uint32_t kPixelBatch = 4;

// Totals over the kGroup lanes of a group (lanes g .. g + kGroup - 1, g a
// multiple of kGroup), every lane of the group getting them.
__device__ __forceinline__ void group_add(uint64_t &lo, uint64_t &hi) {
#pragma unroll
  for (uint32_t d = 1; d < kGroup; d <<= 1)
    acc128(lo, hi, __shfl_xor(lo, static_cast<int>(d)),
           __shfl_xor(hi, static_cast<int>(d)));
}
template <typename V, typename Op>
__device__ __forceinline__ V group_reduce(V v, Op op) {
#pragma unroll
  for (uint32_t d = 1; d < kGroup; d <<= 1)
    v = op(v, static_cast<V>(__shfl_xor(v, static_cast<int>(d))));
  return v;
}

// Exact moments, a group of kGroup lanes per component: each lane sums every
// kGroup-th of the component's runs (in component order) and their pixels
// (consecutive ranks of the ordered list) in registers, the group adds the
// lanes' sums, and its first lane writes the record whole -- no atomics --
// while every lane writes its pixels' component.
__global__ __launch_bounds__(256) void accumulate(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t nc = f.counter[2];
  const SF_GLOBAL uint32_t *first_run = gmem(f.comp_runs);
  const SF_GLOBAL uint32_t *order = gmem(f.parent);
  const SF_GLOBAL run4 *runs = run_list(f);
  const SF_GLOBAL sf_signal_pixel *ordered = gmem(f.ordered);
  SF_GLOBAL uint32_t *comp_of = gmem(f.comp_of);
  const uint32_t sub = threadIdx.x % kGroup;
  const uint32_t groups = gridDim.x * blockDim.x / kGroup;
  // c is the same for every lane of a group, so groups stay whole.
  for (uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) / kGroup; c < nc;
       c += groups) {
    uint32_t n = 0, first = 0xFFFFFFFFu, x0 = 0xFFFFFFFFu, x1 = 0,
             y0 = 0xFFFFFFFFu, y1 = 0;
    uint64_t peak = 0, sum = 0;
    uint64_t sq[2] = {0, 0}, sx[2] = {0, 0}, sy[2] = {0, 0}, sxx[2] = {0, 0},
             syy[2] = {0, 0};
    const uint32_t end = first_run[c + 1];
    uint32_t j = first_run[c] + sub;
    run4 next = {0u, 0u, 0u, 0u};
    if (j < end)
      next = runs[order[j]];
    for (; j < end; j += kGroup) {
      const run4 r = next; // r0, x0, y, length
      if (j + kGroup < end) // the lane's next run, loaded during this one
        next = runs[order[j + kGroup]];
      const uint32_t start = r.z * f.width + r.y;
      first = min(first, start);
      x0 = min(x0, r.y);
      x1 = max(x1, r.y + r.w);
      y0 = min(y0, r.z);
      y1 = max(y1, r.z + 1);
      n += r.w;
      uint64_t run_sum = 0;
      // The run's values, kPixelBatch loads in flight at a time.
      for (uint32_t p0 = 0; p0 < r.w; p0 += kPixelBatch) {
        uint32_t vals[kPixelBatch];
#pragma unroll
        for (uint32_t k = 0; k < kPixelBatch; k++)
          vals[k] = p0 + k < r.w ? ordered[r.x + p0 + k].value : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kPixelBatch; k++) {
          const uint32_t p = p0 + k;
          if (p >= r.w)
            break;
          comp_of[r.x + p] = c;
          const uint64_t v = vals[k];
          const uint64_t ax = 2ull * (r.y + p) + 1ull;
          const uint64_t key = (v << 32) | (0xFFFFFFFFu - (start + p));
          peak = key > peak ? key : peak;
          run_sum += v;
          acc128(sq[0], sq[1], v * v, 0);
          acc128(sx[0], sx[1], v * ax, 0);
          const uint64_t xx = ax * ax;
          acc128(sxx[0], sxx[1], v * xx, __umul64hi(v, xx));
        }
      }
      sum += run_sum;
      // sum of value * ay and value * ay^2 over the run: ay is the run's
      // constant, so these are products of the run's sum.
      const uint64_t ay = 2ull * r.z + 1ull;
      acc128(sy[0], sy[1], run_sum * ay, __umul64hi(run_sum, ay));
      const uint64_t yy = ay * ay;
      acc128(syy[0], syy[1], run_sum * yy, __umul64hi(run_sum, yy));
    }
    auto add = [](auto a, auto b) { return a + b; };
    auto lo = [](auto a, auto b) { return a < b ? a : b; };
    auto hi = [](auto a, auto b) { return a > b ? a : b; };
    n = group_reduce(n, add);
    sum = group_reduce(sum, add);
    peak = group_reduce(peak, hi);
    first = group_reduce(first, lo);
    x0 = group_reduce(x0, lo);
    x1 = group_reduce(x1, hi);
    y0 = group_reduce(y0, lo);
    y1 = group_reduce(y1, hi);
    group_add(sq[0], sq[1]);
    group_add(sx[0], sx[1]);
    group_add(sy[0], sy[1]);
    group_add(sxx[0], sxx[1]);
    group_add(syy[0], syy[1]);
    if (sub == 0) {
      sf_component z;
      z.first = first;
      z.n = n;
      z.x0 = x0;
      z.x1 = x1;
      z.y0 = y0;
      z.y1 = y1;
      z.peak = peak;
      z.sum = sum;
      z.sum_sq[0] = sq[0], z.sum_sq[1] = sq[1];
      z.sum_x[0] = sx[0], z.sum_x[1] = sx[1];
      z.sum_y[0] = sy[0], z.sum_y[1] = sy[1];
      z.sum_xx[0] = sxx[0], z.sum_xx[1] = sxx[1];
      z.sum_yy[0] = syy[0], z.sum_yy[1] = syy[1];
      f.comps[c] = z;
    }
  }
}

// Links a lane keeps in registers while its run is walked.
constexpr uint32_t kLinksHeld = 4;

// Pairs (component here, component on the previous frame) for the pixels set
// in both frames: per run, the previous frame's components along its columns,
// in column order, a component met again straight after itself listed once
// (the pairs feed unions, for which a repeat is harmless).  One atomic per
// wave on the frame's link count.
__global__ __launch_bounds__(256) void link_frames(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  if (f.prev_mask == nullptr)
    return;
  // A previous frame refused as overfull has no component list to link to.
  const uint32_t prev_n = f.prev_counter[0];
  if (prev_n > f.prev_capacity)
    return;
  const uint32_t nr = f.counter[4];
  const uint32_t lane = threadIdx.x % kLanes;
  const SF_GLOBAL run4 *runs = run_list(f);
  const SF_GLOBAL uint32_t *comp_of = gmem(f.comp_of);
  const SF_GLOBAL uint32_t *prev_comp = gmem(f.prev_comp_of);
  // The components found along the run, consecutive repeats dropped, from
  // its columns of the previous frame's bitmask row; `emit(there)` is called
  // for each.
  auto walk = [&](const run4 &r, auto emit) {
    const size_t off = static_cast<size_t>(r.z) * f.wpr;
    const SF_GLOBAL uint32_t *row = gmem(f.prev_mask) + off;
    const SF_GLOBAL uint32_t *pre = gmem(f.prev_prefix) + off;
    const uint32_t x1 = r.y + r.w;
    uint32_t carry = 0, last = 0xFFFFFFFFu;
    for (uint32_t j = r.y >> 5; j <= (x1 - 1u) >> 5; j++) {
      const uint32_t full = row[j];
      const uint32_t base = pre[j]; // read beside the word, used if it is set
      const uint32_t a = full & columns(j, r.y, x1);
      uint32_t segs = a & ~((a << 1) | carry);
      carry = a >> 31;
      while (segs) {
        const uint32_t b = static_cast<uint32_t>(__ffs(segs)) - 1u;
        segs &= segs - 1u;
        const uint32_t q = base + static_cast<uint32_t>(__popc(full & ((1u << b) - 1u)));
        const uint32_t there = q < prev_n ? prev_comp[q] : last;
        if (there != last)
          emit(there);
        last = there;
      }
    }
  };
  // Wave-uniform steps over the runs, so that whole waves meet at the scan.
  for (uint32_t step = blockIdx.x * blockDim.x + (threadIdx.x & ~(kLanes - 1u));
       step < nr; step += gridDim.x * blockDim.x) {
    const uint32_t k = step + lane;
    const bool active = k < nr;
    run4 r = {0u, 0u, 0u, 0u};
    if (active)
      r = runs[k];
    // One walk finds the run's links and keeps the first kLinksHeld in
    // registers; only a run with more walks again to write them.
    uint32_t here = 0, count = 0;
    uint32_t held[kLinksHeld];
#pragma unroll
    for (uint32_t i = 0; i < kLinksHeld; i++)
      held[i] = 0;
    if (active) {
      here = comp_of[r.x];
      walk(r, [&](uint32_t there) {
#pragma unroll
        for (uint32_t i = 0; i < kLinksHeld; i++)
          if (count == i)
            held[i] = there;
        count++;
      });
    }
    uint32_t total;
    uint32_t at = wave_prefix(count, lane, &total);
    if (total == 0)
      continue;
    uint32_t base = 0;
    if (lane == 0)
      base = __hip_atomic_fetch_add(gmem(f.counter) + 3, total, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    at += __shfl(base, 0);
    auto put = [&](uint32_t there) {
      if (at < f.link_capacity) {
        gmem(f.links)[2 * at] = here;
        gmem(f.links)[2 * at + 1] = there;
      }
      at++;
    };
    if (count <= kLinksHeld) {
#pragma unroll
      for (uint32_t i = 0; i < kLinksHeld; i++)
        if (i < count)
          put(held[i]);
    } else {
      walk(r, put);
    }
  }
}

} // namespace

namespace sfi {

// frames_device: the frames' descriptors in device memory, nframes long.
void launch_label(const FrameDev *frames_device, uint32_t nframes,
                  hipStream_t stream) {
  if (nframes == 0)
    return;
  const dim3 grid(256, nframes);
  count_runs<<<grid, 256, 0, stream>>>(frames_device);
  scan_runs<<<nframes, 1024, 0, stream>>>(frames_device);
  emit_runs<<<grid, 256, 0, stream>>>(frames_device);
  unite_runs<<<grid, 256, 0, stream>>>(frames_device);
  flatten<<<grid, 256, 0, stream>>>(frames_device);
  number_roots<<<nframes, 1024, 0, stream>>>(frames_device);
  count_component_runs<<<grid, 256, 0, stream>>>(frames_device);
  scan_components<<<nframes, 1024, 0, stream>>>(frames_device);
  place_runs<<<grid, 256, 0, stream>>>(frames_device);
  accumulate<<<grid, 256, 0, stream>>>(frames_device);
  link_frames<<<grid, 256, 0, stream>>>(frames_device);
  check(hipGetLastError(), "launching the labelling");
}

} // namespace sfi
