// Putting a frame's signal pixels into ascending index order on the device.
//
// The reference sorts the device's unordered emit on the host
// (signal_order::by_index, src/spots/signal_order.cc:19-72).  Here the
// threshold also writes a strong-pixel bitmask, so a pixel's place in the
// ordered list is simply its rank: the number of set bits before it.  An
// exclusive scan of the per-word popcounts gives every word's rank base, and
// one scatter puts each unordered record at
//     base[word] + popcount(word & bits below it).
// O(n) in the frame's words plus the list, no comparisons, and the order is a
// property of the frame alone, not of the order the waves finished in.
#include "sf_device.hh"
#include "sf_internal.hh"

namespace {

// The scan: blocks of 256 threads, 16 words a thread, 4096 words a block.
constexpr int kScanThreads = 256;
constexpr int kPerThread = 16;
constexpr int kPerBlock = kScanThreads * kPerThread;

// Inclusive scan of one value across the block; returns this thread's
// inclusive value and the block total through *total.
__device__ __forceinline__ uint32_t block_inclusive(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wave_sums[kScanThreads / 64];
  const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u = __shfl_up(v, d);
    if (lane >= static_cast<uint32_t>(d))
      v += u;
  }
  if (lane == 63)
    wave_sums[wave] = v;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; w++) {
    const uint32_t x = wave_sums[w];
    before += static_cast<uint32_t>(w) < wave ? x : 0u;
    all += x;
  }
  __syncthreads();
  *total = all;
  return v + before;
}

// Words in use: all of them, or with a live bit count those that hold it
// (none for a count above capacity, which the caller refuses anyway).
__device__ __forceinline__ size_t live_words(size_t words, const uint32_t *live,
                                             uint32_t capacity) {
  if (live == nullptr)
    return words;
  const uint32_t n = *live;
  const size_t w = n > capacity ? 0 : (static_cast<size_t>(n) + 31) / 32;
  return w < words ? w : words;
}

// Per-block totals of popcounts.
__global__ __launch_bounds__(kScanThreads) void popcount_totals(
    const uint32_t *__restrict__ mask, size_t words, const uint32_t *live,
    uint32_t capacity, uint32_t *__restrict__ totals) {
  words = live_words(words, live, capacity);
  const size_t base = static_cast<size_t>(blockIdx.x) * kPerBlock +
                      static_cast<size_t>(threadIdx.x) * kPerThread;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < kPerThread; i++)
    if (base + i < words)
      c += static_cast<uint32_t>(__popc(mask[base + i]));
  uint32_t total;
  (void)block_inclusive(c, &total);
  if (threadIdx.x == 0)
    totals[blockIdx.x] = total;
}

// Exclusive scan of the block totals in place, by one block.
__global__ __launch_bounds__(kScanThreads) void scan_totals(uint32_t *totals,
                                                            uint32_t n) {
  uint32_t carry = 0;
  for (uint32_t start = 0; start < n; start += kScanThreads) {
    const uint32_t i = start + threadIdx.x;
    const uint32_t v = i < n ? totals[i] : 0u;
    uint32_t total;
    const uint32_t inc = block_inclusive(v, &total);
    if (i < n)
      totals[i] = carry + inc - v;
    carry += total;
  }
}

// Exclusive prefix of popcounts per word.
__global__ __launch_bounds__(kScanThreads) void popcount_prefix(
    const uint32_t *__restrict__ mask, size_t words, const uint32_t *live,
    uint32_t capacity, const uint32_t *__restrict__ totals,
    uint32_t *__restrict__ prefix) {
  words = live_words(words, live, capacity);
  const size_t base = static_cast<size_t>(blockIdx.x) * kPerBlock +
                      static_cast<size_t>(threadIdx.x) * kPerThread;
  uint32_t c[kPerThread];
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < kPerThread; i++) {
    c[i] = base + i < words ? static_cast<uint32_t>(__popc(mask[base + i])) : 0u;
    sum += c[i];
  }
  uint32_t total;
  uint32_t run = block_inclusive(sum, &total) - sum + totals[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kPerThread; i++) {
    if (base + i < words)
      prefix[base + i] = run;
    run += c[i];
  }
}

__global__ __launch_bounds__(256) void scatter_by_rank(
    const sf_signal_pixel *__restrict__ in, sf_signal_pixel *__restrict__ out,
    const uint32_t *__restrict__ counter, uint32_t capacity,
    const uint32_t *__restrict__ mask, const uint32_t *__restrict__ base,
    uint32_t width, uint32_t wpr) {
  const uint32_t n = *counter;
  if (n > capacity) // the list is incomplete; the caller reports -2
    return;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const sf_signal_pixel rec = in[i];
    const uint32_t y = rec.index / width;
    const uint32_t x = rec.index - y * width;
    const size_t word = static_cast<size_t>(y) * wpr + x / 32;
    const uint32_t below = mask[word] & ((1u << (x & 31u)) - 1u);
    out[base[word] + static_cast<uint32_t>(__popc(below))] = rec;
  }
}

__global__ __launch_bounds__(256) void set_bits(
    const sf_signal_pixel *__restrict__ list, const uint32_t *__restrict__ counter,
    uint32_t capacity, uint32_t width, uint32_t *__restrict__ mask, uint32_t wpr) {
  const uint32_t n = *counter;
  if (n > capacity)
    return;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t idx = list[i].index;
    const uint32_t y = idx / width, x = idx - y * width;
    atomicOr(&mask[static_cast<size_t>(y) * wpr + x / 32], 1u << (x & 31u));
  }
}

__global__ __launch_bounds__(256) void fill_words(uint32_t *p, size_t n,
                                                  uint32_t value) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
       i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = value;
}

} // namespace

namespace sfi {

size_t scan_tmp_bytes(size_t words) {
  return ((words + kPerBlock - 1) / kPerBlock + 1) * sizeof(uint32_t);
}

size_t order_scan_bytes(size_t words) { return scan_tmp_bytes(words); }

void launch_popcount_scan(const uint32_t *bits, size_t words,
                          const uint32_t *live, uint32_t capacity,
                          uint32_t *prefix, void *tmp, hipStream_t stream) {
  const uint32_t blocks =
      static_cast<uint32_t>((words + kPerBlock - 1) / kPerBlock);
  if (blocks == 0)
    return;
  uint32_t *totals = static_cast<uint32_t *>(tmp);
  popcount_totals<<<blocks, kScanThreads, 0, stream>>>(bits, words, live,
                                                       capacity, totals);
  scan_totals<<<1, kScanThreads, 0, stream>>>(totals, blocks);
  popcount_prefix<<<blocks, kScanThreads, 0, stream>>>(bits, words, live,
                                                       capacity, totals, prefix);
  check(hipGetLastError(), "launching a popcount scan");
}

void launch_order(const FrameArgs &a, const sf_signal_pixel *unordered,
                  sf_signal_pixel *ordered, uint32_t *prefix, void *scan_tmp,
                  size_t scan_tmp_bytes, hipStream_t stream) {
  const size_t words = static_cast<size_t>(a.height) * a.wpr;
  if (scan_tmp_bytes < order_scan_bytes(words))
    throw device_error("the rank scan's scratch is too small");
  launch_popcount_scan(a.mask, words, nullptr, 0, prefix, scan_tmp, stream);
  // Enough blocks to cover a dense frame in a few strides; the count itself
  // is only known on the device.
  scatter_by_rank<<<1024, 256, 0, stream>>>(unordered, ordered, a.counter,
                                            a.capacity, a.mask, prefix,
                                            a.width, a.wpr);
  check(hipGetLastError(), "launching the rank scatter");
}

void launch_set_bits(const sf_signal_pixel *list, const uint32_t *counter,
                     uint32_t capacity, uint32_t width, uint32_t *mask,
                     uint32_t wpr, hipStream_t stream) {
  set_bits<<<512, 256, 0, stream>>>(list, counter, capacity, width, mask, wpr);
  check(hipGetLastError(), "launching set_bits");
}

void launch_fill(uint32_t *p, size_t n, uint32_t value, hipStream_t stream) {
  if (n == 0)
    return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096)
    blocks = 4096;
  fill_words<<<static_cast<unsigned>(blocks), 256, 0, stream>>>(p, n, value);
  check(hipGetLastError(), "launching a fill");
}

} // namespace sfi
