// Putting each frame's signal pixels into ascending index order on the
// device, for a batch of frames at once.
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
//
// Every kernel here takes an array of per-frame descriptors and works on
// frame blockIdx.y, so a batch costs the same few launches as one frame; the
// scans are segmented by frame, each frame's prefix starting at zero.
#include "sf_device.hh"
#include "sf_internal.hh"

namespace {

using sfi::FrameDev;
using sfi::kScanPerBlock;
constexpr int kScanThreads = 256;
constexpr int kPerThread = kScanPerBlock / kScanThreads;

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

// The words of frame blockIdx.y's bit array that are in use: all of them,
// or with a live count only those that hold it (none for a count above the
// capacity, which the caller refuses anyway).
__device__ __forceinline__ size_t live_words(const sfi::ScanSeg &s) {
  if (s.live == nullptr)
    return s.words;
  const uint32_t n = *s.live;
  const size_t w = n > s.capacity ? 0 : (static_cast<size_t>(n) + 31) / 32;
  return w < s.words ? w : s.words;
}

// Four words as one 16-byte access (a plain vector type: the HIP class
// cannot be copied out of the global address space).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// This thread's kPerThread consecutive words of the block's share of a bit
// array, popcounted: four 16-byte loads where the share is whole, words past
// the live ones reading as zero.
__device__ __forceinline__ void thread_counts(const sfi::ScanSeg &s, size_t w0,
                                              size_t words,
                                              uint32_t (&c)[kPerThread]) {
  const SF_GLOBAL uint32_t *bits = sfd::gmem(s.bits);
  if (w0 + kPerThread <= words) {
    const SF_GLOBAL u32x4 *v = reinterpret_cast<const SF_GLOBAL u32x4 *>(bits + w0);
#pragma unroll
    for (int i = 0; i < kPerThread / 4; i++) {
      const u32x4 x = v[i];
#pragma unroll
      for (int k = 0; k < 4; k++)
        c[4 * i + k] = static_cast<uint32_t>(__popc(x[k]));
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < kPerThread; i++)
    c[i] = w0 + i < words ? static_cast<uint32_t>(__popc(bits[w0 + i])) : 0u;
}

// Per-block totals of popcounts; block (x, frame), each thread kPerThread
// consecutive words.
__global__ __launch_bounds__(kScanThreads) void popcount_totals(
    const sfi::ScanSeg *__restrict__ segs) {
  const sfi::ScanSeg s = segs[blockIdx.y];
  const size_t base = static_cast<size_t>(blockIdx.x) * kScanPerBlock;
  if (base >= s.words)
    return;
  uint32_t c[kPerThread];
  thread_counts(s, base + static_cast<size_t>(threadIdx.x) * kPerThread,
                live_words(s), c);
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < kPerThread; i++)
    sum += c[i];
  uint32_t total;
  (void)block_inclusive(sum, &total);
  if (threadIdx.x == 0)
    s.totals[blockIdx.x] = total;
}

// Exclusive scan of one frame's block totals, in place; one block a frame.
__global__ __launch_bounds__(kScanThreads) void scan_totals(
    const sfi::ScanSeg *__restrict__ segs) {
  const sfi::ScanSeg s = segs[blockIdx.y];
  const uint32_t n =
      static_cast<uint32_t>((s.words + kScanPerBlock - 1) / kScanPerBlock);
  uint32_t carry = 0;
  for (uint32_t start = 0; start < n; start += kScanThreads) {
    const uint32_t i = start + threadIdx.x;
    const uint32_t v = i < n ? s.totals[i] : 0u;
    uint32_t total;
    const uint32_t inc = block_inclusive(v, &total);
    if (i < n)
      s.totals[i] = carry + inc - v;
    carry += total;
  }
}

// Exclusive prefix of popcounts per word, frame-local.  Each thread takes
// kPerThread consecutive words of the block's share, a block scan of the
// threads' sums gives each its start, and the thread writes its words'
// prefixes (only those of words with a set bit, for a sparse array).
__global__ __launch_bounds__(kScanThreads) void popcount_prefix(
    const sfi::ScanSeg *__restrict__ segs) {
  const sfi::ScanSeg s = segs[blockIdx.y];
  const size_t base = static_cast<size_t>(blockIdx.x) * kScanPerBlock;
  if (base >= s.words)
    return;
  const size_t words = live_words(s);
  const size_t w0 = base + static_cast<size_t>(threadIdx.x) * kPerThread;
  uint32_t c[kPerThread];
  thread_counts(s, w0, words, c);
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < kPerThread; i++)
    sum += c[i];
  uint32_t total;
  uint32_t run = s.totals[blockIdx.x] + block_inclusive(sum, &total) - sum;
  SF_GLOBAL uint32_t *prefix = sfd::gmem(s.prefix);
  if (!s.sparse && w0 + kPerThread <= words) {
    SF_GLOBAL u32x4 *v = reinterpret_cast<SF_GLOBAL u32x4 *>(prefix + w0);
#pragma unroll
    for (int i = 0; i < kPerThread / 4; i++) {
      u32x4 x;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        x[k] = run;
        run += c[4 * i + k];
      }
      v[i] = x;
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < kPerThread; i++) {
    if (w0 + i < words && (c[i] != 0 || !s.sparse))
      prefix[w0 + i] = run;
    run += c[i];
  }
}

__global__ __launch_bounds__(256) void scatter_by_rank(
    const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t n = f.counter[0];
  if (n > f.capacity) // the list is incomplete; the caller reports -2
    return;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const sf_signal_pixel rec = f.unordered[i];
    const uint32_t y = rec.index / f.width;
    const uint32_t x = rec.index - y * f.width;
    const size_t word = static_cast<size_t>(y) * f.wpr + x / 32;
    const uint32_t below = f.mask[word] & ((1u << (x & 31u)) - 1u);
    f.ordered[f.prefix[word] + static_cast<uint32_t>(__popc(below))] = rec;
  }
}

__global__ __launch_bounds__(256) void set_bits(const FrameDev *__restrict__ frames) {
  const FrameDev f = frames[blockIdx.y];
  const uint32_t n = f.counter[0];
  if (n > f.capacity)
    return;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t idx = f.ordered[i].index;
    const uint32_t y = idx / f.width, x = idx - y * f.width;
    atomicOr(&f.mask[static_cast<size_t>(y) * f.wpr + x / 32], 1u << (x & 31u));
  }
}

__global__ __launch_bounds__(256) void clear_frames(
    const FrameDev *__restrict__ frames, int masks, int first,
    uint32_t *__restrict__ queue) {
  if (queue && blockIdx.x == 0 && blockIdx.y == 0 &&
      threadIdx.x < static_cast<unsigned>(sfi::kQueueParts))
    queue[sfi::kQueueStride * threadIdx.x] = 0;
  const FrameDev f = frames[blockIdx.y];
  if (blockIdx.x == 0 && static_cast<int>(threadIdx.x) >= first &&
      threadIdx.x < sfi::kCounters)
    f.counter[threadIdx.x] = 0;
  if (!masks)
    return;
  const size_t words = static_cast<size_t>(f.height) * f.wpr;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
       i < words; i += static_cast<size_t>(gridDim.x) * blockDim.x)
    f.mask[i] = 0;
  const size_t blocks = (words + kScanPerBlock - 1) / kScanPerBlock;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
       i < blocks; i += static_cast<size_t>(gridDim.x) * blockDim.x)
    f.totals[i] = 0;
}

__global__ __launch_bounds__(256) void fill_words(uint32_t *p, size_t n,
                                                  uint32_t value) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
       i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = value;
}

// Copies `words` 16-byte words, then `tail` (< 16) bytes, from pinned host
// memory mapped into the device's address space.  Every load crosses PCIe,
// so the grid only needs enough loads in flight to cover its latency.
__global__ __launch_bounds__(256) void pull_words(const uint4 *__restrict__ src,
                                                  uint4 *__restrict__ dst,
                                                  size_t words, uint32_t tail) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
       i < words; i += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
  if (blockIdx.x == 0 && threadIdx.x < tail)
    reinterpret_cast<uint8_t *>(dst + words)[threadIdx.x] =
        reinterpret_cast<const uint8_t *>(src + words)[threadIdx.x];
}

} // namespace

namespace sfi {

size_t scan_tmp_words(size_t words) {
  return (words + kScanPerBlock - 1) / kScanPerBlock + 1;
}

void launch_popcount_scan(const ScanSeg *segs_device, uint32_t nframes,
                          size_t max_words, hipStream_t stream,
                          bool totals_ready) {
  const uint32_t blocks =
      static_cast<uint32_t>((max_words + kScanPerBlock - 1) / kScanPerBlock);
  if (blocks == 0 || nframes == 0)
    return;
  if (!totals_ready)
    popcount_totals<<<dim3(blocks, nframes), kScanThreads, 0, stream>>>(segs_device);
  scan_totals<<<dim3(1, nframes), kScanThreads, 0, stream>>>(segs_device);
  popcount_prefix<<<dim3(blocks, nframes), kScanThreads, 0, stream>>>(segs_device);
  check(hipGetLastError(), "launching a popcount scan");
}

void launch_clear(const FrameDev *frames_device, uint32_t nframes, bool masks,
                  int first, hipStream_t stream, uint32_t *queue) {
  if (nframes == 0)
    return;
  clear_frames<<<dim3(masks ? 256 : 1, nframes), 256, 0, stream>>>(
      frames_device, masks ? 1 : 0, first, queue);
  check(hipGetLastError(), "launching the frame clear");
}

void launch_scatter(const FrameDev *frames_device, uint32_t nframes,
                    hipStream_t stream) {
  if (nframes == 0)
    return;
  // Enough blocks to cover a dense frame in a few strides; the count itself
  // is only known on the device.
  scatter_by_rank<<<dim3(256, nframes), 256, 0, stream>>>(frames_device);
  check(hipGetLastError(), "launching the rank scatter");
}

void launch_set_bits(const FrameDev *frames_device, uint32_t nframes,
                     hipStream_t stream) {
  if (nframes == 0)
    return;
  set_bits<<<dim3(256, nframes), 256, 0, stream>>>(frames_device);
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

const void *mapped_host(const void *p) {
  hipPointerAttribute_t attr{};
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError(); // pageable memory can report an error here
    return nullptr;
  }
  return attr.type == hipMemoryTypeHost ? attr.devicePointer : nullptr;
}

void launch_pull(const void *mapped_src, void *dst, size_t bytes,
                 hipStream_t stream) {
  if (bytes == 0)
    return;
  // 256 blocks of 256 threads: 64K loads of 16 bytes in flight, which
  // measures at the link's rate (57 GB/s for a 72 MB frame); more blocks
  // only add contention.
  pull_words<<<256, 256, 0, stream>>>(static_cast<const uint4 *>(mapped_src),
                                      static_cast<uint4 *>(dst), bytes / 16,
                                      static_cast<uint32_t>(bytes % 16));
  check(hipGetLastError(), "launching a host pull");
}

} // namespace sfi
