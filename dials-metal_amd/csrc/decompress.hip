// Decompression of frames on the device: the chunks an EIGER's HDF5 filters
// write (src/spots/decompress.cc:1-136), decoded straight into frames in HBM,
// so that only the compressed bytes cross PCIe and no host thread decodes.
//
//   bslz4  bitshuffle + LZ4 (bshuf_decompress_lz4, decompress.cc:111-121):
//          independent blocks of `block` elements, each a big-endian length
//          and an LZ4 block of the bit-transposed elements, then the count % 8
//          leftover elements raw.  Where each block starts follows from the one
//          before, so the host walks the lengths (it holds the chunk anyway)
//          and every block goes to a wave of its own: the wave decodes the LZ4
//          block into LDS and transposes its bits into the frame.
//   lz4    the HDF5 LZ4 filter (decompress.cc:43-72): blocks decoded in order
//          by one wave a chunk, its matches read from a 64 KB LDS window.
//   ""     stored uncompressed: a copy.
//
// LZ4 decoding is wave-uniform: one sequence at a time, its token, lengths and
// offset read from a 256-byte window of the input held in registers (4 bytes a
// lane, read by readlane), its literals and match copied 64 bytes a step, one
// a lane.  Every read and write is checked against the block's bounds, so a
// malformed block ends as an error status, never as an access outside it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "sf_device.hh"
#include "sf_internal.hh"

namespace {

using sfi::check;

constexpr uint32_t kLanes = 64;
constexpr uint32_t kWindow = 256;      // input bytes a wave holds, 4 a lane
constexpr uint32_t kSmallBlock = 8192; // bytes of a bslz4 block four waves
                                       // of a workgroup decode in LDS each
constexpr uint32_t kLargeBlock = 65536; // a wave of its own, up to this
constexpr uint32_t kRing = 65536;      // the LZ4 window: a match reaches back
                                       // at most 65535 bytes
constexpr uint32_t kPad = 512;         // staging bytes past the last chunk, so
                                       // that the input window never reads
                                       // beyond the allocation
constexpr uint32_t kLeftover = 0xFFFFFFFFu;

// What a chunk's waves need, in device memory.
struct ChunkDev {
  uint64_t payload;       // device address of the bytes after the header
  uint64_t frame;         // device address of the frame
  uint32_t payload_bytes;
  uint32_t elem;          // bytes per element
  uint32_t block;         // bslz4: elements per block; lz4: bytes per block
  uint32_t count;         // bslz4: elements in the frame; lz4: frame bytes
};

// One bslz4 block (or the raw leftover) for one wave.
struct BlockItem {
  uint32_t chunk;
  uint32_t index; // block number in the chunk, or kLeftover
  uint32_t at;    // payload offset of the LZ4 bytes (after the length)
  uint32_t bytes; // the LZ4 block's length, or the leftover's
};

// Per-chunk status words, written by the first wave that finds a fault.
enum : uint32_t {
  kOk = 0,
  kMalformed = 1,     // an LZ4 block of a bslz4 chunk is malformed
  kWrongSize = 2,     // ... or decodes to other than its block's bytes
  kLz4Failed = 3,     // an LZ4 block of an lz4 chunk is malformed
  kLz4Truncated = 4,  // decompress.cc:50-51
  kLz4BlockRange = 5, // decompress.cc:56-58
};

__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t lane) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(lane)));
}

__device__ __forceinline__ uint32_t bpermute(uint32_t v, uint32_t lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(
      static_cast<int>(lane << 2), static_cast<int>(v)));
}

// A stream of input bytes seen through a 256-byte window: byte p of the stream
// is at origin + shift + p, origin 4-byte aligned; the window holds origin +
// [base, base + 256), 4 bytes a lane.
struct Window {
  const uint8_t *origin;
  uint32_t shift, base, word, lane;

  __device__ Window(const uint8_t *p, uint32_t lane_) : lane(lane_) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    origin = reinterpret_cast<const uint8_t *>(a & ~uintptr_t(3));
    shift = static_cast<uint32_t>(a & 3);
    base = 0;
    word = load(0);
  }
  __device__ uint32_t load(uint32_t b) const {
    return *sfd::gmem(reinterpret_cast<const uint32_t *>(origin + b + 4 * lane));
  }
  // Brings stream bytes [p, p + need) into the window (need <= 253).
  __device__ void hold(uint32_t p, uint32_t need) {
    const uint32_t q = p + shift;
    if (q < base || q + need > base + kWindow) {
      base = q & ~3u;
      word = load(base);
    }
  }
  __device__ uint32_t byte(uint32_t p) {
    hold(p, 1);
    const uint32_t d = p + shift - base;
    return (readlane(word, d >> 2) >> (8 * (d & 3))) & 0xFFu;
  }
  // Bytes p .. p + 3, byte p lowest (those past the stream are the staging's).
  __device__ uint32_t peek4(uint32_t p) {
    hold(p, 4);
    const uint32_t d = p + shift - base;
    const uint32_t lo = readlane(word, d >> 2);
    if ((d & 3) == 0)
      return lo;
    // hold(p, 4) leaves d <= 252, so a straddled word is not the last
    const uint32_t hi = readlane(word, (d >> 2) + 1);
    return static_cast<uint32_t>(
        ((static_cast<uint64_t>(hi) << 32) | lo) >> (8 * (d & 3)));
  }
  // Byte p + lane, for the lanes that hold(p, n) covered.
  __device__ uint32_t lane_byte(uint32_t p) const {
    const uint32_t d = p + shift - base + lane;
    return (bpermute(word, (d >> 2) & (kLanes - 1)) >> (8 * (d & 3))) & 0xFFu;
  }
};

// Output held whole in LDS (a bslz4 block): `cap` bytes, then kLanes bytes
// that lanes past the end of a copy write instead, so no write needs a branch.
struct BlockSink {
  uint8_t *buf;
  uint32_t cap;
  __device__ void put(uint32_t pos, uint32_t b, bool ok, uint32_t lane) {
    buf[ok ? pos : cap + lane] = static_cast<uint8_t>(b);
  }
  __device__ uint32_t get(uint32_t pos) const { return buf[pos]; }
};

// Output written to the frame, its last kRing bytes kept in LDS for the
// matches to read (an lz4 block of any length).
struct RingSink {
  uint8_t *ring;
  uint8_t *out;
  __device__ void put(uint32_t pos, uint32_t b, bool ok, uint32_t) {
    if (ok) {
      ring[pos & (kRing - 1)] = static_cast<uint8_t>(b);
      sfd::gmem(out)[pos] = static_cast<uint8_t>(b);
    }
  }
  __device__ uint32_t get(uint32_t pos) const { return ring[pos & (kRing - 1)]; }
};

// A match of `len` bytes from `off` back, written at op.  Below 64 the match
// repeats its first `off` bytes, which the lanes below `off` hold in a
// register, so nothing the match writes is read back; from 64 up, byte i comes
// from op - off + i, which an earlier step (or the output before the match)
// wrote, and within a step every lane reads before any lane writes.  In a ring
// the slot a step writes can only be one that this step has already read.
constexpr uint32_t kBatch = 4; // steps of 64 bytes whose reads go out together

template <class Sink>
__device__ __forceinline__ void copy_match(Sink &s, uint32_t op, uint32_t off,
                                           uint32_t len, uint32_t lane) {
  const uint32_t steps = (len + kLanes - 1) / kLanes;
  if (off < kLanes) {
    const uint32_t lm = lane % off;
    const uint32_t pattern = s.get(op - off + lm);
    const uint32_t step = kLanes % off;
    uint32_t r = 0; // (64 k) % off
    for (uint32_t k = 0; k < steps; k++) {
      uint32_t m = lm + r;
      m = m >= off ? m - off : m;
      const uint32_t i = k * kLanes + lane;
      s.put(op + i, bpermute(pattern, m), i < len, lane);
      r += step;
      r = r >= off ? r - off : r;
    }
  } else {
    // Steps whose sources all precede this batch's first output byte: the
    // first off / 64 of them.  (A lane past the end reads below op + len.)
    const uint32_t m = min(kBatch, off / kLanes);
    for (uint32_t k = 0; k < steps; k += m) {
      const uint32_t mm = min(m, steps - k);
      uint32_t b[kBatch];
#pragma unroll
      for (uint32_t j = 0; j < kBatch; j++)
        if (j < mm)
          b[j] = s.get(op - off + (k + j) * kLanes + lane);
#pragma unroll
      for (uint32_t j = 0; j < kBatch; j++)
        if (j < mm) {
          const uint32_t i = (k + j) * kLanes + lane;
          s.put(op + i, b[j], i < len, lane);
        }
    }
  }
}

// One LZ4 block of n bytes into at most cap bytes: the bytes written, or -1
// for a malformed block.  The checks are the block format's, held as
// LZ4_decompress_safe holds them: the last sequence is literals only and ends
// the input; any other leaves at least 8 input bytes after its literals; a
// match starts at least 12 and ends at least 5 bytes before cap; its offset is
// 1 .. the bytes written so far.
template <class Sink>
__device__ __forceinline__ int64_t lz4_block(Window &in, uint32_t n, Sink &s, uint32_t cap,
                             uint32_t lane) {
  if (n == 0)
    return -1;
  uint32_t ip = 0, op = 0;
  for (;;) {
    if (ip >= n)
      return -1;
    const uint32_t token = in.byte(ip++);
    uint32_t lit = token >> 4;
    if (lit == 15) {
      uint32_t b;
      do {
        if (ip >= n)
          return -1;
        b = in.byte(ip++);
        lit += b;
        if (lit > n)
          return -1;
      } while (b == 255);
    }
    if (lit > n - ip || lit > cap - op)
      return -1;
    // Up to three steps a window: 192 bytes, all of them inside one hold.
    for (uint32_t done = 0; done < lit; done += 3 * kLanes) {
      const uint32_t cnt = min(3 * kLanes, lit - done);
      in.hold(ip + done, cnt);
      for (uint32_t j = 0; j * kLanes < cnt; j++) {
        const uint32_t i = j * kLanes + lane;
        s.put(op + done + i, in.lane_byte(ip + done + j * kLanes), i < cnt, lane);
      }
    }
    ip += lit;
    op += lit;
    if (ip == n)
      return op;
    if (cap < 12 || op > cap - 12 || n - ip < 8)
      return -1;
    // The offset and the first two length bytes in one read (n - ip >= 8).
    const uint32_t o4 = in.peek4(ip);
    const uint32_t offset = o4 & 0xFFFFu;
    ip += 2;
    uint32_t len = token & 15;
    if (len == 15) {
      uint32_t b, k = 2;
      do {
        if (ip >= n)
          return -1;
        b = k < 4 ? (o4 >> (8 * k)) & 0xFFu : in.byte(ip);
        ip++;
        k++;
        len += b;
        if (len > cap)
          return -1;
      } while (b == 255);
    }
    len += 4;
    if (offset == 0 || offset > op)
      return -1;
    if (len > cap - 5 - op)
      return -1;
    copy_match(s, op, offset, len, lane);
    op += len;
  }
}

// One round of the transpose: rows k and k + J (bit J of k clear) swap the
// high J bits of each 2J-bit block of row k with the low J bits of row k + J.
template <int J>
__device__ __forceinline__ void swap_blocks(uint32_t (&a)[32], uint32_t m) {
#pragma unroll
  for (int k = 0; k < 32; k++) {
    if ((k & J) == 0) {
      const uint32_t t = ((a[k] >> J) ^ a[k + J]) & m;
      a[k] ^= t << J;
      a[k + J] ^= t;
    }
  }
}

// After it, bit r of a[t] is bit t of the old a[r].
__device__ __forceinline__ void transpose32(uint32_t (&a)[32]) {
  swap_blocks<16>(a, 0x0000FFFFu);
  swap_blocks<8>(a, 0x00FF00FFu);
  swap_blocks<4>(a, 0x0F0F0F0Fu);
  swap_blocks<2>(a, 0x33333333u);
  swap_blocks<1>(a, 0x55555555u);
}

// The inverse bit transpose of one block of `count` elements of E bytes
// (count % 8 == 0): bit-row 8j + k of `rows` (count / 8 bytes each) holds bit
// k of byte j of every element, element i at bit i % 8 of byte i / 8.  Each
// lane takes 32 elements at a time: one 32-bit word from each of the 8E rows,
// transposed, is those elements.  `out` is 8-byte aligned.
template <int E>
__device__ __forceinline__ void unshuffle(const uint8_t *rows, uint32_t count, uint8_t *out,
                          uint32_t lane) {
  const uint32_t row = count / 8;
  const uint32_t groups = (count + 31) / 32;
  const bool words = row % 4 == 0;
  for (uint32_t g = lane; g < groups; g += kLanes) {
    const uint32_t nb = min(32u, count - 32 * g); // 8, 16, 24 or 32
    uint32_t a[32];
    if (words && nb == 32) {
#pragma unroll
      for (int r = 0; r < 32; r++)
        a[r] = r < 8 * E
                   ? *reinterpret_cast<const uint32_t *>(rows + r * row + 4 * g)
                   : 0u;
    } else { // a short last group, or rows of other than whole words
      const uint32_t have = nb / 8;
#pragma unroll
      for (int r = 0; r < 32; r++) {
        uint32_t v = 0;
        if (r < 8 * E) {
#pragma unroll
          for (int b = 0; b < 4; b++)
            if (static_cast<uint32_t>(b) < have)
              v |= static_cast<uint32_t>(rows[r * row + 4 * g + b]) << (8 * b);
        }
        a[r] = v;
      }
    }
    transpose32(a);
    SF_GLOBAL uint64_t *o = sfd::gmem(reinterpret_cast<uint64_t *>(out + 32 * g * E));
    if constexpr (E == 4) {
#pragma unroll
      for (int t = 0; t < 16; t++)
        if (2u * t < nb)
          o[t] = a[2 * t] | (static_cast<uint64_t>(a[2 * t + 1]) << 32);
    } else if constexpr (E == 2) {
#pragma unroll
      for (int t = 0; t < 8; t++)
        if (4u * t < nb)
          o[t] = (a[4 * t] | (a[4 * t + 1] << 16)) |
                 (static_cast<uint64_t>(a[4 * t + 2] | (a[4 * t + 3] << 16)) << 32);
    } else {
#pragma unroll
      for (int t = 0; t < 4; t++)
        if (8u * t < nb) {
          uint64_t v = 0;
#pragma unroll
          for (int b = 0; b < 8; b++)
            v |= static_cast<uint64_t>(a[8 * t + b] & 0xFFu) << (8 * b);
          o[t] = v;
        }
    }
  }
}

__device__ __forceinline__ void fail(uint32_t *status, uint32_t chunk,
                                     uint32_t code, uint32_t lane) {
  if (lane == 0)
    atomicMax(status + chunk, code);
}

// bslz4 blocks of E-byte elements, one a wave, Cap bytes of LDS each.
template <uint32_t Cap, int Waves, int E>
__global__ __launch_bounds__(Waves * kLanes) void decode_bslz4(
    const ChunkDev *__restrict__ chunks, const BlockItem *__restrict__ items,
    uint32_t n, uint32_t *status) {
  __shared__ __align__(16) uint8_t lds[Waves][Cap + kLanes];
  const uint32_t lane = threadIdx.x % kLanes;
  // The wave's number, proven uniform, so that its item, its chunk and every
  // position the parse derives from them live in scalar registers.
  const uint32_t wave = static_cast<uint32_t>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kLanes)));
  const uint32_t k = blockIdx.x * Waves + wave;
  if (k >= n)
    return;
  const BlockItem it = items[k];
  const ChunkDev c = chunks[it.chunk];
  const uint8_t *payload = reinterpret_cast<const uint8_t *>(c.payload);
  uint8_t *frame = reinterpret_cast<uint8_t *>(c.frame);
  if (it.index == kLeftover) { // the count % 8 elements after the blocks
    const size_t at = static_cast<size_t>(c.count - c.count % 8) * c.elem;
    if (lane < it.bytes)
      sfd::gmem(frame)[at + lane] = sfd::gmem(payload)[it.at + lane];
    return;
  }
  const uint32_t first = it.index * c.block;
  const uint32_t count = min(c.block, c.count - first) / 8 * 8;
  const uint32_t cap = count * E; // at most Cap (the host sorts)
  Window in(payload + it.at, lane);
  BlockSink s{lds[wave], cap};
  const int64_t got = lz4_block(in, it.bytes, s, cap, lane);
  if (got != static_cast<int64_t>(cap)) {
    fail(status, it.chunk, got < 0 ? kMalformed : kWrongSize, lane);
    return;
  }
  unshuffle<E>(lds[wave], count, frame + static_cast<size_t>(first) * E, lane);
}

// lz4 chunks, one a wave: decompress.cc:43-72's loop.
__global__ __launch_bounds__(kLanes) void decode_lz4(
    const ChunkDev *__restrict__ chunks, const uint32_t *__restrict__ ids,
    uint32_t *status) {
  __shared__ __align__(16) uint8_t ring[kRing];
  const uint32_t lane = threadIdx.x;
  const uint32_t id = ids[blockIdx.x];
  const ChunkDev c = chunks[id];
  const uint8_t *payload = reinterpret_cast<const uint8_t *>(c.payload);
  uint8_t *frame = reinterpret_cast<uint8_t *>(c.frame);
  const uint32_t plen = c.payload_bytes, expected = c.count, bs = c.block;
  Window head(payload, lane);
  uint32_t read = 0, written = 0;
  while (written < expected) {
    if (plen - read < 4) {
      fail(status, id, kLz4Truncated, lane);
      return;
    }
    const uint32_t len = (head.byte(read) << 24) | (head.byte(read + 1) << 16) |
                         (head.byte(read + 2) << 8) | head.byte(read + 3);
    read += 4;
    if (len == 0 || len > plen - read) {
      fail(status, id, kLz4BlockRange, lane);
      return;
    }
    Window in(payload + read, lane);
    RingSink s{ring, frame + written};
    const int64_t got = lz4_block(in, len, s, min(expected - written, bs), lane);
    if (got < 0) {
      fail(status, id, kLz4Failed, lane);
      return;
    }
    read += len;
    written += static_cast<uint32_t>(got);
  }
}

__global__ void clear_status(uint32_t *status, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    status[i] = 0;
}

uint32_t be32(const uint8_t *p) {
  return (static_cast<uint32_t>(p[0]) << 24) | (static_cast<uint32_t>(p[1]) << 16) |
         (static_cast<uint32_t>(p[2]) << 8) | static_cast<uint32_t>(p[3]);
}

size_t align16(size_t n) { return (n + 15) & ~size_t(15); }

// A chunk's data does not describe its frame: the reference's exception, with
// its message.
struct chunk_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

std::string bitshuffle_code(long code) {
  return "bitshuffle decompression failed with code " + std::to_string(code);
}

} // namespace

struct sf_decoder {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // Uploads go on a stream of their own, so a load crosses PCIe while the
  // work before it on `stream` runs: the upload waits for the last run to be
  // done with the device staging (`consumed`), and the next run waits for
  // the upload (`uploaded`).
  hipStream_t upload = nullptr;
  hipEvent_t consumed = nullptr;
  bool ran_once = false;
  sfi::DeviceBuffer dev;
  uint8_t *host = nullptr; // pinned staging, `host_bytes` long
  size_t host_bytes = 0;
  hipEvent_t uploaded = nullptr, t0 = nullptr, t1 = nullptr;
  bool timing = false, timed = false, loaded = false, ran = false;

  // What the last load holds: bslz4 items in six groups, small or large
  // blocks by 1-, 2- or 4-byte elements (group = 3 * large + log2 elem).
  uint32_t n = 0, n_lz4 = 0;
  uint32_t n_items[6] = {};
  size_t off_items[6] = {};
  size_t off_lz4 = 0, off_status = 0;
  struct Raw {
    size_t at;  // staging offset
    void *frame;
    size_t bytes;
  };
  std::vector<Raw> raws;
  std::vector<int> kinds; // 0 raw, 1 bslz4, 2 lz4
  std::vector<uint32_t> status;

  ~sf_decoder() {
    if (upload)
      (void)hipStreamSynchronize(upload);
    if (stream)
      (void)hipStreamSynchronize(stream);
    if (host)
      (void)hipHostFree(host);
    for (hipEvent_t e : {uploaded, consumed, t0, t1})
      if (e)
        (void)hipEventDestroy(e);
    if (upload)
      (void)hipStreamDestroy(upload);
    if (own_stream && stream)
      (void)hipStreamDestroy(stream);
  }

  void load(const sf_chunk *chunks, uint32_t count);
  void run();
  int sync(int *out);
  // The last run's status once `after` has fired, read on `read_stream`;
  // each chunk's message (empty when it decoded) into `messages`.
  int status_after(hipStream_t read_stream, hipEvent_t after, int *out,
                   std::vector<std::string> *messages);

private:
  int report(int *out, std::vector<std::string> *messages);
};

namespace {

enum Kind { kRawChunk = 0, kBslz4 = 1, kLz4 = 2 };

// A chunk checked against its frame (decompress.cc:76-133) and its blocks
// walked: where its payload goes in the staging, and its waves' work.
struct Planned {
  Kind kind;
  const uint8_t *payload;
  size_t payload_bytes;
  ChunkDev dev;
  std::vector<BlockItem> small, large;
};

Planned plan_chunk(const sf_chunk &c, uint32_t index) {
  Planned p{};
  if (c.frame == nullptr || (reinterpret_cast<uintptr_t>(c.frame) & 7) != 0)
    throw chunk_error("a chunk's device frame is null or not 8-byte aligned");
  if (c.data == nullptr && c.size != 0)
    throw chunk_error("a chunk's data pointer is null");
  const size_t expected = sf_frame_bytes(c.height, c.width, c.bit_depth);
  const size_t element = c.bit_depth / 8;
  if (element == 0)
    throw chunk_error("output buffer too small for the frame");
  if (expected > 0xFFFFFFF0u || c.size > 0xFFFFFFF0u)
    throw chunk_error("a frame or chunk of 4 GiB or more");
  const std::string algorithm = c.algorithm ? c.algorithm : "";
  if (algorithm.empty()) {
    if (c.size != expected)
      throw chunk_error("uncompressed frame is " + std::to_string(c.size) +
                        " bytes, expected " + std::to_string(expected));
    p.kind = kRawChunk;
    p.payload = c.data;
    p.payload_bytes = c.size;
    return p;
  }
  if (c.size < 12)
    throw chunk_error("too short for a filter header");
  uint64_t size = 0;
  for (int i = 0; i < 8; i++)
    size = (size << 8) | c.data[i];
  const uint32_t header_block = be32(c.data + 8);
  if (size != expected)
    throw chunk_error("filter header says " + std::to_string(size) +
                      " bytes uncompressed, expected " + std::to_string(expected));
  p.payload = c.data + 12;
  p.payload_bytes = c.size - 12;
  p.dev.frame = reinterpret_cast<uint64_t>(c.frame);
  p.dev.payload_bytes = static_cast<uint32_t>(p.payload_bytes);
  p.dev.elem = static_cast<uint32_t>(element);
  if (algorithm == "lz4") {
    p.kind = kLz4;
    p.dev.block = header_block > 0 ? header_block : static_cast<uint32_t>(expected);
    p.dev.count = static_cast<uint32_t>(expected);
    return p;
  }
  if (algorithm != "bslz4")
    throw chunk_error("unknown compression algorithm " + algorithm);
  if (element != 1 && element != 2 && element != 4)
    throw chunk_error("bit depth " + std::to_string(c.bit_depth) +
                      " is beyond the device decoder (8, 16 or 32)");
  p.kind = kBslz4;
  const uint32_t n = static_cast<uint32_t>(expected / element);
  uint32_t block = header_block / static_cast<uint32_t>(element);
  if (block == 0) // bitshuffle's default: 8192 bytes, a multiple of 8, >= 128
    block = std::max<uint32_t>((8192 / static_cast<uint32_t>(element)) / 8 * 8, 128);
  if (block % 8 != 0)
    throw chunk_error(bitshuffle_code(-81));
  const uint64_t block_bytes = static_cast<uint64_t>(block) * element;
  p.dev.block = block;
  p.dev.count = n;
  const uint8_t *in = p.payload;
  const size_t plen = p.payload_bytes;
  size_t ip = 0;
  const uint32_t full = n / block;
  const uint32_t last = (n % block) / 8 * 8;
  const uint32_t blocks = full + (last ? 1 : 0);
  const uint64_t biggest = static_cast<uint64_t>(full ? block : last) * element;
  if (blocks && biggest > kLargeBlock)
    throw chunk_error("a bitshuffle block of " + std::to_string(block_bytes) +
                      " bytes is beyond the device decoder's " +
                      std::to_string(kLargeBlock));
  for (uint32_t b = 0; b < blocks; b++) {
    // bshuf_decompress_lz4 trusts the lengths; a length that runs off the
    // chunk is an error here (code -92) rather than a read past it.
    if (plen - ip < 4)
      throw chunk_error(bitshuffle_code(-92));
    const uint32_t len = be32(in + ip);
    if (len > plen - ip - 4)
      throw chunk_error(bitshuffle_code(-92));
    const uint32_t elems = b < full ? block : last;
    const BlockItem item{index, b, static_cast<uint32_t>(ip + 4), len};
    if (static_cast<uint64_t>(elems) * element <= kSmallBlock)
      p.small.push_back(item);
    else
      p.large.push_back(item);
    ip += 4 + static_cast<size_t>(len);
  }
  const uint32_t left = (n % 8) * static_cast<uint32_t>(element);
  if (left > plen - ip)
    throw chunk_error(bitshuffle_code(-92));
  if (left)
    p.small.push_back(BlockItem{index, kLeftover, static_cast<uint32_t>(ip), left});
  return p;
}

} // namespace

void sf_decoder::load(const sf_chunk *chunks, uint32_t count) {
  std::vector<Planned> plans;
  plans.reserve(count);
  for (uint32_t i = 0; i < count; i++)
    plans.push_back(plan_chunk(chunks[i], i));

  // Staging: [chunks][bslz4 items, six groups][lz4 ids][status]
  // [payloads, 16-aligned each][pad]
  size_t groups[6] = {}, nz = 0, payload = 0;
  auto group_of = [](const Planned &p, bool large) {
    return 3 * (large ? 1 : 0) + (p.dev.elem == 4 ? 2 : p.dev.elem == 2 ? 1 : 0);
  };
  for (const Planned &p : plans) {
    if (p.kind == kBslz4) {
      groups[group_of(p, false)] += p.small.size();
      groups[group_of(p, true)] += p.large.size();
    }
    nz += p.kind == kLz4;
    payload += align16(p.payload_bytes);
  }
  const size_t off_chunks = 0;
  size_t o_items[6];
  size_t end = align16(off_chunks + count * sizeof(ChunkDev));
  for (int g = 0; g < 6; g++) {
    o_items[g] = end;
    end = align16(end + groups[g] * sizeof(BlockItem));
  }
  const size_t o_lz4 = end;
  const size_t o_status = align16(o_lz4 + nz * sizeof(uint32_t));
  const size_t o_payload = align16(o_status + count * sizeof(uint32_t));
  const size_t total = o_payload + payload + kPad;

  // The staging may still be on its way up from the last load.
  if (uploaded) {
    check(hipEventSynchronize(uploaded), "waiting for the last upload");
  } else {
    check(hipEventCreateWithFlags(&uploaded, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&consumed, hipEventDisableTiming), "hipEventCreate");
    check(hipStreamCreateWithFlags(&upload, hipStreamNonBlocking), "hipStreamCreate");
  }
  dev.ensure(total, stream);
  if (total > host_bytes) {
    if (host)
      check(hipHostFree(host), "hipHostFree");
    host = nullptr;
    host_bytes = 0;
    check(hipHostMalloc(reinterpret_cast<void **>(&host), total, hipHostMallocDefault),
          "hipHostMalloc");
    host_bytes = total;
  }
  uint8_t *base = static_cast<uint8_t *>(dev.ptr);
  ChunkDev *cd = reinterpret_cast<ChunkDev *>(host + off_chunks);
  uint32_t *zi = reinterpret_cast<uint32_t *>(host + o_lz4);
  uint32_t filled[6] = {};
  std::memset(host + o_status, 0, count * sizeof(uint32_t));
  raws.clear();
  kinds.assign(count, 0);
  size_t at = o_payload;
  uint32_t kz = 0;
  std::vector<sfi::CopyPiece> copies;
  copies.reserve(count);
  for (uint32_t i = 0; i < count; i++) {
    Planned &p = plans[i];
    kinds[i] = p.kind;
    if (p.payload_bytes)
      copies.push_back(sfi::CopyPiece{host + at, p.payload, p.payload_bytes});
    p.dev.payload = reinterpret_cast<uint64_t>(base + at);
    cd[i] = p.dev;
    if (p.kind == kRawChunk)
      raws.push_back(Raw{at, chunks[i].frame, p.payload_bytes});
    for (int large = 0; large < 2; large++) {
      const int g = group_of(p, large != 0);
      BlockItem *dst = reinterpret_cast<BlockItem *>(host + o_items[g]);
      for (const BlockItem &b : large ? p.large : p.small)
        dst[filled[g]++] = b;
    }
    if (p.kind == kLz4)
      zi[kz++] = i;
    at += align16(p.payload_bytes);
  }
  std::memset(host + at, 0, total - at);
  sfi::host_copy(copies.data(), copies.size());
  if (ran_once) // the last run's kernels are done reading the device staging
    check(hipStreamWaitEvent(upload, consumed, 0), "hipStreamWaitEvent");
  check(hipMemcpyAsync(dev.ptr, host, total, hipMemcpyHostToDevice, upload),
        "uploading the chunks");
  check(hipEventRecord(uploaded, upload), "hipEventRecord");
  n = count;
  for (int g = 0; g < 6; g++) {
    n_items[g] = filled[g];
    off_items[g] = o_items[g];
  }
  n_lz4 = kz;
  off_lz4 = o_lz4;
  off_status = o_status;
  loaded = true;
  ran = false;
}

void sf_decoder::run() {
  uint8_t *base = static_cast<uint8_t *>(dev.ptr);
  const ChunkDev *cd = reinterpret_cast<const ChunkDev *>(base);
  uint32_t *st = reinterpret_cast<uint32_t *>(base + off_status);
  check(hipStreamWaitEvent(stream, uploaded, 0), "hipStreamWaitEvent");
  if (timing) {
    if (!t0) {
      check(hipEventCreate(&t0), "hipEventCreate");
      check(hipEventCreate(&t1), "hipEventCreate");
    }
    check(hipEventRecord(t0, stream), "hipEventRecord");
  }
  if (n)
    clear_status<<<1, 256, 0, stream>>>(st, n);
  auto items = [&](int g) {
    return reinterpret_cast<const BlockItem *>(base + off_items[g]);
  };
  const uint32_t *ni = n_items;
  if (ni[0])
    decode_bslz4<kSmallBlock, 4, 1><<<(ni[0] + 3) / 4, 4 * kLanes, 0, stream>>>(cd, items(0), ni[0], st);
  if (ni[1])
    decode_bslz4<kSmallBlock, 4, 2><<<(ni[1] + 3) / 4, 4 * kLanes, 0, stream>>>(cd, items(1), ni[1], st);
  if (ni[2])
    decode_bslz4<kSmallBlock, 4, 4><<<(ni[2] + 3) / 4, 4 * kLanes, 0, stream>>>(cd, items(2), ni[2], st);
  if (ni[3])
    decode_bslz4<kLargeBlock, 1, 1><<<ni[3], kLanes, 0, stream>>>(cd, items(3), ni[3], st);
  if (ni[4])
    decode_bslz4<kLargeBlock, 1, 2><<<ni[4], kLanes, 0, stream>>>(cd, items(4), ni[4], st);
  if (ni[5])
    decode_bslz4<kLargeBlock, 1, 4><<<ni[5], kLanes, 0, stream>>>(cd, items(5), ni[5], st);
  if (n_lz4)
    decode_lz4<<<n_lz4, kLanes, 0, stream>>>(
        cd, reinterpret_cast<const uint32_t *>(base + off_lz4), st);
  check(hipGetLastError(), "launching the decoders");
  for (const Raw &r : raws)
    if (r.bytes)
      check(hipMemcpyAsync(r.frame, base + r.at, r.bytes, hipMemcpyDeviceToDevice,
                           stream),
            "copying an uncompressed chunk");
  if (timing)
    check(hipEventRecord(t1, stream), "hipEventRecord");
  check(hipEventRecord(consumed, stream), "hipEventRecord");
  ran_once = true;
  timed = timing;
  ran = true;
}

int sf_decoder::sync(int *out) {
  status.assign(n, 0);
  if (n)
    check(hipMemcpyAsync(status.data(), static_cast<uint8_t *>(dev.ptr) + off_status,
                         n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
          "reading the decoders' status");
  check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  return report(out, nullptr);
}

int sf_decoder::status_after(hipStream_t read_stream, hipEvent_t after, int *out,
                             std::vector<std::string> *messages) {
  status.assign(n, 0);
  if (n) {
    check(hipStreamWaitEvent(read_stream, after, 0), "hipStreamWaitEvent");
    check(hipMemcpyAsync(status.data(), static_cast<uint8_t *>(dev.ptr) + off_status,
                         n * sizeof(uint32_t), hipMemcpyDeviceToHost, read_stream),
          "reading the decoders' status");
    check(hipStreamSynchronize(read_stream), "hipStreamSynchronize");
  }
  return report(out, messages);
}

int sf_decoder::report(int *out, std::vector<std::string> *messages) {
  if (messages)
    messages->assign(n, std::string());
  std::string first;
  for (uint32_t i = 0; i < n; i++) {
    std::string msg;
    switch (status[i]) {
    case kOk:
      break;
    case kMalformed: // LZ4_decompress_safe's failure, as bitshuffle returns it
      msg = bitshuffle_code(-1001);
      break;
    case kWrongSize:
      msg = bitshuffle_code(-91);
      break;
    case kLz4Failed:
      msg = "lz4 decompression failed with code -1";
      break;
    case kLz4Truncated:
      msg = "lz4 data truncated";
      break;
    case kLz4BlockRange:
      msg = "lz4 block size out of range";
      break;
    default:
      msg = "decoder status " + std::to_string(status[i]);
    }
    if (out)
      out[i] = msg.empty() ? 0 : -1;
    if (messages)
      (*messages)[i] = msg;
    if (!msg.empty() && first.empty())
      first = msg;
  }
  if (!first.empty()) {
    sfi::set_last_error(first);
    return -1;
  }
  return 0;
}

namespace {

template <typename F> int dguard(F &&body) {
  try {
    return body();
  } catch (const chunk_error &e) {
    sfi::set_last_error(e.what());
    return -1;
  } catch (const sfi::device_error &e) {
    sfi::set_last_error(e.what());
    return -3;
  } catch (const std::bad_alloc &) {
    sfi::set_last_error("out of host memory");
    return -4;
  } catch (const std::exception &e) {
    sfi::set_last_error(e.what());
    return -3;
  }
}

std::unique_ptr<sf_decoder> &thread_decoder() {
  static thread_local std::unique_ptr<sf_decoder> d;
  return d;
}

} // namespace

extern "C" {

size_t sf_frame_bytes(size_t height, size_t width, unsigned bit_depth) {
  return height * width * bit_depth / 8;
}

sf_decoder *sf_decoder_create(void *stream) {
  auto d = std::make_unique<sf_decoder>();
  if (stream) {
    d->stream = static_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
      sfi::set_last_error("hipStreamCreate failed");
      return nullptr;
    }
    d->own_stream = true;
  }
  return d.release();
}

void sf_decoder_destroy(sf_decoder *d) { delete d; }

void *sf_decoder_stream(sf_decoder *d) { return d ? d->stream : nullptr; }

int sf_decoder_load(sf_decoder *d, const sf_chunk *chunks, size_t n) {
  if (d == nullptr || (n && chunks == nullptr) || n > 0xFFFFFFu) {
    sfi::set_last_error("bad argument: decoder, chunks or count");
    return -1;
  }
  return dguard([&]() -> int {
    d->load(chunks, static_cast<uint32_t>(n));
    return 0;
  });
}

int sf_decoder_run(sf_decoder *d) {
  if (d == nullptr || !d->loaded) {
    sfi::set_last_error("bad argument: no chunks loaded");
    return -1;
  }
  return dguard([&]() -> int {
    d->run();
    return 0;
  });
}

int sf_decoder_sync(sf_decoder *d, int *status) {
  if (d == nullptr || !d->ran) {
    sfi::set_last_error("bad argument: nothing run");
    return -1;
  }
  return dguard([&]() -> int { return d->sync(status); });
}

void sf_decoder_set_timing(sf_decoder *d, int on) {
  if (d)
    d->timing = on != 0;
}

int sf_decoder_kernel_ms(sf_decoder *d, double *ms) {
  if (d == nullptr || !d->timed || ms == nullptr) {
    sfi::set_last_error("bad argument: no timed run");
    return -1;
  }
  return dguard([&]() -> int {
    check(hipEventSynchronize(d->t1), "hipEventSynchronize");
    float m = 0;
    check(hipEventElapsedTime(&m, d->t0, d->t1), "hipEventElapsedTime");
    *ms = m;
    return 0;
  });
}

int sf_decompress_image(const uint8_t *data, size_t size, const char *algorithm,
                        unsigned bit_depth, uint32_t height, uint32_t width,
                        void *device_frame, size_t device_frame_bytes) {
  return dguard([&]() -> int {
    if (device_frame_bytes < sf_frame_bytes(height, width, bit_depth) ||
        bit_depth / 8 == 0)
      throw chunk_error("output buffer too small for the frame");
    auto &d = thread_decoder();
    if (!d) {
      d.reset(sf_decoder_create(nullptr));
      if (!d)
        throw sfi::device_error("no decoder stream");
    }
    const sf_chunk c{data, size, algorithm, bit_depth, height, width, device_frame};
    d->load(&c, 1);
    d->run();
    return d->sync(nullptr);
  });
}

} // extern "C"

namespace sfi {

void decoder_load(sf_decoder *d, const sf_chunk *chunks, size_t n) {
  d->load(chunks, static_cast<uint32_t>(n));
}

void decoder_run(sf_decoder *d) { d->run(); }

int decoder_status_after(sf_decoder *d, hipStream_t read_stream, hipEvent_t after,
                         int *status, std::vector<std::string> *messages) {
  return d->status_after(read_stream, after, status, messages);
}

void check_chunk(const sf_chunk &c) {
  // The frame is only named in a plan; any aligned address stands in.
  sf_chunk probe = c;
  probe.frame = reinterpret_cast<void *>(uintptr_t(256));
  (void)plan_chunk(probe, 0);
}

} // namespace sfi
