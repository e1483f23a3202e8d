// The fused threshold: all three stages of src/spots/dext.cc:53-222 in one
// pass of each wave down a column strip of the frame, reading every pixel
// from HBM once.
//
// Geometry.  A wave owns a strip of 256 columns, four consecutive columns per
// lane, and slides down a segment of rows.  The three windows nest: the
// output needs stage 2's 11 x 11 sums (halo 5), which need stage 1's 5 x 5
// erode (halo 2 more), which needs stage 0's 7 x 7 sums (halo 3 more), so the
// strip carries 12 columns of halo on each side (10 needed, rounded to a lane)
// and writes the middle 232; rows run from 10 above the segment to 10 below.
// Pixels outside the frame load as masked, which is exact (a masked pixel adds
// nothing to a sum, is set in scr0 so it cannot break an erode, and stays out
// of stage 2's background), the same argument as dext_fused.hh:16-24.
//
// Per row entering at the bottom, the wave
//   * updates stage 0's vertical 7-row sums per column (registers), with the
//     row leaving at the top read back from a 17-row ring of raw pixels in LDS;
//   * sums 7 columns across lanes (DPP lane shifts), tests, and so finishes
//     scr0 three rows up;
//   * erodes with bit operations on nibbles of scr0 (5 columns by 5 rows), so
//     finishing scr1 five rows up;
//   * updates stage 2's vertical 11-row sums of pixels that are valid and not
//     scr1, sums 11 columns across lanes, and tests the candidates ten rows up,
//     emitting signal pixels with one atomic per wave and setting their bits
//     in the strong-pixel bitmask.
//
// Arithmetic.  The count and the sum of a window travel together in one
// 32-bit word (count in the top bits), the sum of squares in a double; both
// are exact while every valid pixel is below 2^18 (49 * (2^18-1) < 2^26 under
// a count at bit 26; 121 * (2^18-1) < 2^25 under a count at bit 25; sums of
// squares below 2^53).  A wave that meets a larger valid pixel raises the
// frame's overflow flag and the caller re-runs that frame with the direct
// kernels, whose sums are in the reference's own wide types.  uint16 frames
// never overflow.  The tests are the reference's expressions in its real type
// on the exact sums; the stage-0 root comes from a host-computed table.
#include "sf_device.hh"
#include "sf_internal.hh"

namespace {

using namespace sfd;

constexpr int kLanes = 64;
constexpr int kQ = 4;                   // columns per lane
constexpr int kStrip = kQ * kLanes;     // 256 columns per wave
constexpr int kHalo = 12;               // columns of halo each side
constexpr int kOut = kStrip - 2 * kHalo; // 232 output columns per strip
constexpr int kRing = 17;               // raw rows held: entering .. 16 above
constexpr int kWaves = 2;               // independent waves per block
constexpr uint32_t kFast = 1u << 18;    // fast-path bound on valid pixels
constexpr uint32_t kCount0 = 1u << 26;  // count unit in stage 0's word
constexpr uint32_t kCount2 = 1u << 25;  // count unit in stage 2's word

struct FusedParams {
  const void *frame;
  uint32_t height, width;
  sf_signal_pixel *out;
  uint32_t *counter;
  uint32_t capacity;
  uint32_t *mask;
  uint32_t wpr;
  const void *table;
  uint32_t strips, segments, seg_rows;
  uint32_t *overflow;
};

// Lane shifts across the whole wave.  DPP wave_shr:1 gives lane i the value of
// lane i-1 (lane 0 gets 0); wave_shl:1 gives lane i the value of lane i+1
// (lane 63 gets 0).  The edge lanes' results are halo and never used.
__device__ __forceinline__ uint32_t from_left(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ uint32_t from_right(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ double from_left(double v) {
  const uint64_t b = __double_as_longlong(v);
  const uint32_t lo = from_left(static_cast<uint32_t>(b));
  const uint32_t hi = from_left(static_cast<uint32_t>(b >> 32));
  return __longlong_as_double(static_cast<long long>(
      (static_cast<uint64_t>(hi) << 32) | lo));
}
__device__ __forceinline__ double from_right(double v) {
  const uint64_t b = __double_as_longlong(v);
  const uint32_t lo = from_right(static_cast<uint32_t>(b));
  const uint32_t hi = from_right(static_cast<uint32_t>(b >> 32));
  return __longlong_as_double(static_cast<long long>(
      (static_cast<uint64_t>(hi) << 32) | lo));
}

template <typename T> struct Row {
  uint32_t v[kQ];
};

// Loads this lane's four pixels of row y; outside the frame they are masked.
template <typename T>
__device__ __forceinline__ void load_row(const T *__restrict__ frame, int height,
                                         int width, int y, int c0,
                                         uint32_t (&v)[kQ]) {
  const uint32_t masked = traits<T>::masked;
  if (y < 0 || y >= height) {
#pragma unroll
    for (int q = 0; q < kQ; q++)
      v[q] = masked;
    return;
  }
  const T *row = frame + static_cast<size_t>(y) * width;
#pragma unroll
  for (int q = 0; q < kQ; q++) {
    const int c = c0 + q;
    v[q] = (c >= 0 && c < width) ? static_cast<uint32_t>(row[c]) : masked;
  }
}

template <typename T>
__device__ __forceinline__ uint32_t pack0(uint32_t v) {
  return v < traits<T>::masked ? v + kCount0 : 0u;
}
template <typename T> __device__ __forceinline__ double square(uint32_t v) {
  const double p = v < traits<T>::masked ? static_cast<double>(v) : 0.0;
  return p * p;
}
template <typename T>
__device__ __forceinline__ uint32_t pack2(uint32_t v, uint32_t eroded) {
  return (v < traits<T>::masked && !eroded) ? v + kCount2 : 0u;
}

// Stage 0's test on the fast path's exact sums, in the reference's real type.
template <typename T>
__device__ __forceinline__ bool stage0_test(uint32_t word, double sum_sq,
                                            const typename traits<T>::R *table);

template <>
__device__ __forceinline__ bool stage0_test<uint32_t>(uint32_t word,
                                                      double sum_sq,
                                                      const double *table) {
  const uint32_t count = word >> 26;
  const double m = static_cast<double>(count);
  const double s = static_cast<double>(word & (kCount0 - 1));
  const double lhs = (m * sum_sq - s * s) - s * (m - 1.0);
  const double rhs = (s * 6.0) * table[count]; // +inf below two: false
  return lhs > rhs;
}

template <>
__device__ __forceinline__ bool stage0_test<uint16_t>(uint32_t word,
                                                      double sum_sq,
                                                      const float *table) {
  const uint32_t count = word >> 26;
  const float m = static_cast<float>(count);
  const float s = static_cast<float>(word & (kCount0 - 1));
  // The reference's accumulator is uint32 and wraps; the exact sum here is
  // below 2^38, so its low 32 bits are the wrapped value.
  const uint32_t wrapped =
      static_cast<uint32_t>(static_cast<uint64_t>(sum_sq));
  const float s2 = static_cast<float>(wrapped);
  const float lhs = (m * s2 - s * s) - s * (m - 1.0f);
  const float rhs = (s * 6.0f) * table[count];
  return lhs > rhs;
}

template <typename T>
__global__ __launch_bounds__(kWaves *kLanes) void fused_kernel(FusedParams a) {
  using R = typename traits<T>::R;
  using Wt = typename traits<T>::W;
  __shared__ uint4 ring_all[kWaves][kRing][kLanes];
  __shared__ R table[128];
  for (int i = threadIdx.x; i < 128; i += blockDim.x)
    table[i] = static_cast<const R *>(a.table)[i];
  __syncthreads();

  const int wave = threadIdx.x / kLanes;
  const uint32_t lane = threadIdx.x % kLanes;
  const uint32_t item = blockIdx.x * kWaves + wave;
  if (item >= a.strips * a.segments)
    return;
  const int height = static_cast<int>(a.height);
  const int width = static_cast<int>(a.width);
  const int strip = static_cast<int>(item % a.strips);
  const int seg = static_cast<int>(item / a.strips);
  const int x0 = strip * kOut;
  const int c0 = x0 - kHalo + kQ * static_cast<int>(lane); // first column
  const int y0 = seg * static_cast<int>(a.seg_rows);
  const int y1 = min(height, y0 + static_cast<int>(a.seg_rows));
  const T *frame = static_cast<const T *>(a.frame);
  uint4(*ring)[kLanes] = ring_all[wave];
  const uint32_t masked = traits<T>::masked;

  // Output columns of this lane inside the strip's middle and the frame.
  uint32_t in_out = 0;
#pragma unroll
  for (int q = 0; q < kQ; q++) {
    const int c = c0 + q;
    if (c >= x0 && c < x0 + kOut && c < width)
      in_out |= 1u << q;
  }

  uint32_t v0w[kQ] = {0, 0, 0, 0};   // stage 0 vertical: count | sum
  double v0s[kQ] = {0.0, 0.0, 0.0, 0.0}; // stage 0 vertical: sum of squares
  uint32_t v2w[kQ] = {0, 0, 0, 0};   // stage 2 vertical: count | sum
  uint32_t mhist = 0;  // masked nibbles, row y_in - k at nibble k
  uint32_t hreg = 0;   // horizontal-erode nibbles, row yc - k at nibble k
  uint64_t s1hist = 0; // scr1 nibbles, row ye - k at nibble k

  const int steps = (y1 - y0) + 20;
  uint32_t next[kQ];
  load_row<T>(frame, height, width, y0 - 10, c0, next);
  int slot = 0; // ring slot of the entering row
  for (int t = 0; t < steps; t++) {
    const int y_in = y0 - 10 + t;
    uint32_t cur[kQ];
#pragma unroll
    for (int q = 0; q < kQ; q++)
      cur[q] = next[q];
    if (t + 1 < steps)
      load_row<T>(frame, height, width, y_in + 1, c0, next);

    // Fast-path guard: a valid pixel at or above 2^18 voids this frame's
    // fast sums; the caller re-runs it exactly.
    bool big = false;
#pragma unroll
    for (int q = 0; q < kQ; q++)
      big |= cur[q] < masked && cur[q] >= kFast;
    if (__ballot(big) != 0 && lane == 0)
      atomicOr(a.overflow, 1u);

    ring[slot][lane] = make_uint4(cur[0], cur[1], cur[2], cur[3]);
    uint32_t mnib = 0;
#pragma unroll
    for (int q = 0; q < kQ; q++)
      mnib |= (cur[q] >= masked ? 1u : 0u) << q;
    mhist = (mhist << 4) | mnib;

    auto at = [&](int back) {
      int s = slot - back;
      if (s < 0)
        s += kRing;
      return ring[s][lane];
    };

    // ---- stage 0: vertical 7-row sums (rows y_in-6 .. y_in) ----
    {
      uint32_t out[kQ] = {masked, masked, masked, masked};
      if (t >= 7) {
        const uint4 o = at(7);
        out[0] = o.x, out[1] = o.y, out[2] = o.z, out[3] = o.w;
      }
#pragma unroll
      for (int q = 0; q < kQ; q++) {
        v0w[q] += pack0<T>(cur[q]) - pack0<T>(out[q]);
        v0s[q] += square<T>(cur[q]) - square<T>(out[q]);
      }
    }

    // ---- stage 0: 7 columns across lanes, test -> scr0 of row y_in-3 ----
    uint32_t scr0 = 0;
    {
      const uint32_t lw1 = from_left(v0w[1]), lw2 = from_left(v0w[2]),
                     lw3 = from_left(v0w[3]);
      const uint32_t rw0 = from_right(v0w[0]), rw1 = from_right(v0w[1]),
                     rw2 = from_right(v0w[2]);
      const double ls1 = from_left(v0s[1]), ls2 = from_left(v0s[2]),
                   ls3 = from_left(v0s[3]);
      const double rs0 = from_right(v0s[0]), rs1 = from_right(v0s[1]),
                   rs2 = from_right(v0s[2]);
      uint32_t hw[kQ];
      double hs[kQ];
      hw[0] = lw1 + lw2 + lw3 + v0w[0] + v0w[1] + v0w[2] + v0w[3];
      hw[1] = hw[0] - lw1 + rw0;
      hw[2] = hw[1] - lw2 + rw1;
      hw[3] = hw[2] - lw3 + rw2;
      hs[0] = ((ls1 + ls2) + (ls3 + v0s[0])) + ((v0s[1] + v0s[2]) + v0s[3]);
      hs[1] = (hs[0] - ls1) + rs0;
      hs[2] = (hs[1] - ls2) + rs1;
      hs[3] = (hs[2] - ls3) + rs2;
      const uint32_t m3 = (mhist >> 12) & 0xFu; // masked nibble of row y_in-3
#pragma unroll
      for (int q = 0; q < kQ; q++)
        if (stage0_test<T>(hw[q], hs[q], table))
          scr0 |= 1u << q;
      scr0 |= m3;
    }

    // ---- stage 1: erode 5 x 5 -> scr1 of row y_in-5 ----
    uint32_t scr1;
    {
      const uint32_t left = from_left(scr0) & 0xFu;
      const uint32_t right = from_right(scr0) & 0xFu;
      // bit i is column c0 - 4 + i
      const uint32_t x = left | (scr0 << 4) | (right << 8);
      const uint32_t y = x & (x >> 1) & (x >> 2) & (x >> 3) & (x >> 4);
      hreg = (hreg << 4) | ((y >> 2) & 0xFu);
      const uint32_t all5 =
          hreg & (hreg >> 4) & (hreg >> 8) & (hreg >> 12) & (hreg >> 16) & 0xFu;
      const uint32_t m5 = (mhist >> 20) & 0xFu; // masked nibble of row y_in-5
      scr1 = all5 | m5;
      s1hist = (s1hist << 4) | scr1;
    }

    // ---- stage 2: vertical 11-row sums of valid, not-scr1 pixels ----
    if (t >= 5) {
      const uint4 e = at(5); // row y_in-5, whose scr1 is nibble 0
      v2w[0] += pack2<T>(e.x, scr1 & 1u);
      v2w[1] += pack2<T>(e.y, scr1 & 2u);
      v2w[2] += pack2<T>(e.z, scr1 & 4u);
      v2w[3] += pack2<T>(e.w, scr1 & 8u);
    }
    if (t >= 16) {
      const uint4 o = at(16); // row y_in-16, whose scr1 is nibble 11
      const uint32_t s = static_cast<uint32_t>(s1hist >> 44) & 0xFu;
      v2w[0] -= pack2<T>(o.x, s & 1u);
      v2w[1] -= pack2<T>(o.y, s & 2u);
      v2w[2] -= pack2<T>(o.z, s & 4u);
      v2w[3] -= pack2<T>(o.w, s & 8u);
    }

    // ---- stage 2: 11 columns across lanes, candidates of row y_in-10 ----
    if (t >= 20) {
      const int y2 = y_in - 10;
      const uint32_t s1 = static_cast<uint32_t>(s1hist >> 20) & 0xFu;
      const uint32_t maybe = s1 & in_out;
      if (__ballot(maybe != 0) != 0) {
        const uint32_t l0 = from_left(v2w[0]), l1 = from_left(v2w[1]),
                       l2 = from_left(v2w[2]), l3 = from_left(v2w[3]);
        const uint32_t r0 = from_right(v2w[0]), r1 = from_right(v2w[1]),
                       r2 = from_right(v2w[2]), r3 = from_right(v2w[3]);
        const uint32_t ll3 = from_left(l3), rr0 = from_right(r0);
        uint32_t h[kQ];
        h[0] = ll3 + l0 + l1 + l2 + l3 + v2w[0] + v2w[1] + v2w[2] + v2w[3] +
               r0 + r1;
        h[1] = h[0] - ll3 + r2;
        h[2] = h[1] - l0 + r3;
        h[3] = h[2] - l1 + rr0;
        const uint4 pv = at(10);
        const uint32_t raw[kQ] = {pv.x, pv.y, pv.z, pv.w};
#pragma unroll
        for (int q = 0; q < kQ; q++) {
          const uint32_t p = raw[q] >= masked ? 0u : raw[q];
          bool signal = false;
          sf_signal_pixel rec{};
          if (((maybe >> q) & 1u) && p > 0) {
            const uint32_t m2 = h[q] >> 25;
            const uint32_t s2 = h[q] & (kCount2 - 1);
            R mean;
            signal = poisson<Wt, R>(p, static_cast<Wt>(m2), static_cast<Wt>(s2),
                                    &mean);
            const int c = c0 + q;
            rec.index = static_cast<uint32_t>(y2) * a.width +
                        static_cast<uint32_t>(c);
            rec.value = p;
            rec.background = static_cast<float>(mean);
            rec.population = static_cast<uint16_t>(m2);
            rec.reserved = 0;
            if (signal)
              atomicOr(&a.mask[static_cast<size_t>(y2) * a.wpr + c / 32],
                       1u << (c & 31));
          }
          emit_wave(signal, rec, a.out, a.counter, a.capacity);
        }
      }
    }
    slot = slot + 1 == kRing ? 0 : slot + 1;
  }
}

template <typename T>
void run(const sfi::FrameArgs &f, uint32_t *overflow, hipStream_t stream) {
  FusedParams p;
  p.frame = f.frame;
  p.height = f.height;
  p.width = f.width;
  p.out = f.out;
  p.counter = f.counter;
  p.capacity = f.capacity;
  p.mask = f.mask;
  p.wpr = f.wpr;
  p.table = f.table;
  p.strips = (f.width + kOut - 1) / kOut;
  // Segments tall enough that the 20 halo rows cost little, short enough
  // that a single frame still spreads over the whole device.
  const uint32_t target_items = 2048;
  uint32_t seg = (f.height * p.strips + target_items - 1) / target_items;
  if (seg < 64)
    seg = 64;
  p.seg_rows = seg;
  p.segments = (f.height + seg - 1) / seg;
  p.overflow = overflow;
  const uint32_t items = p.strips * p.segments;
  const uint32_t blocks = (items + kWaves - 1) / kWaves;
  fused_kernel<T><<<blocks, kWaves * kLanes, 0, stream>>>(p);
  sfi::check(hipGetLastError(), "launching the fused threshold");
}

} // namespace

namespace sfi {

void launch_fused(const FrameArgs &a, hipStream_t stream) {
  // The caller clears the bitmask (it is or-ed into bit by bit) and the
  // overflow word before this launch.
  if (a.bits == 16)
    run<uint16_t>(a, a.overflow, stream);
  else
    run<uint32_t>(a, a.overflow, stream);
}

} // namespace sfi
