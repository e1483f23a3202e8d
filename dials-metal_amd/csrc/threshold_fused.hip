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
constexpr uint32_t kStage = 128;        // signal pixels staged per wave
constexpr uint32_t kFast = 1u << 18;    // fast-path bound on valid pixels
constexpr uint32_t kCount0 = 1u << 26;  // count unit in stage 0's word
constexpr uint32_t kCount2 = 1u << 25;  // count unit in stage 2's word

using sfi::FusedParams;

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

// What one wave keeps in LDS: the ring of raw rows and the signal staging.
struct WaveLds {
  uint4 ring[kRing][kLanes];
  sf_signal_pixel stage[kStage];
};

// One wave's sweep of one strip segment of one frame: `item` numbers the
// frame's segments strip-fastest.
template <typename T>
__device__ void sweep(const FusedParams &a, uint32_t item, uint32_t lane,
                      WaveLds &w, const typename traits<T>::R *table) {
  using R = typename traits<T>::R;
  using Wt = typename traits<T>::W;
  const int height = static_cast<int>(a.height);
  const int width = static_cast<int>(a.width);
  const int strip = static_cast<int>(item % a.strips);
  const int seg = static_cast<int>(item / a.strips);
  const int x0 = strip * kOut;
  const int c0 = x0 - kHalo + kQ * static_cast<int>(lane); // first column
  const int y0 = seg * static_cast<int>(a.seg_rows);
  const int y1 = min(height, y0 + static_cast<int>(a.seg_rows));
  const T *frame = static_cast<const T *>(a.frame);
  uint4(*ring)[kLanes] = w.ring;
  const uint32_t masked = traits<T>::masked;

  // Signal pixels wait in LDS and go out kStage at a time, one atomic on the
  // frame's counter per flush: a returning atomic per row on that one word
  // serialises every wave of the device behind it.
  sf_signal_pixel *stage = w.stage;
  uint32_t staged = 0; // wave-uniform
  auto flush = [&]() {
    if (staged == 0)
      return;
    uint32_t base = 0;
    if (lane == 0)
      base = atomicAdd(a.counter, staged);
    base = __shfl(base, 0);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t i = lane; i < staged; i += kLanes)
      if (base + i < a.capacity)
        a.out[base + i] = stage[i];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    staged = 0;
  };
  auto put = [&](bool signal, const sf_signal_pixel &rec) {
    const uint64_t b = __ballot(signal);
    if (b == 0)
      return;
    const uint32_t n = static_cast<uint32_t>(__popcll(b));
    if (staged + n > kStage)
      flush();
    if (signal)
      stage[staged + static_cast<uint32_t>(__popcll(b & ((1ull << lane) - 1ull)))] =
          rec;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    staged += n;
  };
  bool saw_big = false;

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
    saw_big = saw_big || __ballot(big) != 0;

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
          put(signal, rec);
        }
      }
    }
    slot = slot + 1 == kRing ? 0 : slot + 1;
  }
  flush();
  if (saw_big && lane == 0)
    atomicOr(a.overflow, 1u);
}

// Every wave takes items in turn, grid-stride, over all the frames of a
// batch (item_base ascending), any mix of sizes and pixel types in one
// launch.  The grid is sized to stay resident, so the launch is persistent
// and every wave leaves when the items run out.
template <int Waves>
__global__ __launch_bounds__(Waves *kLanes) void fused_kernel(
    const FusedParams *__restrict__ frames, uint32_t nframes,
    uint32_t total_items, const float *__restrict__ tf,
    const double *__restrict__ td) {
  __shared__ WaveLds lds[Waves];
  __shared__ float table_f[128];
  __shared__ double table_d[128];
  for (int i = threadIdx.x; i < 128; i += blockDim.x) {
    table_f[i] = tf[i];
    table_d[i] = td[i];
  }
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const uint32_t lane = threadIdx.x % kLanes;
  for (uint32_t item = blockIdx.x * Waves + wave; item < total_items;
       item += gridDim.x * Waves) {
    uint32_t lo = 0, hi = nframes;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (frames[mid].item_base <= item)
        lo = mid;
      else
        hi = mid;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const FusedParams p = frames[lo];
    if (p.bits == 16)
      sweep<uint16_t>(p, item - p.item_base, lane, lds[wave], table_f);
    else
      sweep<uint32_t>(p, item - p.item_base, lane, lds[wave], table_d);
  }
}

void launch(const FusedParams *frames, uint32_t nframes, uint32_t total_items,
            const float *tf, const double *td, int waves_per_block,
            uint32_t max_waves, hipStream_t stream) {
  if (total_items == 0)
    return;
  const uint32_t waves = total_items < max_waves ? total_items : max_waves;
  switch (waves_per_block) {
  case 1:
    fused_kernel<1><<<waves, kLanes, 0, stream>>>(frames, nframes, total_items,
                                                  tf, td);
    break;
  case 4:
    fused_kernel<4><<<(waves + 3) / 4, 4 * kLanes, 0, stream>>>(
        frames, nframes, total_items, tf, td);
    break;
  default:
    fused_kernel<2><<<(waves + 1) / 2, 2 * kLanes, 0, stream>>>(
        frames, nframes, total_items, tf, td);
    break;
  }
  sfi::check(hipGetLastError(), "launching the fused threshold");
}

} // namespace

namespace sfi {

uint32_t strips_of(uint32_t width) { return (width + kOut - 1) / kOut; }

uint32_t segment_rows(uint64_t strip_rows, const Tuning &t) {
  if (t.seg_rows > 0)
    return static_cast<uint32_t>(t.seg_rows);
  // Segments tall enough that the 20 halo rows cost little, short enough
  // that the batch still spreads over the whole device: about target_waves
  // items, each 32 to 256 rows.
  const uint64_t target = static_cast<uint64_t>(t.target_waves);
  uint64_t seg = (strip_rows + target - 1) / target;
  seg = seg < 32 ? 32 : (seg > 256 ? 256 : seg);
  return static_cast<uint32_t>(seg);
}

uint32_t plan_fused(FusedPlan &p, uint32_t height, uint32_t width,
                    uint32_t seg_rows) {
  p.strips = strips_of(width);
  p.seg_rows = seg_rows;
  p.segments = (height + seg_rows - 1) / seg_rows;
  return p.strips * p.segments;
}

void launch_fused_batch(const FusedParams *frames_device, uint32_t nframes,
                        uint32_t total_items, const float *tf, const double *td,
                        hipStream_t stream) {
  const Tuning t = tuning();
  // Resident waves: sized so that every CU holds several; the items are
  // spread over them grid-stride.
  int cus = 256;
  int device = 0;
  if (hipGetDevice(&device) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                device);
  const uint32_t max_waves = static_cast<uint32_t>(cus) * 16u;
  launch(frames_device, nframes, total_items, tf, td, t.waves_per_block,
         max_waves, stream);
}

FusedParams to_params(const FrameArgs &a, const FusedPlan &plan,
                      uint32_t item_base) {
  FusedParams p;
  p.frame = a.frame;
  p.height = a.height;
  p.width = a.width;
  p.out = a.out;
  p.counter = a.counter;
  p.capacity = a.capacity;
  p.mask = a.mask;
  p.wpr = a.wpr;
  p.table = a.table;
  p.strips = plan.strips;
  p.segments = plan.segments;
  p.seg_rows = plan.seg_rows;
  p.overflow = a.overflow;
  p.bits = static_cast<uint32_t>(a.bits);
  p.item_base = item_base;
  return p;
}

} // namespace sfi
