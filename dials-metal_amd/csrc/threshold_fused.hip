// The fused threshold: all three stages of src/spots/dext.cc:53-222 in one
// pass of each wave down a column strip of the frame, reading every pixel
// from HBM once.
//
// Geometry.  A wave owns a strip of 64 KQ columns, KQ consecutive columns per
// lane (KQ is 4 or 8, a launch-shape choice), and slides down a segment of
// rows.  The three windows nest: the output needs stage 2's 11 x 11 sums (halo
// 5), which need stage 1's 5 x 5 erode (halo 2 more), which needs stage 0's
// 7 x 7 sums (halo 3 more), so the strip carries 10 columns of halo on each
// side and writes the middle 64 KQ - 20; rows run from 10 above the segment to
// 10 below.  Pixels outside the frame count as masked, which is exact (a
// masked pixel adds nothing to a sum, is set in scr0 so it cannot break an
// erode, and stays out of stage 2's background), the same argument as
// dext_fused.hh:16-24.
//
// Representation.  A masked pixel is held as the value 0 with its bit in a
// register history of masked nibbles (one nibble a row, a bit a column, one
// history per four columns), so every window sum is a plain sum of values and
// a window's count is the window's size less the masked bits in it.  The
// entering row is read from HBM two rows ahead of use; the rows leaving the
// windows are read back from a 16-row ring in LDS that holds each pixel's low
// 16 bits, with bits 16-17 in a second ring written and read only for the
// rare rows that have such a pixel (a wave-uniform row mask says which).
// Sixteen rows suffice because the row 16 above is read before the entering
// row takes its slot.  The ring starts at zero, so the rows before the first
// read back as nothing and no step needs a prologue test.
//
// Per row entering at the bottom, the wave
//   * updates stage 0's vertical 7-row sums of values and of squares per
//     column (registers), with the square's change as one 24-bit product
//     (in - out)(in + out) modulo 2^32;
//   * sums 7 columns across lanes (DPP lane shifts) and tests, finishing scr0
//     three rows up.  Where no masked pixel lies in any of the wave's windows
//     (a wave-uniform test) every count is 49; elsewhere the counts come from
//     the masked history by popcount and travel in the top bits of the sums;
//   * erodes with bit operations on scr0 (5 columns by 5 rows), finishing
//     scr1 five rows up, and records which pixels stage 2 leaves out (masked
//     or in scr1) in an exclusion history;
//   * updates stage 2's vertical 11-row sums of values outside scr1 (only
//     unmasked scr1 pixels need excluding, the masked ones being zero, and a
//     wave-uniform record of the rows holding any skips even that for the
//     rest), and, on rows with a candidate ten rows up, forms the counts from
//     the exclusion history, sums 11 columns across lanes and tests, emitting
//     signal pixels to a per-wave LDS stage (one atomic per flush) and setting
//     their bits in the strong-pixel bitmask.
//
// Arithmetic.  Sums of values are exact in 32 bits while every valid pixel is
// below 2^18 (49 * (2^18-1) < 2^26 under a count at bit 26; 121 * (2^18-1) <
// 2^25 under a count at bit 25).  Sums of squares are 32-bit too: while the
// window's pixels are below 2^13, 49 squares below 2^26 cannot wrap; a row
// with a pixel of 2^13 or more (a wave-uniform record says which) moves its
// squares' parts from bit 26 up into a second sum, and the test rebuilds
// the exact sum (below 2^53) as a double from the two.  A wave that meets a
// valid pixel of 2^18 or more raises the frame's overflow flag and the
// caller re-runs that frame with the direct kernels, whose sums are in the
// reference's own wide types.  uint16 frames never overflow.  The tests are the reference's expressions in its real type
// on the exact sums; for uint32 frames every intermediate of the left side is
// an integer below 2^53, so it is formed with a fused multiply-add that is
// exact, and the right side keeps the reference's two roundings; the stage-0
// root comes from a host-computed table.
#include <type_traits>
#include <utility>

#include "sf_device.hh"
#include "sf_internal.hh"

namespace {

using namespace sfd;

constexpr int kLanes = 64;
constexpr int kHalo = 10;               // columns of halo each side: 3 + 2 + 5
constexpr int kAhead = 2;               // rows a wave keeps in flight from HBM
constexpr int kRing = 16;               // raw rows held: entering .. 15 above
constexpr uint32_t kStage = 48;         // candidates staged per wave
constexpr uint32_t kChunk = 8;          // consecutive items dealt together
constexpr uint32_t kFast = 1u << 18;    // fast-path bound on valid pixels
constexpr uint32_t kWide = 1u << 16;    // pixels at or above this use the high ring
constexpr int kTable = 64;              // root table entries used (m <= 49)
constexpr uint32_t kCount0 = 1u << 26;  // count unit in stage 0's word
constexpr uint32_t kCount2 = 1u << 25;  // count unit in stage 2's word
constexpr uint32_t kHot = 1u << 13;     // uint32 pixels from here split their squares
constexpr uint32_t kSplit = 1u << 26;   // where a square splits into low and high parts
// Strips start kPhase columns left of a multiple of their output width, so
// that every lane's first column is a multiple of four: a lane is then wholly
// inside or outside a frame whose width is a multiple of four, and the
// lattice columns 4k + 2 of the frame lie two columns before a module gap
// that starts on a multiple of four and two after one that ends before one.
constexpr int kPhase = 2;

// Output columns of one strip with KQ columns a lane.
__host__ __device__ constexpr int out_columns(int kq) {
  return kq * kLanes - 2 * kHalo;
}

using sfi::FusedParams;

// Lane shifts across the whole wave.  DPP wave_shr:1 gives lane i the value of
// lane i-1 (lane 0 gets 0); wave_shl:1 gives lane i the value of lane i+1
// (lane 63 gets 0): bound_ctrl makes a source lane outside the wave read as
// zero, so no old value is needed and an add can take the shift as its
// operand.  The edge lanes' results are halo and never used.
__device__ __forceinline__ uint32_t from_left(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ uint32_t from_right(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x130, 0xf, 0xf, true));
}
// The value that lane (this lane + D) holds.
template <int D, typename V> __device__ __forceinline__ V from_lane(V v) {
  if constexpr (D < 0)
    return from_lane<D + 1>(from_left(v));
  else if constexpr (D > 0)
    return from_lane<D - 1>(from_right(v));
  else
    return v;
}

// The lane, relative to this one, that holds column c0 + I.
template <int KQ, int I> constexpr int lane_of() {
  return I >= 0 ? I / KQ : -((-I + KQ - 1) / KQ);
}

// Sum of columns [A, B] of v, numbered from this lane's first column: one
// partial sum per lane that holds some of them, moved across by DPP (the
// move folds into the add that takes it).  Partial sums run from each lane's
// last column down, so that the sums shared between windows are formed once.
template <int KQ, int Lo, int Hi>
__device__ __forceinline__ uint32_t run_sum(const uint32_t (&v)[KQ]) {
  if constexpr (Lo == Hi)
    return v[Lo];
  else
    return v[Lo] + run_sum<KQ, Lo + 1, Hi>(v);
}
template <int KQ, int A, int B, int D = lane_of<KQ, A>()>
__device__ __forceinline__ uint32_t window_sum(const uint32_t (&v)[KQ]) {
  constexpr int first = D * KQ;
  constexpr int lo = (A > first ? A : first) - first;
  constexpr int hi = (B < first + KQ - 1 ? B : first + KQ - 1) - first;
  const uint32_t part = from_lane<D>(run_sum<KQ, lo, hi>(v));
  if constexpr (first + KQ - 1 >= B)
    return part;
  else
    return window_sum<KQ, A, B, D + 1>(v) + part;
}

// The value of column I (numbered as above).
template <int KQ, int I>
__device__ __forceinline__ uint32_t column(const uint32_t (&v)[KQ]) {
  return from_lane<lane_of<KQ, I>()>(v[I - lane_of<KQ, I>() * KQ]);
}

// One window's sum from its neighbour's: o - out + in for the columns Out
// leaving and In entering, formed as in - (out - o), so that each lane shift
// is the minuend of a subtraction and folds into it as v_sub_u32 with DPP.
// The other order folds into v_subrev_u32 with DPP, which on this device
// took the wrong operand in a probe of exactly this code; the empty asm
// keeps the compiler from reassociating back into it.
template <int KQ, int Out, int In>
__device__ __forceinline__ uint32_t slide(uint32_t o, const uint32_t (&v)[KQ]) {
  uint32_t t = column<KQ, Out>(v) - o;
  __asm__("" : "+v"(t));
  return column<KQ, In>(v) - t;
}

// Window sums of radius R at every column of the lane, given those at the
// lattice columns 4h + 2: each next one from its neighbour by the column
// leaving and the one entering.  The values are integers modulo 2^32, so the
// order of the additions does not change the sums.
template <int R, int KQ, int G>
__device__ __forceinline__ void slide_group(const uint32_t (&v)[KQ],
                                            uint32_t (&o)[KQ]) {
  constexpr int c = 4 * G + 2;
  o[c - 1] = slide<KQ, c + R, c - 1 - R>(o[c], v);
  o[c - 2] = slide<KQ, c - 1 + R, c - 2 - R>(o[c - 1], v);
  o[c + 1] = slide<KQ, c - R, c + 1 + R>(o[c], v);
}
template <int R, int KQ, int... G>
__device__ __forceinline__ void slide_all(const uint32_t (&v)[KQ],
                                          uint32_t (&o)[KQ],
                                          std::integer_sequence<int, G...>) {
  (slide_group<R, KQ, G>(v, o), ...);
}
template <int R, int KQ, int... G>
__device__ __forceinline__ void lattice_sums(const uint32_t (&v)[KQ],
                                             uint32_t (&o)[KQ],
                                             std::integer_sequence<int, G...>) {
  ((o[4 * G + 2] = window_sum<KQ, 4 * G + 2 - R, 4 * G + 2 + R>(v)), ...);
}
template <int R, int KQ>
__device__ __forceinline__ void row_sums(const uint32_t (&v)[KQ],
                                         uint32_t (&o)[KQ]) {
  lattice_sums<R>(v, o, std::make_integer_sequence<int, KQ / 4>{});
  slide_all<R>(v, o, std::make_integer_sequence<int, KQ / 4>{});
}

// A lane's KQ pixels as vector loads; dword alignment is all the hardware
// asks of a multi-dword global load (and uint16 rows get none: unaligned
// global access is on).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4), aligned(4)));
typedef uint16_t u16x4 __attribute__((ext_vector_type(4), aligned(2)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8), aligned(2)));
typedef uint32_t u32x4a __attribute__((ext_vector_type(4))); // 16-byte aligned

// Loads KQ pixels of in-frame row y from column `cl`, where cl + KQ <= width.
// One load instruction on every path, so the wait for a row can count the
// loads issued since.
template <typename T, int KQ>
__device__ __forceinline__ void load_lane(const SF_GLOBAL T *__restrict__ frame,
                                          int width, int y, int cl,
                                          uint32_t (&v)[KQ]) {
  const SF_GLOBAL T *row = frame + static_cast<size_t>(y) * width;
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int k = 0; k < KQ; k += 4) {
      const u32x4 x = *reinterpret_cast<const SF_GLOBAL u32x4 *>(row + cl + k);
      v[k] = x[0], v[k + 1] = x[1], v[k + 2] = x[2], v[k + 3] = x[3];
    }
  } else if constexpr (KQ == 8) {
    const u16x8 x = *reinterpret_cast<const SF_GLOBAL u16x8 *>(row + cl);
#pragma unroll
    for (int q = 0; q < 8; q++)
      v[q] = x[q];
  } else {
#pragma unroll
    for (int k = 0; k < KQ; k += 4) {
      const u16x4 x = *reinterpret_cast<const SF_GLOBAL u16x4 *>(row + cl + k);
      v[k] = x[0], v[k + 1] = x[1], v[k + 2] = x[2], v[k + 3] = x[3];
    }
  }
}

// A lane with columns outside the frame loaded from the clamped column cl =
// c0 - sh: its column q is x[q + sh], and its outside columns read as 0.
template <int KQ>
__device__ __forceinline__ void shift_lane(uint32_t (&x)[KQ], int sh,
                                           uint32_t outside) {
  // The shift's conditions are formed again for every row (the empty asm
  // hides that they are the same each time): kept across the loop, they
  // would hold a pair of scalar registers each.
  __asm__("" : "+v"(sh), "+v"(outside));
  uint32_t r[KQ];
#pragma unroll
  for (int q = 0; q < KQ; q++) {
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < KQ; j++)
      v = q + sh == j ? x[j] : v;
    r[q] = (outside >> q) & 1u ? 0u : v;
  }
#pragma unroll
  for (int q = 0; q < KQ; q++)
    x[q] = r[q];
}

// The high ring's form of a lane's pixels (each below 2^18): bits 16-17 of
// column q at bits 2q, 2q + 1.
template <int KQ>
__device__ __forceinline__ uint32_t pack_high(const uint32_t (&v)[KQ]) {
  uint32_t h = 0;
#pragma unroll
  for (int q = 0; q < KQ; q++)
    h |= (v[q] >> 16) << (2 * q);
  return h;
}
template <int KQ>
__device__ __forceinline__ void add_high(uint32_t h, uint32_t (&v)[KQ]) {
#pragma unroll
  for (int q = 0; q < KQ; q++)
    v[q] |= ((h >> (2 * q)) & 3u) << 16;
}

// 1 when `c` holds in any lane, 0 otherwise: a wave-uniform word formed in
// scalar registers (the ballot's population count, at most one).
__device__ __forceinline__ uint32_t bit_any(bool c) {
  return min(static_cast<uint32_t>(__popcll(__ballot(c))), 1u);
}

// Stage 0's test (dext.cc:99-110) on exact window sums: count m (with m - 1,
// both in the real type), sum S of the values and the sum of the squares s2
// in the real type.  For uint16 frames s2 is the reference's wrapped uint32
// sum Q; for uint32 frames Q is the sum of the squares' low 26 bits and H
// that of their high parts, nonzero only near a pixel of 2^13 or more.  The
// root is R(sqrt(2 (m - 1))), +inf below two, so that m < 2 tests false.
template <typename T> struct Stage0;

template <> struct Stage0<uint32_t> {
  __device__ static __forceinline__ double squares(uint32_t Q) {
    return static_cast<double>(Q);
  }
  __device__ static __forceinline__ double squares(uint32_t Q, uint32_t H) {
    return __fma_rn(static_cast<double>(H), 0x1p26, static_cast<double>(Q));
  }
  __device__ static __forceinline__ bool test(double m, double mm1, uint32_t S,
                                              double s2, double root) {
    const double s = static_cast<double>(S);
    // (m * s2 - s * s) - s * (m - 1): every term and partial result is an
    // integer below 2^53, so this one-rounding form is the same exact value.
    const double lhs = __fma_rn(m, s2, -(s * (s + mm1)));
    const double rhs = (s * 6.0) * root;
    return lhs > rhs;
  }
};

template <> struct Stage0<uint16_t> {
  // The reference's accumulator is uint32 and wraps; Q is that wrapped sum.
  __device__ static __forceinline__ float squares(uint32_t Q) {
    return static_cast<float>(Q);
  }
  __device__ static __forceinline__ float squares(uint32_t Q, uint32_t) {
    return static_cast<float>(Q);
  }
  __device__ static __forceinline__ bool test(float m, float mm1, uint32_t S,
                                              float s2, float root) {
    const float s = static_cast<float>(S);
    const float lhs = (m * s2 - s * s) - s * mm1;
    const float rhs = (s * 6.0f) * root;
    return lhs > rhs;
  }
};

// What one wave keeps in LDS: the rings of raw rows (the low ring column by
// column, so that a lane's 16-bit reads fall in distinct banks) and the
// staged stage-2 candidates.  KQ = 4 takes 9984 bytes (four blocks of four
// waves fit a CU's 160 KB), KQ = 8 takes 19200 (two blocks of four).
template <int KQ> struct WaveLds {
  using High = std::conditional_t<KQ == 4, uint8_t, uint16_t>;
  uint16_t ring[kRing][KQ][kLanes];
  High high[kRing][kLanes];
  uint4 stage[kStage];
};

// The wave-uniform row records of a sweep, in one 64-bit word: each field is
// a history of one bit a row, the newest at the field's first bit, and the
// whole word shifts by one at the start of a row (kRecKeep clears the bit
// each field's neighbour below shifts into it), so that every test is one
// scalar AND.  Field by field, row r - k at bit k of the field:
//   kRecW  rows with a pixel at or above 2^16, which the high ring holds;
//          r = y_in, k <= 16;
//   kRecH  rows with a pixel at or above 2^13 (uint32 frames: their squares
//          are split); r = y_in, k <= 7;
//   kRecY  rows with a horizontal-erode bit anywhere in the wave;
//          r = y_in - 3, k <= 4;
//   kRecU  of them, rows with such a bit at a pixel unmasked in its row;
//          r = y_in - 3, k <= 2;
//   kRecX  rows with an unmasked scr1 pixel anywhere in the wave;
//          r = y_in - 5, k <= 11.
constexpr int kRecW = 0, kRecH = 17, kRecY = 25, kRecU = 30, kRecX = 33;
constexpr uint64_t kRecKeep = ~((1ull << kRecW) | (1ull << kRecH) |
                                (1ull << kRecY) | (1ull << kRecU) |
                                (1ull << kRecX));
__host__ __device__ constexpr uint64_t rec_bit(int field, int k) {
  return 1ull << (field + k);
}

// Bits of column q in rows of a nibble history: nibbles k0 .. k0 + n - 1.
__host__ __device__ constexpr uint64_t column_bits(int k0, int n, int q) {
  uint64_t m = 0;
  for (int k = k0; k < k0 + n; k++)
    m |= 1ull << (4 * k + q);
  return m;
}

// One wave's sweep of one strip segment of one frame: `item` numbers the
// frame's segments strip-fastest.
template <typename T, int KQ>
__device__ __forceinline__ void sweep(const FusedParams &a, uint32_t item,
                                      uint32_t lane, WaveLds<KQ> &w,
                                      const typename traits<T>::R *table) {
  using R = typename traits<T>::R;
  using Wt = typename traits<T>::W;
  using S0 = Stage0<T>;
  constexpr int P = KQ / 4;          // nibble histories, one per 4 columns
  constexpr int kOut = out_columns(KQ);
  constexpr bool kU32 = sizeof(T) == 4;
  constexpr uint32_t kAll = (1u << KQ) - 1u;
  const int height = static_cast<int>(a.height);
  const int width = static_cast<int>(a.width);
  const int strip = static_cast<int>(item % a.strips);
  const int seg = static_cast<int>(item / a.strips);
  const int x0 = strip * kOut - kPhase;
  const int c0 = x0 - kHalo + KQ * static_cast<int>(lane); // first column
  const int y0 = seg * static_cast<int>(a.seg_rows);
  const int y1 = min(height, y0 + static_cast<int>(a.seg_rows));
  const SF_GLOBAL T *frame = gmem(static_cast<const T *>(a.frame));
  // records move as 16-byte vectors
  SF_GLOBAL u32x4a *out = reinterpret_cast<SF_GLOBAL u32x4a *>(gmem(a.out));
  SF_GLOBAL uint32_t *mask = gmem(a.mask);
  uint16_t(*ring)[KQ][kLanes] = w.ring;
  auto *high = w.high;
  auto ring_row = [&](int s, uint32_t(&v)[KQ]) {
#pragma unroll
    for (int q = 0; q < KQ; q++)
      v[q] = ring[s][q][lane];
  };
  const uint32_t masked = traits<T>::masked;
  // A row whose pixels are all below this holds no masked pixel, none that
  // needs the high ring or the exact path, and (uint32 frames) no square of
  // 2^26 or more.
  const uint32_t plain = kU32 ? kHot : masked;

  // Stage-2 candidates wait in LDS, kStage (one a lane) at a time, each as
  // {index, value | mask bit << 18, count | sum word, mask word}.  A flush
  // tests them one a lane, so the test's divide and square root run at full
  // width, and appends the signal pixels with one atomic on the frame's
  // counter: a returning atomic per row on that one word serialises every
  // wave of the device behind it.
  uint4 *stage = w.stage;
  uint32_t staged = 0; // wave-uniform
  auto flush = [&]() {
    if (staged == 0)
      return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bool signal = false;
    u32x4a rec = {0u, 0u, 0u, 0u};
    uint32_t mword = 0, mbit = 0;
    if (lane < staged) {
      const uint4 c = stage[lane];
      const uint32_t p = c.y & (kFast - 1);
      const uint32_t m2 = c.z >> 25;
      R mean;
      signal = poisson<Wt, R>(p, static_cast<Wt>(m2),
                              static_cast<Wt>(c.z & (kCount2 - 1)), &mean);
      // {index, value, background, population | reserved 0}
      rec = u32x4a{c.x, p, __float_as_uint(static_cast<float>(mean)), m2};
      mword = c.w;
      mbit = c.y >> 18;
    }
    const uint64_t b = __ballot(signal);
    if (b != 0) {
      uint32_t base = 0;
      if (lane == 0)
        base = __hip_atomic_fetch_add(gmem(a.counter),
                                      static_cast<uint32_t>(__popcll(b)),
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      base = __shfl(base, 0);
      if (signal) {
        const uint32_t i =
            base + static_cast<uint32_t>(__popcll(b & ((1ull << lane) - 1ull)));
        if (i < a.capacity)
          out[i] = rec;
        __hip_atomic_fetch_or(mask + mword, 1u << mbit, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
      }
      // The ordering scan's block totals: the bits set in each block of
      // kScanPerBlock mask words, one atomic per block the flush touches
      // (its pixels lie in a few rows, so one or two blocks).
      const uint32_t block = mword / static_cast<uint32_t>(sfi::kScanPerBlock);
      for (uint64_t todo = b; todo != 0;) {
        const uint32_t first = static_cast<uint32_t>(__builtin_ctzll(todo));
        const uint32_t fb = static_cast<uint32_t>(__builtin_amdgcn_readlane(
            static_cast<int>(block), static_cast<int>(first)));
        const uint64_t same = __ballot(signal && block == fb) & todo;
        if (lane == first)
          __hip_atomic_fetch_add(gmem(a.totals) + fb,
                                 static_cast<uint32_t>(__popcll(same)),
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        todo &= ~same;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    staged = 0;
  };
  // One ballot can hold more candidates (up to a wave) than the stage has
  // room for, so they go in chunks that fit, flushing in between: `staged`
  // never exceeds kStage, and a flush reads only entries written since the
  // last one.
  auto put = [&](bool cand, const uint4 &entry) {
    const uint64_t b = __ballot(cand);
    if (b == 0)
      return;
    const uint32_t n = static_cast<uint32_t>(__popcll(b));
    const uint32_t rank =
        static_cast<uint32_t>(__popcll(b & ((1ull << lane) - 1ull)));
    for (uint32_t done = 0; done < n;) {
      if (staged == kStage)
        flush();
      const uint32_t take = min(kStage - staged, n - done);
      if (cand && rank >= done && rank < done + take)
        stage[staged + (rank - done)] = entry;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      staged += take;
      done += take;
    }
  };
  bool big = false; // a valid pixel of this lane at or above kFast

  // Columns of this lane outside the frame (masked in every row), those it
  // outputs (inside the strip's middle and the frame), and those whose 5 x 5
  // erode reaches an output through stage 2's 11 x 11 window.
  uint32_t outside = 0, in_out = 0, reach = 0;
#pragma unroll
  for (int q = 0; q < KQ; q++) {
    const int c = c0 + q;
    if (c < 0 || c >= width)
      outside |= 1u << q;
    else if (c >= x0 && c < x0 + kOut)
      in_out |= 1u << q;
    if (c >= x0 - 5 && c < x0 + kOut + 5)
      reach |= 1u << q;
  }

  // The ring starts at zero: rows before the first add and remove nothing.
  for (int s = 0; s < kRing; s++)
#pragma unroll
    for (int q = 0; q < KQ; q++)
      ring[s][q][lane] = 0;

  uint32_t v0w[KQ] = {}; // stage 0 vertical: sum
  // Stage 0 vertical sums of squares: uint16 frames, the wrapped uint32 sum
  // (the reference's own accumulator); uint32 frames, the squares' low 26
  // bits (49 of them stay below 2^32) and, beside, their high parts, nonzero
  // only while a row with a pixel of 2^13 or more is in the window.
  uint32_t v0q[KQ] = {};
  uint32_t v0h[KQ] = {};
  uint32_t v2w[KQ] = {}; // stage 2 vertical: sum
  // Per four columns (columns 4h .. 4h + 3 of the lane):
  //   mh[h]  masked nibbles, row y_in - k at nibble k (k < 8); rows before
  //          the first count as masked (their ring values are zero);
  //   xh[h]  excluded from stage 2 (masked or in scr1) nibbles, row
  //          (y_in - 5) - k at nibble k (k < 12), the rows before the first
  //          excluded;
  //   hr[h]  horizontal-erode nibbles, row (y_in - 3) - k at nibble k (k < 5).
  uint32_t mh[P], hr[P];
  uint64_t xh[P];
#pragma unroll
  for (int h = 0; h < P; h++) {
    mh[h] = ~0u;
    xh[h] = ~0ull;
    hr[h] = 0;
  }
  // The wave-uniform row records, one word (kRecW .. kRecX above), all
  // shifted once at the start of a row.
  uint64_t recs = 0;

  // Masks that stay put.  While every lane's masked columns are the same in
  // the 7 rows of stage 0's window (no masked pixel at all being the common
  // case), each column's count is fixed: sm[q] = 7 x its window's unmasked
  // columns.  The plan holds them, and for the lattice columns 4h + 2 of
  // every lane their counts and roots in the real type.  The lattice
  // pre-test applies when every unmasked column whose erode reaches an
  // output has an unmasked lattice column within two of it (s_ok): a column
  // unmasked in the row tested is unmasked in every row of the erodes
  // through it, so each such erode window then holds an unmasked lattice
  // pixel in each of its rows, and when all of those test negative no such
  // window is all scr0, whatever the row's other pixels test.
  uint32_t s_pat = 0;   // the pattern planned for (per lane)
  uint32_t s_y = 0;     // the horizontal erode of the masked columns alone
  uint32_t sm[KQ];
  R lat_m[P], lat_mm1[P], lat_root[P];
  uint32_t s_ok = 0, s_yany = 0, s_all49 = 0; // wave-uniform
#pragma unroll
  for (int q = 0; q < KQ; q++)
    sm[q] = 0;
#pragma unroll
  for (int h = 0; h < P; h++) {
    lat_m[h] = R(0);
    lat_mm1[h] = R(-1);
    lat_root[h] = table[0];
  }
  auto plan = [&](uint32_t pat) {
    s_pat = pat;
    // bit j: column c0 - 3 + j masked
    const uint32_t x = ((from_left(pat) >> (KQ - 3)) & 7u) | (pat << 3) |
                       ((from_right(pat) & 7u) << (KQ + 3));
    bool not49 = false;
#pragma unroll
    for (int q = 0; q < KQ; q++) {
      sm[q] = 7u * (7u - static_cast<uint32_t>(__popc((x >> q) & 0x7Fu)));
      not49 |= sm[q] != 49u;
    }
    s_all49 = __builtin_amdgcn_readfirstlane(bit_any(not49) ^ 1u);
    uint32_t lat = 0; // unmasked lattice columns of this lane
#pragma unroll
    for (int h = 0; h < P; h++) {
      const uint32_t m = sm[4 * h + 2];
      lat_m[h] = static_cast<R>(m);
      lat_mm1[h] = static_cast<R>(m) - R(1);
      lat_root[h] = table[m];
      lat |= ((pat >> (4 * h + 2)) & 1u ? 0u : 1u) << (4 * h + 2);
    }
    // bit j: column c0 - 4 + j is an unmasked lattice column
    const uint32_t t = (from_left(lat) >> (KQ - 4)) | (lat << 4) |
                       (from_right(lat) << (KQ + 4));
    const uint32_t near = t | (t >> 1) | (t >> 2) | (t >> 3) | (t >> 4);
    const uint32_t need = ~pat & reach & kAll;
    s_ok = __builtin_amdgcn_readfirstlane(bit_any(((near >> 2) & need) != need) ^ 1u);
    const uint32_t e = ((from_left(pat) >> (KQ - 2)) & 3u) | (pat << 2) |
                       ((from_right(pat) & 3u) << (KQ + 2));
    s_y = e & (e >> 1) & (e >> 2) & (e >> 3) & (e >> 4) & kAll;
    s_yany = __builtin_amdgcn_readfirstlane(bit_any(s_y != 0u));
  };
  // The plan's state, one wave-uniform word: bits 0-3 count the rows still
  // to come before the masked columns have stayed put through a whole
  // stage-0 window (zero once they have), bit 4 says the plan is older than
  // them.  `irregular` says the previous row had masked pixels beyond the
  // frame's edges (its masked columns are then the history's newest
  // nibbles); the rows before the first count as all masked.
  uint32_t pstate = 6u | 16u;
  uint32_t irregular = 1;
  auto last_mbits = [&]() {
    uint32_t m = 0;
#pragma unroll
    for (int h = 0; h < P; h++)
      m |= (mh[h] & 0xFu) << (4 * h);
    return m;
  };

  const int steps = (y1 - y0) + 20;
  // Rows arrive kAhead ahead of use, in three buffers: the row read from one
  // while the row kAhead below loads into the one read the step before, so
  // no value is copied out of the way of a load.  Rows outside the frame
  // load a clamped row (their values are replaced below), and rows past the
  // segment's last reload that last row, from the cache.  Columns start at a
  // multiple of four, so a lane is wholly inside or wholly outside the
  // frame unless the width is not one (`straddle`).
  const int cl = min(max(c0, 0), width - KQ);
  const int sh = c0 - cl;
  const bool edge = __ballot(outside != 0) != 0;           // wave-uniform
  const bool lane_out = outside == kAll;
  const bool straddle = __ballot(outside != 0 && !lane_out) != 0;
  const int y_last = min(height, y1 + 10) - 1;
  auto issue = [&](int y, uint32_t(&v)[KQ]) {
    load_lane<T, KQ>(frame, width, max(0, min(y, y_last)), cl, v);
  };
  uint32_t ahead[kAhead + 1][KQ];
#pragma unroll
  for (int k = 0; k < kAhead; k++)
    issue(y0 - 10 + k, ahead[k]);
  int slot = 0; // ring slot of the entering row
  auto row = [&](int t, uint32_t(&val)[KQ], uint32_t(&next)[KQ]) {
    const int y_in = y0 - 10 + t;
    recs = (recs << 1) & kRecKeep;

    // ---- the entering row: values with masked as zero, masked bits ----
    uint32_t mbits = outside;
    uint32_t wide = 0, hot = 0; // wave-uniform
    uint32_t changed;           // wave-uniform: masked columns moved
    if (static_cast<uint32_t>(y_in) >= static_cast<uint32_t>(height)) {
#pragma unroll
      for (int q = 0; q < KQ; q++)
        val[q] = 0;
      mbits = kAll;
      changed = bit_any(mbits != (irregular ? last_mbits() : outside));
      irregular = bit_any(mbits != outside);
    } else {
      if (edge) {
        if (straddle) {
          shift_lane<KQ>(val, sh, outside);
        } else {
#pragma unroll
          for (int q = 0; q < KQ; q++)
            val[q] = lane_out ? 0u : val[q];
        }
      }
      uint32_t top = 0;
#pragma unroll
      for (int q = 0; q < KQ; q++)
        top = max(top, val[q]);
      if (__builtin_expect(__ballot(top >= plain) != 0, 0)) {
        // Masked pixels become zero; a valid pixel at or above kFast voids
        // this frame's fast sums (the caller re-runs it exactly) and is
        // clamped so that no sum spills into its count bits.
        uint32_t vmax = 0;
#pragma unroll
        for (int q = 0; q < KQ; q++) {
          const bool m = val[q] >= masked;
          mbits |= (m ? 1u : 0u) << q;
          big |= !m && val[q] >= kFast;
          val[q] = m ? 0u : min(val[q], kFast - 1u);
          vmax = max(vmax, val[q]);
        }
        if constexpr (kU32) {
          wide = bit_any(vmax >= kWide);
          hot = bit_any(vmax >= kHot);
          recs |= (static_cast<uint64_t>(wide) << kRecW) |
                  (static_cast<uint64_t>(hot) << kRecH);
        }
        changed = bit_any(mbits != (irregular ? last_mbits() : outside));
        irregular = bit_any(mbits != outside);
      } else {
        changed = irregular;
        irregular = 0;
      }
    }
    issue(y_in + kAhead, next);
    if (__builtin_expect((changed | pstate) != 0u, 0))
      pstate = changed ? (6u | 16u) : pstate - min(pstate & 15u, 1u);

    // Ring rows leaving stage 0 (y_in - 7), entering stage 2 (y_in - 5) and
    // leaving it (y_in - 16, whose slot the entering row takes once it has
    // been read), low 16 bits.  The changes to the sums are formed in the
    // block of the reads (the empty asm holds them there), so that the
    // 16-bit loads' zero extension folds into them, and the rare rows with a
    // pixel at or above 2^13 or 2^16 correct the sums afterwards rather than
    // redefine the changes, so that the common path copies nothing.
    const int s7 = (slot - 7) & (kRing - 1);
    const int s5 = (slot - 5) & (kRing - 1);
    const int s16 = slot;
    uint32_t out0[KQ], in2[KQ], out2[KQ];
    ring_row(s7, out0);
    ring_row(s5, in2);
    ring_row(s16, out2);
#pragma unroll
    for (int q = 0; q < KQ; q++) {
      uint32_t d0 = val[q] - out0[q];
      uint32_t p0 = val[q] + out0[q];
      uint32_t d2 = in2[q] - out2[q];
      __asm__ volatile("" : "+v"(d0), "+v"(p0), "+v"(d2));
      // ---- stage 0: vertical 7-row sums (rows y_in-6 .. y_in) ----
      // The change of the squares is one 24-bit product (the pixels are
      // below 2^18), exact modulo 2^32: for uint16 frames the reference's
      // wrapped sum, for uint32 frames the sum of the squares' low 26 bits
      // once the rare rows below take the high parts out.
      v0w[q] += d0;
      v0q[q] += static_cast<uint32_t>(
          __mul24(static_cast<int>(d0), static_cast<int>(p0)));
      // Stage 2's entering and leaving rows, as a change to its sums.
      v2w[q] += d2;
    }
#pragma unroll
    for (int q = 0; q < KQ; q++)
      ring[slot][q][lane] = static_cast<uint16_t>(val[q]);
#pragma unroll
    for (int h = 0; h < P; h++)
      mh[h] = (mh[h] << 4) | ((mbits >> (4 * h)) & 0xFu);
    if (kU32 && __builtin_expect((recs & (rec_bit(kRecW, 5) | rec_bit(kRecW, 16) |
                                           rec_bit(kRecH, 0) | rec_bit(kRecH, 7))) != 0,
                                 0)) {
      // Rows with a pixel at or above 2^16 (before row y_in is in, row
      // y_in - k is bit k of the records): the high ring's bits 16-17 of
      // stage 2's entering and leaving rows.
      if (recs & (rec_bit(kRecW, 5) | rec_bit(kRecW, 16))) {
        uint32_t hin[KQ] = {}, hout[KQ] = {};
        if (recs & rec_bit(kRecW, 5))
          add_high<KQ>(high[s5][lane], hin);
        if (recs & rec_bit(kRecW, 16))
          add_high<KQ>(high[s16][lane], hout);
#pragma unroll
        for (int q = 0; q < KQ; q++)
          v2w[q] += hin[q] - hout[q];
      }
      // A pixel of 2^13 or more entering or leaving stage 0: the leaving
      // row whole, and the squares' parts from bit 26 up moved from the
      // low sums to the high ones.
      if (recs & (rec_bit(kRecH, 0) | rec_bit(kRecH, 7))) {
        uint32_t o[KQ];
#pragma unroll
        for (int q = 0; q < KQ; q++)
          o[q] = out0[q];
        if (recs & rec_bit(kRecW, 7)) {
          add_high<KQ>(high[s7][lane], o);
#pragma unroll
          for (int q = 0; q < KQ; q++) {
            const uint32_t hi = o[q] - out0[q];
            v0w[q] -= hi;
            v0q[q] -= static_cast<uint32_t>(
                __mul24(static_cast<int>(hi), static_cast<int>(o[q] + out0[q])));
          }
        }
#pragma unroll
        for (int q = 0; q < KQ; q++) {
          const uint64_t pi = static_cast<uint64_t>(val[q] & (kFast - 1u)) *
                              (val[q] & (kFast - 1u));
          const uint64_t po = static_cast<uint64_t>(o[q] & (kFast - 1u)) *
                              (o[q] & (kFast - 1u));
          const uint32_t dh =
              static_cast<uint32_t>(pi >> 26) - static_cast<uint32_t>(po >> 26);
          v0h[q] += dh;
          v0q[q] -= dh << 26;
        }
      }
    }
    // ---- stage 0: 7 columns across lanes, test -> scr0 of row y_in-3 ----
    uint32_t scr0 = 0; // starts as the masked bits of row y_in-3
#pragma unroll
    for (int h = 0; h < P; h++)
      scr0 |= ((mh[h] >> 12) & 0xFu) << (4 * h);
    const uint32_t scr0_masked = scr0;
    // Rows above y0 - 7 are never read by an output's windows: their stage 0
    // is skipped (scr0 the masked pixels, no erode bits), which changes
    // nothing that reaches an output.  A pixel of 2^13 or more in rows
    // y_in-6 .. y_in (uint32 frames) brings the squares' high parts in.
    // The plan's countdown starts at six, so rows before t = 6 never take
    // the static path, and the other path skips them.
    bool quiet = true;
    {
      const uint64_t hwin = kU32 ? (recs & (0x7Full << kRecH)) : 0;
      if ((pstate & 15u) == 0u) {
        // Masks that stay put (or none): the plan's counts.
        if (__builtin_expect(pstate != 0u, 0)) {
          plan(mbits);
          pstate = 0;
        }
        uint32_t hw[KQ], hq[KQ], hh[KQ];
        lattice_sums<3>(v0w, hw, std::make_integer_sequence<int, P>{});
        lattice_sums<3>(v0q, hq, std::make_integer_sequence<int, P>{});
        uint32_t lattice = 0;
        if (__builtin_expect(hwin == 0, 1)) {
#pragma unroll
          for (int h = 0; h < P; h++) {
            const int c = 4 * h + 2;
            if (S0::test(lat_m[h], lat_mm1[h], hw[c], S0::squares(hq[c]),
                         lat_root[h]))
              lattice |= 1u << c;
          }
        } else {
          lattice_sums<3>(v0h, hh, std::make_integer_sequence<int, P>{});
#pragma unroll
          for (int h = 0; h < P; h++) {
            const int c = 4 * h + 2;
            if (S0::test(lat_m[h], lat_mm1[h], hw[c],
                         S0::squares(hq[c], hh[c]), lat_root[h]))
              lattice |= 1u << c;
          }
        }
        if (s_ok == 0u || __ballot((lattice & ~s_pat) != 0) != 0) {
          quiet = false;
          scr0 |= lattice;
          slide_all<3>(v0w, hw, std::make_integer_sequence<int, P>{});
          slide_all<3>(v0q, hq, std::make_integer_sequence<int, P>{});
          if (__builtin_expect(hwin == 0, 1)) {
            if (s_all49) {
              const R root49 = table[49];
#pragma unroll
              for (int q = 0; q < KQ; q++)
                if ((q & 3) != 2 &&
                    S0::test(R(49), R(48), hw[q], S0::squares(hq[q]), root49))
                  scr0 |= 1u << q;
            } else {
#pragma unroll
              for (int q = 0; q < KQ; q++) {
                const R m = static_cast<R>(sm[q]);
                if ((q & 3) != 2 &&
                    S0::test(m, m - R(1), hw[q], S0::squares(hq[q]), table[sm[q]]))
                  scr0 |= 1u << q;
              }
            }
          } else {
            slide_all<3>(v0h, hh, std::make_integer_sequence<int, P>{});
#pragma unroll
            for (int q = 0; q < KQ; q++) {
              const R m = static_cast<R>(sm[q]);
              if ((q & 3) != 2 && S0::test(m, m - R(1), hw[q],
                                           S0::squares(hq[q], hh[q]), table[sm[q]]))
                scr0 |= 1u << q;
            }
          }
        }
      } else if (t >= 6) {
        // Masks that move: each column's unmasked pixels in rows
        // y_in-6 .. y_in counted from the clear bits of the masked history,
        // carried in the top bits of the sums.
        quiet = false;
        uint32_t word[KQ], hw[KQ], hq[KQ], hh[KQ];
#pragma unroll
        for (int q = 0; q < KQ; q++)
          word[q] = v0w[q] +
                    (static_cast<uint32_t>(__popc(
                         ~mh[q / 4] & static_cast<uint32_t>(column_bits(0, 7, q % 4))))
                     << 26);
        row_sums<3>(word, hw);
        row_sums<3>(v0q, hq);
        if (hwin)
          row_sums<3>(v0h, hh);
#pragma unroll
        for (int q = 0; q < KQ; q++) {
          const uint32_t m = hw[q] >> 26;
          const R mr = static_cast<R>(m);
          const auto s2 = hwin ? S0::squares(hq[q], hh[q]) : S0::squares(hq[q]);
          if (S0::test(mr, mr - R(1), hw[q] & (kCount0 - 1), s2, table[m]))
            scr0 |= 1u << q;
        }
      }
    }

    // ---- stage 1: erode 5 x 5 -> scr1 of row y_in-5 ----
    uint32_t x1 = 0, ex = 0; // scr1 and not masked; scr1 or masked
    {
      const uint32_t m3 = scr0_masked;
      uint32_t y = 0, ybit = 0, yubit = 0;
      if (quiet) {
        // (Before t = 6 no plan has been made in this sweep: s_y is empty.)
        y = s_y;
        ybit = s_yany;
      } else if (__ballot(scr0 != 0) != 0) {
        const uint32_t left = from_left(scr0);
        const uint32_t right = from_right(scr0);
        const uint32_t x = ((left >> (KQ - 2)) & 3u) | (scr0 << 2) |
                           ((right & 3u) << (KQ + 2));
        y = x & (x >> 1) & (x >> 2) & (x >> 3) & (x >> 4) & kAll;
        ybit = bit_any(y != 0);
        yubit = bit_any((y & ~m3) != 0);
      }
      recs |= (static_cast<uint64_t>(ybit) << kRecY) |
              (static_cast<uint64_t>(yubit) << kRecU);
#pragma unroll
      for (int h = 0; h < P; h++)
        hr[h] = (hr[h] << 4) | ((y >> (4 * h)) & 0xFu);
      if ((recs & (0x1Full << kRecY)) == (0x1Full << kRecY) &&
          (recs & rec_bit(kRecU, 2))) {
#pragma unroll
        for (int h = 0; h < P; h++) {
          const uint32_t all5 = hr[h] & (hr[h] >> 4) & (hr[h] >> 8) &
                                (hr[h] >> 12) & (hr[h] >> 16) & 0xFu;
          const uint32_t m5 = (mh[h] >> 20) & 0xFu;
          x1 |= (all5 & ~m5) << (4 * h);
          ex |= (all5 | m5) << (4 * h);
        }
        recs |= static_cast<uint64_t>(bit_any(x1 != 0)) << kRecX;
      } else {
#pragma unroll
        for (int h = 0; h < P; h++)
          ex |= ((mh[h] >> 20) & 0xFu) << (4 * h);
      }
    }

    // ---- stage 2: vertical 11-row sums of values outside scr1 ----
    {
      // The entering row's value less the leaving row's is in; the two
      // corrections follow, each only where the wave-uniform record says
      // its row holds an unmasked scr1 pixel: an entering one must not be
      // added, a leaving one was never added (masked pixels are zero
      // anyway).  Their rows are whole here (high ring bits added).
      if (__builtin_expect((recs & (rec_bit(kRecX, 0) | rec_bit(kRecX, 11))) != 0, 0)) {
        if (recs & rec_bit(kRecX, 0)) {
          uint32_t w[KQ];
#pragma unroll
          for (int q = 0; q < KQ; q++)
            w[q] = in2[q];
          if (recs & rec_bit(kRecW, 5))
            add_high<KQ>(high[s5][lane], w);
#pragma unroll
          for (int q = 0; q < KQ; q++)
            v2w[q] -= ((x1 >> q) & 1u) ? w[q] : 0u;
        }
        if (recs & rec_bit(kRecX, 11)) {
          uint32_t xout = 0; // row y_in-16's exclusion bits
#pragma unroll
          for (int h = 0; h < P; h++)
            xout |= (static_cast<uint32_t>(xh[h] >> 40) & 0xFu) << (4 * h);
          uint32_t w[KQ];
#pragma unroll
          for (int q = 0; q < KQ; q++)
            w[q] = out2[q];
          if (recs & rec_bit(kRecW, 16))
            add_high<KQ>(high[s16][lane], w);
#pragma unroll
          for (int q = 0; q < KQ; q++)
            v2w[q] += ((xout >> q) & 1u) ? w[q] : 0u;
        }
      }
#pragma unroll
      for (int h = 0; h < P; h++)
        xh[h] = (xh[h] << 4) | ((ex >> (4 * h)) & 0xFu);
    }

    // ---- stage 2: candidates of row y_in-10 ----
    if (__builtin_expect((recs & rec_bit(kRecX, 5)) != 0, 0) && t >= 20) {
      const int y2 = y_in - 10;
      // Excluded pixels of row y_in-10 in the output columns; of them the
      // candidates are the unmasked ones with a value above zero, and a
      // masked pixel's ring value is zero.
      uint32_t maybe = 0;
#pragma unroll
      for (int h = 0; h < P; h++)
        maybe |= (static_cast<uint32_t>(xh[h] >> 20) & 0xFu) << (4 * h);
      const int s10 = (slot - 10) & (kRing - 1);
      uint32_t raw[KQ];
      ring_row(s10, raw);
      if (recs & rec_bit(kRecW, 10))
        add_high<KQ>(high[s10][lane], raw);
      uint32_t nz = 0;
#pragma unroll
      for (int q = 0; q < KQ; q++)
        nz |= min(raw[q], 1u) << q;
      const uint32_t cand = maybe & in_out & nz;
      if (__ballot(cand != 0) != 0) {
        // Counts: the pixels of rows y_in-15 .. y_in-5 not excluded, from
        // the clear bits of the exclusion history.
        uint32_t word[KQ];
#pragma unroll
        for (int q = 0; q < KQ; q++)
          word[q] = v2w[q] + (static_cast<uint32_t>(__popcll(
                                  ~xh[q / 4] & column_bits(0, 11, q % 4)))
                              << 25);
        uint32_t hsum[KQ];
        row_sums<5>(word, hsum);
        auto entry = [&](int q) {
          const int c = c0 + q;
          return make_uint4(
              static_cast<uint32_t>(y2) * a.width + static_cast<uint32_t>(c),
              raw[q] | (static_cast<uint32_t>(c & 31) << 18), hsum[q],
              static_cast<uint32_t>(y2) * a.wpr + static_cast<uint32_t>(c) / 32);
        };
        // One wave scan of the lanes' candidate counts places them all when
        // the stage has room; a row with more goes column by column through
        // put's chunks.
        const uint32_t k = static_cast<uint32_t>(__popc(cand));
        const uint32_t incl = wave_inclusive(k);
        const uint32_t total = last_lane(incl);
        if (staged + total <= kStage) {
          uint32_t at = staged + incl - k;
#pragma unroll
          for (int q = 0; q < KQ; q++)
            if ((cand >> q) & 1u)
              stage[at++] = entry(q);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          staged += total;
        } else {
#pragma unroll
          for (int q = 0; q < KQ; q++)
            put(((cand >> q) & 1u) != 0, entry(q));
        }
      }
    }
    // The entering row's high bits take the slot of row y_in - 16 only now,
    // after stage 2's correction for that row has read them.
    if (__builtin_expect(wide != 0u, 0))
      high[slot][lane] = static_cast<typename WaveLds<KQ>::High>(pack_high<KQ>(val));
    slot = (slot + 1) & (kRing - 1);
  };
  static_assert(kAhead == 2, "the buffers rotate over three steps");
  for (int t = 0; t < steps; t += 3) {
    row(t, ahead[0], ahead[2]);
    if (t + 1 < steps)
      row(t + 1, ahead[1], ahead[0]);
    if (t + 2 < steps)
      row(t + 2, ahead[2], ahead[1]);
  }
  flush();
  if (__ballot(big) != 0 && lane == 0)
    __hip_atomic_fetch_or(gmem(a.overflow), 1u, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
}

// Every wave takes items from its partition's counter (one atomic a wave per
// item after its first), over all the frames of a batch (item_base
// ascending), any mix of sizes and pixel types in one launch: items of
// unequal cost then finish within about one item of each other in each
// partition, and the partitions hold like shares of every frame.  The grid
// is the device's resident waves, so the launch is persistent, and every
// wave leaves when its partition runs past the items.
template <int Waves, int KQ>
__global__ __launch_bounds__(Waves *kLanes)
__attribute__((amdgpu_waves_per_eu(KQ == 4 ? 4 : 2))) void fused_kernel(
    const FusedParams *__restrict__ frames, uint32_t nframes,
    uint32_t total_items, const float *__restrict__ tf,
    const double *__restrict__ td, uint32_t *__restrict__ queue) {
  __shared__ WaveLds<KQ> lds[Waves];
  __shared__ float table_f[kTable];
  __shared__ double table_d[kTable];
  for (int i = threadIdx.x; i < kTable; i += blockDim.x) {
    table_f[i] = tf[i];
    table_d[i] = td[i];
  }
  __syncthreads();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const uint32_t lane = threadIdx.x % kLanes;
  // Items come in chunks of kChunk consecutive ones (neighbouring strips of
  // one band of rows, which share their halo columns through the cache),
  // dealt round robin over `parts` partitions: blocks b = x mod parts take
  // partition x's chunks from its own counter, in a line of its own.  The
  // device deals blocks round robin over its XCDs, so a partition is an
  // XCD: its waves share one L2, and no counter is hit by the whole device
  // at once (atomics on one address from every XCD serialise at the memory
  // side, about 15 ns each).  A wave's first item is its place among its
  // partition's waves, taken without an atomic; the counter hands out the
  // rest.  A partition's items rise with its local index, so the first one
  // past the end ends the wave.  Every partition has a block (b = x), and
  // every block runs until its partition is done, so no item is missed.
  const uint32_t parts = min(gridDim.x, static_cast<uint32_t>(sfi::kQueueParts));
  const uint32_t x = blockIdx.x % parts;
  const uint32_t part_waves = ((gridDim.x - x + parts - 1) / parts) * Waves;
  SF_GLOBAL uint32_t *counter = gmem(queue) + sfi::kQueueStride * x;
  uint32_t local = (blockIdx.x / parts) * Waves + wave;
  for (;;) {
    const uint32_t item =
        ((local / kChunk) * parts + x) * kChunk + local % kChunk;
    if (item >= total_items)
      break;
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
      sweep<uint16_t, KQ>(p, item - p.item_base, lane, lds[wave], table_f);
    else
      sweep<uint32_t, KQ>(p, item - p.item_base, lane, lds[wave], table_d);
    uint32_t next = 0;
    if (lane == 0)
      next = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    local = __builtin_amdgcn_readfirstlane(next) + part_waves;
  }
}

// Resident blocks of a kernel variant on one CU, from the occupancy the
// compiler's register and LDS use allow; computed once.
template <int Waves, int KQ> uint32_t resident_blocks() {
  static const uint32_t blocks = [] {
    int n = 0;
    sfi::check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &n, fused_kernel<Waves, KQ>, Waves * kLanes, 0),
               "querying the fused kernel's occupancy");
    return static_cast<uint32_t>(n > 0 ? n : 1);
  }();
  return blocks;
}

template <int Waves, int KQ>
void launch_waves(const FusedParams *frames, uint32_t nframes,
                  uint32_t total_items, const float *tf, const double *td,
                  uint32_t *queue, uint32_t cus, hipStream_t stream) {
  const uint32_t resident = cus * resident_blocks<Waves, KQ>();
  const uint32_t need = (total_items + Waves - 1) / Waves;
  const uint32_t blocks = need < resident ? need : resident;
  fused_kernel<Waves, KQ><<<blocks, Waves * kLanes, 0, stream>>>(
      frames, nframes, total_items, tf, td, queue);
}

template <int KQ>
void launch_shape(int waves, const FusedParams *frames, uint32_t nframes,
                  uint32_t total_items, const float *tf, const double *td,
                  uint32_t *queue, uint32_t cus, hipStream_t stream) {
  switch (waves) {
  case 1:
    launch_waves<1, KQ>(frames, nframes, total_items, tf, td, queue, cus, stream);
    break;
  case 4:
    launch_waves<4, KQ>(frames, nframes, total_items, tf, td, queue, cus, stream);
    break;
  default:
    launch_waves<2, KQ>(frames, nframes, total_items, tf, td, queue, cus, stream);
    break;
  }
}

} // namespace

namespace sfi {

uint32_t strips_of(uint32_t width, int lane_columns) {
  const uint32_t out = static_cast<uint32_t>(out_columns(lane_columns));
  return (width + kPhase + out - 1) / out;
}

uint32_t segment_rows(uint64_t strip_rows, const Tuning &t) {
  // Segments tall enough that the 20 halo rows cost little, short enough
  // that the launch still spreads over the whole device: about target_waves
  // items, each 40 to 256 rows.  (A lone 16M frame measured fastest at about
  // 40: shorter segments add halo rows faster than the extra waves hide the
  // row chain's latency.)
  const uint64_t target = static_cast<uint64_t>(t.target_waves);
  uint64_t seg = (strip_rows + target - 1) / target;
  seg = seg < 40 ? 40 : (seg > 256 ? 256 : seg);
  // A set length is a ceiling: a launch too small to fill the device with
  // segments that long (a lone frame, under a shape tuned for batches) takes
  // the shorter ones.
  if (t.seg_rows > 0 && static_cast<uint64_t>(t.seg_rows) < seg)
    seg = static_cast<uint64_t>(t.seg_rows);
  return static_cast<uint32_t>(seg);
}

uint32_t plan_fused(FusedPlan &p, uint32_t height, uint32_t width,
                    uint32_t seg_rows, int lane_columns) {
  p.strips = strips_of(width, lane_columns);
  p.seg_rows = seg_rows;
  p.segments = (height + seg_rows - 1) / seg_rows;
  p.lane_columns = lane_columns;
  // A frame narrower than one lane's columns is left to the exact kernels.
  if (fused_narrow(width, lane_columns))
    return 0;
  return p.strips * p.segments;
}

void launch_fused_batch(const FusedParams *frames_device, uint32_t nframes,
                        uint32_t total_items, const float *tf, const double *td,
                        uint32_t *queue, int lane_columns, int waves_per_block,
                        hipStream_t stream) {
  if (total_items == 0)
    return;
  int cus = 256;
  int device = 0;
  if (hipGetDevice(&device) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                device);
  const uint32_t n = static_cast<uint32_t>(cus);
  if (lane_columns == 8)
    launch_shape<8>(waves_per_block, frames_device, nframes, total_items, tf, td,
                    queue, n, stream);
  else
    launch_shape<4>(waves_per_block, frames_device, nframes, total_items, tf, td,
                    queue, n, stream);
  check(hipGetLastError(), "launching the fused threshold");
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
  p.totals = a.totals;
  p.bits = static_cast<uint32_t>(a.bits);
  p.item_base = item_base;
  return p;
}

} // namespace sfi
