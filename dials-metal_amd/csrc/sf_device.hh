// Device-side definitions shared by every kernel of the spot-finding path:
// the pixel-type traits, the SignalPixel record, and the exact tests.
//
// The tests reproduce src/spots/dext.cc:99-110 (stage 0) and :193-207
// (stage 2) operation for operation: the window sums are integers in the wide
// type W, converted to the real type R (correctly rounded), and the
// expressions are evaluated in R with the reference's operation order.  All
// code is compiled with -ffp-contract=off, so no multiply-add is fused; HIP's
// double division and square root, and float ones under
// -fhip-fp32-correctly-rounded-divide-sqrt, are correctly rounded, as the
// CPU's are.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spotfinder_hip.h"

namespace sfd {

// accumulator_t / real_t of src/spots/dext.hh:18-38.
template <typename T> struct traits;
template <> struct traits<uint16_t> {
  using W = uint32_t;
  using R = float;
  static constexpr uint16_t masked = 0xFFFEu;
};
template <> struct traits<uint32_t> {
  using W = uint64_t;
  using R = double;
  static constexpr uint32_t masked = 0xFFFFFFFEu;
};

// Global-memory views of pointers that arrive through descriptors in memory,
// so that accesses through them compile to global rather than flat
// instructions: a flat access also counts against the LDS wait counter, which
// makes every wait for an LDS read wait for loads in flight too.
#define SF_GLOBAL __attribute__((address_space(1)))
template <typename P> __device__ __forceinline__ SF_GLOBAL P *gmem(P *p) {
  return (SF_GLOBAL P *)p;
}

__device__ __forceinline__ float root(float v) { return __fsqrt_rn(v); }
__device__ __forceinline__ double root(double v) { return __dsqrt_rn(v); }
__device__ __forceinline__ float quotient(float a, float b) {
  return __fdiv_rn(a, b);
}
__device__ __forceinline__ double quotient(double a, double b) {
  return __ddiv_rn(a, b);
}

// Stage 0's dispersion test on exact window sums (dext.cc:99-110).  `table`
// holds R(sqrt(R(2) * (R(m) - 1))) for m = 0..121, computed on the host by
// the same correctly rounded square root the reference calls.
template <typename W, typename R>
__device__ __forceinline__ bool dispersion(W m_sum, W i_sum, W i2_sum,
                                           const R *table) {
  if (m_sum < 2)
    return false;
  const R m = static_cast<R>(m_sum);
  const R s = static_cast<R>(i_sum);
  const R s2 = static_cast<R>(i2_sum);
  const R lhs = (m * s2 - s * s) - s * (m - R(1));
  const R rhs = (s * R(6)) * table[m_sum];
  return lhs > rhs;
}

// Stage 2's Poisson test at a candidate (dext.cc:193-207).  Returns the mean
// through *mean_out whether or not the pixel is signal.
template <typename W, typename R>
__device__ __forceinline__ bool poisson(uint32_t p, W m_sum, W i_sum,
                                        R *mean_out) {
  const R mean = m_sum >= 2 ? quotient(static_cast<R>(i_sum),
                                       static_cast<R>(m_sum))
                            : R(0);
  *mean_out = mean;
  return static_cast<R>(p) >= mean + R(3) * root(mean);
}

// Lane helpers for wave64.
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// v from a DPP source lane, 0 where the source is outside the row or the
// row is not in `RowMask` (the destination keeps the old value, 0).
template <int Ctrl, int RowMask>
__device__ __forceinline__ uint32_t dpp_or_zero(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      0, static_cast<int>(v), Ctrl, RowMask, 0xf, false));
}

// Inclusive sum of v over lanes 0 .. this lane of the wave (every lane
// active), all in VALU lane moves: shifts by 1, 2, 4 and 8 within each row of
// 16 lanes (row_shr), then lane 15 of each row added to the next row and lane
// 31 to the upper half (row_bcast:15, row_bcast:31).
__device__ __forceinline__ uint32_t wave_inclusive(uint32_t v) {
  v += dpp_or_zero<0x111, 0xf>(v); // row_shr:1
  v += dpp_or_zero<0x112, 0xf>(v); // row_shr:2
  v += dpp_or_zero<0x114, 0xf>(v); // row_shr:4
  v += dpp_or_zero<0x118, 0xf>(v); // row_shr:8
  v += dpp_or_zero<0x142, 0xa>(v); // row_bcast:15 into rows 1 and 3
  v += dpp_or_zero<0x143, 0xc>(v); // row_bcast:31 into rows 2 and 3
  return v;
}

// The value of lane 63.
__device__ __forceinline__ uint32_t last_lane(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// Appends this lane's record if `signal`, one atomic per wave.  Every lane of
// the wave must call it.  Writes past `capacity` are dropped but still
// counted, so the caller can report the overflow (dext_cuda.cu:294-297).
__device__ __forceinline__ void emit_wave(bool signal, const sf_signal_pixel &rec,
                                          sf_signal_pixel *out,
                                          uint32_t *counter,
                                          uint32_t capacity) {
  const uint64_t ballot = __ballot(signal);
  if (ballot == 0)
    return;
  const uint32_t lane = lane_id();
  const int leader = __ffsll(static_cast<unsigned long long>(ballot)) - 1;
  uint32_t base = 0;
  if (static_cast<int>(lane) == leader)
    base = atomicAdd(counter, static_cast<uint32_t>(__popcll(ballot)));
  base = __shfl(base, leader);
  if (signal) {
    const uint32_t at =
        base + static_cast<uint32_t>(__popcll(ballot & ((1ull << lane) - 1ull)));
    if (at < capacity)
      out[at] = rec;
  }
}

} // namespace sfd
