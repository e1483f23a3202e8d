/*
 * Oracle: the 16-bit narrowing of --gpu-force and mxi_max's frame scan.
 *
 *   oracle_narrow_to_16   narrow_to_16 (src/spots/find_spots.cc:162-182): every
 *                         pixel's low 16 bits (so the 32-bit markers 0xfffffffe
 *                         and 0xffffffff become 0xfffe and 0xffff); the largest
 *                         pixel that is not a marker; true when it is below
 *                         0xFFFD, the first value 16 bits reserve.
 *   oracle_scan_u16/_u32  scan<T> (apps/mxi_max.cc:45-68): the largest count
 *                         that is not a marker (max() - 1 a bad pixel, max() a
 *                         tile join), how many pixels are each marker, and how
 *                         many counts reach 0xFFFD.
 *
 * Neither source file compiles on its own here (find_spots.cc and mxi_max.cc
 * need HDF5 and the reader), so these are restatements; tests/test_narrow.py
 * pins them.  Test infrastructure only: the product never links them.
 */
#include <stdint.h>

#include "oracle.h"

int oracle_narrow_to_16(const uint32_t *in, size_t n, uint16_t *out,
                        uint32_t *offending) {
  uint32_t worst = 0;
  for (size_t i = 0; i < n; i++) {
    const uint32_t v = in[i];
    const uint32_t real = v >= 0xFFFFFFFEu ? 0u : v;
    if (real > worst)
      worst = real;
    out[i] = (uint16_t)v;
  }
  *offending = worst;
  return worst < 0xFFFDu;
}

#define SCAN(T, NAME)                                                          \
  void NAME(const T *px, size_t n, oracle_frame_scan *s) {                     \
    const T join = (T)~(T)0, dead = (T)(join - 1);                             \
    uint64_t best = 0, dead_n = 0, join_n = 0, above = 0;                      \
    for (size_t i = 0; i < n; i++) {                                           \
      const T v = px[i];                                                       \
      dead_n += v == dead;                                                     \
      join_n += v == join;                                                     \
      const uint64_t real = v >= dead ? 0u : (uint64_t)v;                      \
      above += real >= 0xFFFDu;                                                \
      if (real > best)                                                         \
        best = real;                                                           \
    }                                                                          \
    s->largest = best;                                                         \
    s->bad = dead_n;                                                           \
    s->joins = join_n;                                                         \
    s->over = above;                                                           \
  }

SCAN(uint16_t, oracle_scan_u16)
SCAN(uint32_t, oracle_scan_u32)
