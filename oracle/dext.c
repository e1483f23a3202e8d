/*
 * The oracle's extended dispersion threshold: a CPU restatement of
 * src/spots/dext.cc:8-225 (dext<T>) for the checker only.
 *
 * Semantics restated, stage by stage (reference line numbers in brackets):
 *   masked        = max(T) - 1; a pixel at or above it is masked   [dext.cc:40]
 *   stage 0       over the clamped 7 x 7 window of valid pixels, count m, sum
 *                 s and sum of squares s2 in the wide type W (u32 for u16
 *                 frames, u64 for u32 frames, both wrapping); if m >= 2 the
 *                 pixel is signal when, in R (float for u16, double for u32),
 *                 (m*s2 - s*s - s*(m-1)) > s*6*sqrt(2*(m-1));
 *                 scr0 = signal || masked                           [:53-112]
 *   stage 1       scr1 = masked ? 1 : scr0 && every in-frame scr0 of the
 *                 5 x 5 neighbourhood                               [:115-142]
 *   stage 2       over the clamped 11 x 11 window of pixels that are valid
 *                 and not scr1, count m and sum s; candidates are p > 0 and
 *                 scr1 (p = 0 for a masked pixel); mean = m >= 2 ? s/m : 0;
 *                 signal when p >= mean + 3*sqrt(mean)              [:144-222]
 *   emit          {index, p, (float)mean, (u16)m, 0}, ascending index.
 *
 * The window sums here are separable running sums rather than the reference's
 * summed-area tables: both are the same sums modulo 2^bits(W), which is all
 * the tests read.  The tests themselves are the reference's expressions in the
 * same operation order; this file must be compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* Clamped sliding-window sum along one line: out[j] = sum of in[j-r..j+r]
 * restricted to [0, n).  W arithmetic wraps, as the reference's does. */
#define DEFINE_SLIDE(NAME, W)                                                  \
  static void NAME(const W *in, W *out, int n, int r) {                       \
    W acc = 0;                                                                 \
    for (int j = 0; j < r && j < n; j++)                                       \
      acc += in[j];                                                            \
    for (int j = 0; j < n; j++) {                                              \
      if (j + r < n)                                                           \
        acc += in[j + r];                                                      \
      if (j - r - 1 >= 0)                                                      \
        acc -= in[j - r - 1];                                                  \
      out[j] = acc;                                                            \
    }                                                                          \
  }

DEFINE_SLIDE(slide_u32, uint32_t)
DEFINE_SLIDE(slide_u64, uint64_t)
DEFINE_SLIDE(slide_i32, int32_t)

/* 1 where every in-frame flag of the (2r+1)^2 neighbourhood is set. */
static void erode_all(const uint8_t *flag, uint8_t *all, int ny, int nx, int r) {
  int32_t *zeros_row = (int32_t *)malloc(sizeof(int32_t) * (size_t)nx);
  int32_t *zeros_col = (int32_t *)calloc((size_t)nx, sizeof(int32_t));
  int32_t *rowz = (int32_t *)malloc(sizeof(int32_t) * (size_t)ny * (size_t)nx);
  int32_t *line = (int32_t *)malloc(sizeof(int32_t) * (size_t)nx);
  for (int i = 0; i < ny; i++) {
    for (int j = 0; j < nx; j++)
      line[j] = flag[(size_t)i * nx + j] ? 0 : 1;
    slide_i32(line, zeros_row, nx, r);
    memcpy(rowz + (size_t)i * nx, zeros_row, sizeof(int32_t) * (size_t)nx);
  }
  /* vertical running sum of the horizontal zero counts */
  for (int i = 0; i < r && i < ny; i++)
    for (int j = 0; j < nx; j++)
      zeros_col[j] += rowz[(size_t)i * nx + j];
  for (int i = 0; i < ny; i++) {
    for (int j = 0; j < nx; j++) {
      if (i + r < ny)
        zeros_col[j] += rowz[(size_t)(i + r) * nx + j];
      if (i - r - 1 >= 0)
        zeros_col[j] -= rowz[(size_t)(i - r - 1) * nx + j];
      all[(size_t)i * nx + j] = zeros_col[j] == 0;
    }
  }
  free(zeros_row);
  free(zeros_col);
  free(rowz);
  free(line);
}

/*
 * T pixel type, W wide accumulator, R real type, SQRT its square root,
 * MASKED the first masked value.
 */
#define DEFINE_DEXT(NAME, T, W, R, SQRT, MASKED, SLIDE)                        \
  int NAME(const T *image, size_t height, size_t width,                       \
           oracle_signal_pixel *out, size_t capacity, size_t *found) {        \
    if (found)                                                                 \
      *found = 0;                                                              \
    if (image == NULL || height == 0 || width == 0)                            \
      return -1;                                                               \
    if (height > (size_t)INT32_MAX / width)                                    \
      return -1;                                                               \
    const size_t n = height * width;                                           \
    const int ny = (int)height, nx = (int)width;                               \
    const T masked = (T)(MASKED);                                              \
    W *cm = (W *)calloc((size_t)nx, sizeof(W));                                \
    W *cs = (W *)calloc((size_t)nx, sizeof(W));                                \
    W *cs2 = (W *)calloc((size_t)nx, sizeof(W));                               \
    W *hm = (W *)malloc(sizeof(W) * (size_t)nx);                               \
    W *hs = (W *)malloc(sizeof(W) * (size_t)nx);                               \
    W *hs2 = (W *)malloc(sizeof(W) * (size_t)nx);                              \
    uint8_t *scr0 = (uint8_t *)malloc(n);                                      \
    uint8_t *scr1 = (uint8_t *)malloc(n);                                      \
    uint8_t *all = (uint8_t *)malloc(n);                                       \
    /* stage 0: rows i-3..i+3 clamped, by running column sums */             \
    for (int i = 0; i < ny + 3; i++) {                                         \
      if (i < ny) {                                                            \
        for (int j = 0; j < nx; j++) {                                         \
          const T v = image[(size_t)i * nx + j];                               \
          const W valid = v < masked ? (W)1 : (W)0;                            \
          const W p = valid ? (W)v : (W)0;                                     \
          cm[j] += valid;                                                      \
          cs[j] += p;                                                          \
          cs2[j] += p * p;                                                     \
        }                                                                      \
      }                                                                        \
      const int leave = i - 7;                                                 \
      if (leave >= 0) {                                                        \
        for (int j = 0; j < nx; j++) {                                         \
          const T v = image[(size_t)leave * nx + j];                           \
          const W valid = v < masked ? (W)1 : (W)0;                            \
          const W p = valid ? (W)v : (W)0;                                     \
          cm[j] -= valid;                                                      \
          cs[j] -= p;                                                          \
          cs2[j] -= p * p;                                                     \
        }                                                                      \
      }                                                                        \
      const int row = i - 3; /* the row whose window is now complete */       \
      if (row < 0)                                                             \
        continue;                                                              \
      SLIDE(cm, hm, nx, 3);                                                    \
      SLIDE(cs, hs, nx, 3);                                                    \
      SLIDE(cs2, hs2, nx, 3);                                                  \
      for (int j = 0; j < nx; j++) {                                           \
        const size_t k = (size_t)row * nx + j;                                 \
        int signal = 0;                                                        \
        if (hm[j] >= 2) {                                                      \
          const R m = (R)hm[j];                                                \
          const R s = (R)hs[j];                                                \
          const R s2 = (R)hs2[j];                                              \
          signal = (m * s2 - s * s - s * (m - (R)1)) >                         \
                   (s * (R)6 * SQRT((R)2 * (m - (R)1)));                       \
        }                                                                      \
        scr0[k] = (signal || image[k] >= masked) ? 1 : 0;                      \
      }                                                                        \
    }                                                                          \
    /* stage 1: 5 x 5 erode; masked pixels are always set */                  \
    erode_all(scr0, all, ny, nx, 2);                                           \
    for (size_t k = 0; k < n; k++)                                             \
      scr1[k] = image[k] >= masked ? 1 : (scr0[k] && all[k]);                  \
    /* stage 2: 11 x 11 over valid and not scr1 */                            \
    memset(cm, 0, sizeof(W) * (size_t)nx);                                     \
    memset(cs, 0, sizeof(W) * (size_t)nx);                                     \
    size_t count = 0;                                                          \
    for (int i = 0; i < ny + 5; i++) {                                         \
      if (i < ny) {                                                            \
        for (int j = 0; j < nx; j++) {                                         \
          const size_t k = (size_t)i * nx + j;                                 \
          const W valid = (image[k] < masked && !scr1[k]) ? (W)1 : (W)0;       \
          cm[j] += valid;                                                      \
          cs[j] += valid ? (W)image[k] : (W)0;                                 \
        }                                                                      \
      }                                                                        \
      const int leave = i - 11;                                                \
      if (leave >= 0) {                                                        \
        for (int j = 0; j < nx; j++) {                                         \
          const size_t k = (size_t)leave * nx + j;                             \
          const W valid = (image[k] < masked && !scr1[k]) ? (W)1 : (W)0;       \
          cm[j] -= valid;                                                      \
          cs[j] -= valid ? (W)image[k] : (W)0;                                 \
        }                                                                      \
      }                                                                        \
      const int row = i - 5;                                                   \
      if (row < 0)                                                             \
        continue;                                                              \
      SLIDE(cm, hm, nx, 5);                                                    \
      SLIDE(cs, hs, nx, 5);                                                    \
      for (int j = 0; j < nx; j++) {                                           \
        const size_t k = (size_t)row * nx + j;                                 \
        const T p = image[k] >= masked ? (T)0 : image[k];                      \
        if (!(p > 0 && scr1[k]))                                               \
          continue;                                                            \
        const R mean = hm[j] >= 2 ? (R)hs[j] / (R)hm[j] : (R)0;                \
        if (!((R)p >= (mean + (R)3 * SQRT(mean))))                             \
          continue;                                                            \
        if (count < capacity) {                                                \
          oracle_signal_pixel *o = out + count;                                \
          o->index = (uint32_t)k;                                              \
          o->value = (uint32_t)p;                                              \
          o->background = (float)mean;                                         \
          o->population = (uint16_t)hm[j];                                     \
          o->reserved = 0;                                                     \
        }                                                                      \
        count++;                                                               \
      }                                                                        \
    }                                                                          \
    free(cm);                                                                  \
    free(cs);                                                                  \
    free(cs2);                                                                 \
    free(hm);                                                                  \
    free(hs);                                                                  \
    free(hs2);                                                                 \
    free(scr0);                                                                \
    free(scr1);                                                                \
    free(all);                                                                 \
    if (found)                                                                 \
      *found = count;                                                          \
    return count > capacity ? -2 : 0;                                          \
  }

DEFINE_DEXT(oracle_dext_u16, uint16_t, uint32_t, float, sqrtf, 0xFFFEu,
            slide_u32)
DEFINE_DEXT(oracle_dext_u32, uint32_t, uint64_t, double, sqrt, 0xFFFFFFFEu,
            slide_u64)

static int by_index(const void *a, const void *b) {
  const uint32_t x = ((const oracle_signal_pixel *)a)->index;
  const uint32_t y = ((const oracle_signal_pixel *)b)->index;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* src/spots/signal_order.cc:19-72 sorts ascending by index; the bucket pass
 * there is an optimisation of exactly this ordering. */
void oracle_sort_by_index(oracle_signal_pixel *pixels, size_t n) {
  if (n > 1)
    qsort(pixels, n, sizeof(*pixels), by_index);
}
