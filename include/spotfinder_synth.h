/*
 * Synthetic detector frames for tests and benchmarks.  No counterpart in the
 * reference, whose device tests use a 16-bit generator of their own
 * (tests/spots/synthetic_frame.hh) and whose end-to-end fixtures need h5py.
 *
 * A frame is a function of (seed, frame number): a Poisson background of mean
 * `background` plus a Gaussian ring of height `ring_mean` at `ring_radius`
 * (width `ring_sigma`) around the detector centre; `spots_per_frame` 2-D
 * Gaussian spots (sigma uniform in [sigma_min, sigma_max], peak uniform in
 * [peak_min, peak_max], position uniform), Poisson-sampled; with `persist`,
 * each spot lasts 1-3 consecutive frames, keeping about spots_per_frame lit;
 * module gaps (module_w/gap_w columns, module_h/gap_h rows) set to the
 * all-ones sentinel (masked); and a fixed `hot_fraction` of pixels at 65535.
 */
#ifndef SPOTFINDER_SYNTH_H
#define SPOTFINDER_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sf_synth_desc {
  uint32_t width, height; /* fast, slow */
  uint32_t bits;          /* 16 or 32 */
  uint32_t module_w, gap_w, module_h, gap_h; /* 0: no gaps */
  uint64_t seed;
  double background, ring_mean, ring_radius, ring_sigma;
  uint32_t spots_per_frame;
  double sigma_min, sigma_max, peak_min, peak_max;
  double hot_fraction;
  uint32_t persist; /* spots last 1-3 frames */
  uint32_t threads; /* 0: all cores, at most 16 */
} sf_synth_desc;

/* Defaults: 0.8 background, ring 1.5 at r = 600 (sigma 40), 300 spots of
 * sigma 0.6-1.5 and peak 20-5000, 0.01 % hot pixels, 32 bits, seed 1. */
void sf_synth_default(sf_synth_desc *d, uint32_t width, uint32_t height);
/* EIGER2 geometry: modules_x x modules_y modules of 1028 x 512 pixels with
 * 12-column and 38-row gaps; 2 x 4 is 2068 x 2162 (4M), 4 x 8 is 4148 x 4362
 * (16M). */
void sf_synth_eiger(sf_synth_desc *d, int modules_x, int modules_y);
/* Writes width * height pixels of `bits` into out; 0, or -1 for a bad desc. */
int sf_synth_frame(const sf_synth_desc *d, int64_t frame, void *out);

#ifdef __cplusplus
}
#endif

#endif
