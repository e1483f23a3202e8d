// Synthetic detector frames for the tests and the benchmark: a Poisson
// background with a radial solvent ring, planted 2-D Gaussian spots that may
// persist over up to three frames of a rotation sweep, module gaps masked at
// the top of the pixel range, and hot pixels.  Deterministic: every value is
// a function of (seed, frame, pixel) through a counter-based hash, so a frame
// can be regenerated anywhere, in any order, by any number of threads.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "spotfinder_synth.h"

namespace {

// SplitMix64's finaliser over a running key: a good 64-bit mix, cheap.
inline uint64_t mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t key(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  return mix(mix(mix(seed ^ 0x5F3759DFull) ^ a) ^ (b * 0x9E3779B97F4A7C15ull)) ^
         c;
}

inline double uniform(uint64_t h) {
  return static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0);
}

// Poisson by inversion of the CDF for small means, normal approximation
// (rounded, clipped at zero) above 64, which is where the tail weight of the
// difference no longer matters to a spot finder's synthetic input.
uint32_t poisson(double mean, uint64_t h) {
  if (mean <= 0.0)
    return 0;
  if (mean < 64.0) {
    const double u = uniform(h);
    double p = std::exp(-mean), cdf = p;
    uint32_t k = 0;
    while (u > cdf && k < 1000) {
      k++;
      p *= mean / k;
      cdf += p;
    }
    return k;
  }
  // Box-Muller from two 26-bit halves of a second hash.
  const uint64_t g = mix(h);
  const double u1 = (static_cast<double>(g >> 38) + 0.5) / 67108864.0;
  const double u2 = (static_cast<double>((g >> 12) & 0x3FFFFFF) + 0.5) / 67108864.0;
  const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  const double v = std::floor(mean + std::sqrt(mean) * z + 0.5);
  return v < 0.0 ? 0u : static_cast<uint32_t>(v);
}

struct Spot {
  double cx, cy, sigma, peak;
};

// Spots that start on frame s: count, then each one's position, width, peak,
// and how many frames it lasts (1..3 when persistence is on).
void spots_starting(const sf_synth_desc &d, int64_t s, std::vector<Spot> &out,
                    std::vector<int> &len) {
  const uint32_t n = d.persist ? (d.spots_per_frame + 1) / 2 : d.spots_per_frame;
  for (uint32_t j = 0; j < n; j++) {
    const uint64_t base = key(d.seed, 0x5A07ull, static_cast<uint64_t>(s), j);
    Spot sp;
    sp.cx = uniform(mix(base + 1)) * d.width;
    sp.cy = uniform(mix(base + 2)) * d.height;
    sp.sigma = d.sigma_min + (d.sigma_max - d.sigma_min) * uniform(mix(base + 3));
    sp.peak = d.peak_min + (d.peak_max - d.peak_min) * uniform(mix(base + 4));
    out.push_back(sp);
    len.push_back(d.persist ? 1 + static_cast<int>(mix(base + 5) % 3) : 1);
  }
}

// The spots lit on frame f, with the rocking-curve weight of f in each.
void spots_on(const sf_synth_desc &d, int64_t f, std::vector<Spot> &lit) {
  static const double weight[4][3] = {
      {0, 0, 0}, {1.0, 0, 0}, {0.7, 0.7, 0}, {0.4, 1.0, 0.4}};
  for (int64_t back = 0; back < (d.persist ? 3 : 1); back++) {
    std::vector<Spot> sp;
    std::vector<int> len;
    spots_starting(d, f - back, sp, len);
    for (size_t j = 0; j < sp.size(); j++) {
      if (back < len[j]) {
        Spot s = sp[j];
        s.peak *= weight[len[j]][back];
        lit.push_back(s);
      }
    }
  }
}

bool in_gap(const sf_synth_desc &d, uint32_t x, uint32_t y) {
  if (d.module_w && d.gap_w) {
    const uint32_t px = d.module_w + d.gap_w;
    if (x % px >= d.module_w)
      return true;
  }
  if (d.module_h && d.gap_h) {
    const uint32_t py = d.module_h + d.gap_h;
    if (y % py >= d.module_h)
      return true;
  }
  return false;
}

template <typename T>
void render(const sf_synth_desc &d, int64_t f, T *out, uint32_t ylo,
            uint32_t yhi, const std::vector<Spot> &lit,
            const std::vector<double> &ring_mean) {
  const T gap = static_cast<T>(~T(0));
  const double cx = 0.5 * d.width, cy = 0.5 * d.height;
  for (uint32_t y = ylo; y < yhi; y++) {
    for (uint32_t x = 0; x < d.width; x++) {
      const size_t k = static_cast<size_t>(y) * d.width + x;
      if (in_gap(d, x, y)) {
        out[k] = gap;
        continue;
      }
      const double dx = x + 0.5 - cx, dy = y + 0.5 - cy;
      const uint32_t r = static_cast<uint32_t>(std::sqrt(dx * dx + dy * dy));
      const double mean = ring_mean[std::min<size_t>(r, ring_mean.size() - 1)];
      const uint64_t v =
          poisson(mean, key(d.seed, static_cast<uint64_t>(f), k, 0xB6));
      const uint64_t top = static_cast<uint64_t>(gap) - 2;
      T value = static_cast<T>(v > top ? top : v);
      // Hot pixels are a property of the detector, the same on every frame:
      // 65535, which is a valid count at 32 bits and the masked sentinel at 16.
      if (uniform(key(d.seed, 0x4075ull, k, 0x0)) < d.hot_fraction)
        value = static_cast<T>(65535);
      out[k] = value;
    }
  }
  // Spots: Poisson counts added over a 4-sigma box, clipped below the
  // masked sentinels; pixels in gaps stay masked.
  for (size_t j = 0; j < lit.size(); j++) {
    const Spot &s = lit[j];
    const double reach = 4.0 * s.sigma + 1.0;
    const int64_t ya = std::max<int64_t>(ylo, static_cast<int64_t>(s.cy - reach));
    const int64_t yb = std::min<int64_t>(yhi, static_cast<int64_t>(s.cy + reach) + 1);
    const int64_t xa = std::max<int64_t>(0, static_cast<int64_t>(s.cx - reach));
    const int64_t xb = std::min<int64_t>(d.width, static_cast<int64_t>(s.cx + reach) + 1);
    const double inv = 1.0 / (2.0 * s.sigma * s.sigma);
    for (int64_t y = ya; y < yb; y++) {
      for (int64_t x = xa; x < xb; x++) {
        const size_t k = static_cast<size_t>(y) * d.width + static_cast<size_t>(x);
        if (in_gap(d, static_cast<uint32_t>(x), static_cast<uint32_t>(y)) ||
            static_cast<uint64_t>(out[k]) + 1 >= static_cast<uint64_t>(gap))
          continue;
        const double ex = x + 0.5 - s.cx, ey = y + 0.5 - s.cy;
        const double mean = s.peak * std::exp(-(ex * ex + ey * ey) * inv);
        if (mean < 0.05)
          continue;
        const uint64_t add =
            poisson(mean, key(d.seed ^ 0xC0FFEEull, static_cast<uint64_t>(f), k, j));
        const uint64_t v = static_cast<uint64_t>(out[k]) + add;
        const uint64_t top = static_cast<uint64_t>(gap) - 2;
        out[k] = static_cast<T>(v > top ? top : v);
      }
    }
  }
}

} // namespace

extern "C" {

void sf_synth_default(sf_synth_desc *d, uint32_t width, uint32_t height) {
  std::memset(d, 0, sizeof(*d));
  d->width = width;
  d->height = height;
  d->bits = 32;
  d->seed = 1;
  d->background = 0.8;
  d->ring_mean = 1.5;
  d->ring_radius = 600.0;
  d->ring_sigma = 40.0;
  d->spots_per_frame = 300;
  d->sigma_min = 0.6;
  d->sigma_max = 1.5;
  d->peak_min = 20.0;
  d->peak_max = 5000.0;
  d->hot_fraction = 1e-4;
  d->persist = 0;
  d->threads = 0;
}

void sf_synth_eiger(sf_synth_desc *d, int modules_x, int modules_y) {
  // EIGER2: 1028 x 512 modules with 12-column and 38-row gaps (4M, 16M).
  sf_synth_default(d, static_cast<uint32_t>(modules_x * 1028 + (modules_x - 1) * 12),
                   static_cast<uint32_t>(modules_y * 512 + (modules_y - 1) * 38));
  d->module_w = 1028;
  d->gap_w = 12;
  d->module_h = 512;
  d->gap_h = 38;
}

int sf_synth_frame(const sf_synth_desc *d, int64_t frame, void *out) {
  if (!d || !out || d->width == 0 || d->height == 0 ||
      (d->bits != 16 && d->bits != 32))
    return -1;
  std::vector<Spot> lit;
  spots_on(*d, frame, lit);
  // Background mean by integer radius from the centre.
  const double half_diag =
      std::sqrt(0.25 * d->width * d->width + 0.25 * d->height * d->height) + 2;
  std::vector<double> ring(static_cast<size_t>(half_diag) + 1);
  for (size_t r = 0; r < ring.size(); r++) {
    const double z = (static_cast<double>(r) - d->ring_radius) / d->ring_sigma;
    ring[r] = d->background + d->ring_mean * std::exp(-0.5 * z * z);
  }
  unsigned threads = d->threads ? d->threads : std::thread::hardware_concurrency();
  if (threads == 0)
    threads = 1;
  threads = std::min<unsigned>(threads, 16);
  threads = std::min<unsigned>(threads, d->height);
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; t++) {
    const uint32_t ylo = static_cast<uint32_t>(static_cast<uint64_t>(d->height) * t / threads);
    const uint32_t yhi =
        static_cast<uint32_t>(static_cast<uint64_t>(d->height) * (t + 1) / threads);
    pool.emplace_back([&, ylo, yhi] {
      if (d->bits == 16)
        render<uint16_t>(*d, frame, static_cast<uint16_t *>(out), ylo, yhi, lit, ring);
      else
        render<uint32_t>(*d, frame, static_cast<uint32_t *>(out), ylo, yhi, lit, ring);
    });
  }
  for (auto &t : pool)
    t.join();
  return 0;
}

} // extern "C"
