// A C surface over the reference's own CPU spot-finding code, so the tests can
// load it with ctypes and pin the oracle against it.  This file is ours; the
// code it calls is compiled from the reference's sources where they lie
// (src/spots/dext.cc, signal_order.cc, dials_spots.cc, refl.cc) by oracle/Makefile into
// oracle/_ref/libspots_ref.so.  Test infrastructure only.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

#include "dext.hh"
#include "dials_spots.hh"
#include "refl.hh"
#include "signal_order.hh"
#include "signal_pixel.hh"

namespace {

void copy_message(const std::exception &e, char *err, std::size_t errlen) {
  if (err && errlen) {
    std::strncpy(err, e.what(), errlen - 1);
    err[errlen - 1] = 0;
  }
}

template <typename T>
int run_dext(const T *image, std::size_t h, std::size_t w, SignalPixel *out,
             std::size_t capacity, std::size_t *found) {
  std::vector<SignalPixel> signal;
  const int rc = dext<T>(image, signal, h, w);
  if (found)
    *found = signal.size();
  if (rc != 0)
    return rc;
  const std::size_t n = signal.size() < capacity ? signal.size() : capacity;
  if (n)
    std::memcpy(out, signal.data(), n * sizeof(SignalPixel));
  return signal.size() > capacity ? -2 : 0;
}

} // namespace

extern "C" {

int ref_dext_u16(const std::uint16_t *image, std::size_t h, std::size_t w,
                 SignalPixel *out, std::size_t capacity, std::size_t *found) {
  return run_dext<std::uint16_t>(image, h, w, out, capacity, found);
}

int ref_dext_u32(const std::uint32_t *image, std::size_t h, std::size_t w,
                 SignalPixel *out, std::size_t capacity, std::size_t *found) {
  return run_dext<std::uint32_t>(image, h, w, out, capacity, found);
}

void ref_sort_by_index(SignalPixel *pixels, std::size_t n) {
  std::vector<SignalPixel> v(pixels, pixels + n);
  signal_order::by_index(v);
  if (n)
    std::memcpy(pixels, v.data(), n * sizeof(SignalPixel));
}

void *ref_labeller_new(std::size_t h, std::size_t w, std::size_t min_size,
                       std::size_t max_size, double max_separation, int two_d,
                       char *err, std::size_t errlen) {
  dials_spots::Options o;
  o.min_spot_size = min_size;
  o.max_spot_size = max_size;
  o.max_separation = max_separation;
  o.two_d = two_d != 0;
  try {
    return new dials_spots::Labeller(h, w, o);
  } catch (const std::exception &e) {
    copy_message(e, err, errlen);
    return nullptr;
  }
}

void ref_labeller_free(void *l) {
  delete static_cast<dials_spots::Labeller *>(l);
}

int ref_labeller_add(void *l, std::int64_t frame, const SignalPixel *signal,
                     std::size_t n, char *err, std::size_t errlen) {
  try {
    std::vector<SignalPixel> v(signal, signal + n);
    static_cast<dials_spots::Labeller *>(l)->add(frame, v);
    return 0;
  } catch (const std::exception &e) {
    copy_message(e, err, errlen);
    return -1;
  }
}

void ref_labeller_finish(void *l) {
  static_cast<dials_spots::Labeller *>(l)->finish();
}

std::size_t ref_labeller_spot_count(void *l) {
  return static_cast<dials_spots::Labeller *>(l)->spots().size();
}

void ref_labeller_spots(void *l, dials_spots::Spot *out) {
  const auto &s = static_cast<dials_spots::Labeller *>(l)->spots();
  if (!s.empty())
    std::memcpy(out, s.data(), s.size() * sizeof(dials_spots::Spot));
}

std::size_t ref_labeller_pixel_count(void *l) {
  return static_cast<dials_spots::Labeller *>(l)->pixels().size();
}

void ref_labeller_pixels(void *l, dials_spots::Pixel *out) {
  const auto &p = static_cast<dials_spots::Labeller *>(l)->pixels();
  if (!p.empty())
    std::memcpy(out, p.data(), p.size() * sizeof(dials_spots::Pixel));
}

void ref_labeller_counts(void *l, dials_spots::Counts *out) {
  *out = static_cast<dials_spots::Labeller *>(l)->counts();
}

std::size_t ref_labeller_live(void *l) {
  return static_cast<dials_spots::Labeller *>(l)->live();
}

std::size_t ref_sizeof_spot() { return sizeof(dials_spots::Spot); }

// refl::write (refl.cc:223-351) of n spots and their pixels to `path`.
int ref_refl_write(const char *path, const dials_spots::Spot *spots,
                   std::size_t n, const dials_spots::Pixel *pixels,
                   std::size_t npixels, std::size_t width, std::size_t panel,
                   int id, const char *identifier, int shoeboxes, char *err,
                   std::size_t errlen) {
  try {
    const std::vector<dials_spots::Spot> s(spots, spots + n);
    const std::vector<dials_spots::Pixel> p(pixels, pixels + npixels);
    refl::Options o;
    o.panel = panel;
    o.id = id;
    o.identifier = identifier ? identifier : "";
    o.shoeboxes = shoeboxes != 0;
    refl::write(path, s, p, width, o);
    return 0;
  } catch (const std::exception &e) {
    copy_message(e, err, errlen);
    return -1;
  }
}

// The reference's threshold on N threads, one frame each at a time, with a
// per-thread scratch as mxi_find keeps one (find_spots.cc:404-405).
std::uint64_t ref_dext_u32_threads(const std::uint32_t *frames,
                                   std::size_t nframes, std::size_t total,
                                   std::size_t h, std::size_t w, int threads,
                                   double *seconds) {
  std::atomic<std::size_t> next{0};
  std::atomic<std::uint64_t> signal{0};
  if (threads < 1)
    threads = 1;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++) {
    pool.emplace_back([&] {
      dext_scratch<std::uint32_t> scratch;
      std::vector<SignalPixel> out;
      for (;;) {
        const std::size_t f = next.fetch_add(1);
        if (f >= total)
          break;
        dext<std::uint32_t>(frames + (f % nframes) * h * w, out, h, w, scratch);
        signal.fetch_add(out.size());
      }
    });
  }
  for (auto &t : pool)
    t.join();
  if (seconds)
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                             t0)
                   .count();
  return signal.load();
}

} // extern "C"
