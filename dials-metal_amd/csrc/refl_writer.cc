// A DIALS reflection table (.refl) of the spots the path found, written as
// src/spots/refl.cc:223-351 writes it: one msgpack document
//
//   [ "dials::af::reflection_table", 2,
//     { "identifiers": { id: identifier } or {},
//       "nrows": N,
//       "data": { name: [ type, [ N, bin(column) ] ], ... } } ]
//
// with the columns in name order, each a little-endian memory image of N
// elements: bbox (int6), flags (std::size_t, Strong), id (int),
// intensity.sum.value and .variance (double), n_signal (int), panel
// (std::size_t), optionally shoebox (Shoebox<>: panel, bbox, version 2, then
// float data, uint8 mask and float background dense over the box), and
// xyzobs.px.value and .variance (vec3<double>).  Every msgpack header takes
// the smallest form that holds its value, as the reference's does, so that
// the bytes agree.
//
// Host code: the table is a few bytes a spot beside the frames the device
// reads, and the caller's file system is on the host.

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "spotfinder_hip.h"

namespace sfi {
void set_last_error(const std::string &message); // the library's sf_last_error
}

namespace {

static_assert(sizeof(size_t) == 8, "std::size_t columns are eight bytes");
static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__,
              "columns are little-endian memory images");
static_assert(sizeof(sf_spot) == 104 && sizeof(sf_pixel) == 12,
              "the reference's Spot and Pixel layouts");

constexpr uint8_t kMaskValid = 1u << 0;      // dials/model/data/mask_code.h
constexpr uint8_t kMaskForeground = 1u << 2;
constexpr uint64_t kFlagStrong = 1u << 5;    // dials::af::Flags::Strong
constexpr uint8_t kShoeboxVersion = 2;       // a uint8 mask

// Where the encoded bytes go: the caller's memory, or a file through a 1 MB
// buffer.
class Sink {
public:
  virtual ~Sink() = default;
  virtual void put(const void *p, size_t n) = 0;
};

// The caller's buffer: bytes past its capacity are counted, not written.
class SpanSink final : public Sink {
public:
  SpanSink(uint8_t *out, size_t capacity) : out_(out), cap_(capacity) {}
  void put(const void *p, size_t n) override {
    if (size_ + n <= cap_)
      std::memcpy(out_ + size_, p, n);
    else if (size_ < cap_)
      std::memcpy(out_ + size_, p, cap_ - size_);
    size_ += n;
  }
  size_t size() const { return size_; }

private:
  uint8_t *out_;
  size_t cap_, size_ = 0;
};

class FileSink final : public Sink {
public:
  explicit FileSink(const char *path) : path_(path) {
    f_ = std::fopen(path, "wb");
    if (!f_)
      throw std::runtime_error(std::string("cannot write ") + path);
    buf_.reserve(kChunk);
  }
  ~FileSink() override {
    if (f_)
      std::fclose(f_);
  }
  void put(const void *p, size_t n) override {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    while (n) {
      const size_t take = std::min(n, kChunk - buf_.size());
      buf_.insert(buf_.end(), b, b + take);
      b += take;
      n -= take;
      if (buf_.size() == kChunk)
        drain();
    }
  }
  void finish() {
    drain();
    const int rc = std::fclose(f_);
    f_ = nullptr;
    if (rc != 0)
      throw std::runtime_error("cannot finish writing " + path_);
  }

private:
  static constexpr size_t kChunk = 1u << 20;
  void drain() {
    if (!buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size())
      throw std::runtime_error("short write to " + path_);
    buf_.clear();
  }
  std::string path_;
  std::FILE *f_ = nullptr;
  std::vector<uint8_t> buf_;
};

// msgpack: headers big-endian, each in the smallest form for its value.
class Encoder {
public:
  explicit Encoder(Sink &s) : s_(s) {}

  template <typename T> void raw(T v) { s_.put(&v, sizeof v); }
  void bytes(const void *p, size_t n) { s_.put(p, n); }

  void uint(uint64_t v) {
    if (v < 0x80)
      tag(static_cast<uint8_t>(v));
    else
      sized(v, 0xcc, 0xcd, 0xce, 0xcf);
  }
  void str(const std::string &t) {
    const size_t n = t.size();
    if (n < 32)
      tag(static_cast<uint8_t>(0xa0 | n));
    else
      sized(n, 0xd9, 0xda, 0xdb, 0);
    s_.put(t.data(), n);
  }
  void array(size_t n) { n < 16 ? tag(static_cast<uint8_t>(0x90 | n)) : sized(n, 0, 0xdc, 0xdd, 0); }
  void map(size_t n) { n < 16 ? tag(static_cast<uint8_t>(0x80 | n)) : sized(n, 0, 0xde, 0xdf, 0); }
  void bin(uint64_t n) {
    if (n > 0xffffffffull)
      throw std::runtime_error(
          "a reflection table column would be over 4 GB, which msgpack's binary "
          "type cannot describe; write without shoeboxes");
    sized(n, 0xc4, 0xc5, 0xc6, 0);
  }
  // [ type, [ rows, bin(bytes) ] ] under its name; `bytes` follow.
  void column(const char *name, const char *type, size_t rows, uint64_t bytes) {
    str(name);
    array(2);
    str(type);
    array(2);
    uint(rows);
    bin(bytes);
  }

private:
  void tag(uint8_t t) { s_.put(&t, 1); }
  // The first of the 1-, 2-, 4- and 8-byte forms (a 0 tag: no such form)
  // that holds v, as its tag and v big-endian.
  void sized(uint64_t v, uint8_t t1, uint8_t t2, uint8_t t4, uint8_t t8) {
    const struct {
      uint8_t tag;
      int width;
      uint64_t top;
    } forms[4] = {{t1, 1, 0xffull}, {t2, 2, 0xffffull}, {t4, 4, 0xffffffffull},
                  {t8, 8, ~0ull}};
    for (const auto &f : forms) {
      if (f.tag == 0 || v > f.top)
        continue;
      uint8_t out[9];
      out[0] = f.tag;
      for (int i = 0; i < f.width; i++)
        out[1 + i] = static_cast<uint8_t>(v >> (8 * (f.width - 1 - i)));
      s_.put(out, 1 + static_cast<size_t>(f.width));
      return;
    }
    throw std::runtime_error("a msgpack length beyond its widest form");
  }
  Sink &s_;
};

// A fixed-width column's rows, `bytes_per_row` each from fill(i, dst), staged
// through a local buffer so that the sink sees a few large writes.
template <typename Fill>
void column_rows(Encoder &e, size_t n, size_t bytes_per_row, Fill fill) {
  constexpr size_t kRows = 4096;
  uint8_t stage[kRows * 24];
  for (size_t i = 0; i < n;) {
    const size_t take = std::min(kRows, n - i);
    for (size_t k = 0; k < take; k++)
      fill(i + k, stage + k * bytes_per_row);
    e.bytes(stage, take * bytes_per_row);
    i += take;
  }
}

uint64_t box_elements(const sf_spot &s) {
  return static_cast<uint64_t>(s.bbox[1] - s.bbox[0]) *
         static_cast<uint64_t>(s.bbox[3] - s.bbox[2]) *
         static_cast<uint64_t>(s.bbox[5] - s.bbox[4]);
}

void write_table(Sink &sink, const sf_spot *spots, size_t n, const sf_pixel *pixels,
                 size_t npixels, size_t width, const sf_refl_options &o) {
  if (width == 0)
    throw std::runtime_error("refl: a frame width of zero");
  for (size_t i = 0; i < n; i++) {
    const sf_spot &s = spots[i];
    if (s.bbox[1] < s.bbox[0] || s.bbox[3] < s.bbox[2] || s.bbox[5] < s.bbox[4])
      throw std::runtime_error("refl: a spot's bounding box is inverted");
    if (o.shoeboxes && (s.first > npixels || s.n_signal > npixels - s.first))
      throw std::runtime_error("refl: a spot's pixels run past the pixel list");
  }
  Encoder e(sink);
  e.array(3);
  e.str("dials::af::reflection_table");
  e.uint(2);
  e.map(3);
  e.str("identifiers");
  const std::string identifier = o.identifier ? o.identifier : "";
  if (identifier.empty()) {
    e.map(0);
  } else {
    e.map(1);
    e.uint(static_cast<uint64_t>(static_cast<int64_t>(o.id)));
    e.str(identifier);
  }
  e.str("nrows");
  e.uint(n);
  e.str("data");
  e.map(o.shoeboxes ? 10 : 9);

  auto put = [](uint8_t *dst, const void *src, size_t k) { std::memcpy(dst, src, k); };
  e.column("bbox", "int6", n, 24ull * n);
  column_rows(e, n, 24, [&](size_t i, uint8_t *d) { put(d, spots[i].bbox, 24); });
  e.column("flags", "std::size_t", n, 8ull * n);
  column_rows(e, n, 8, [&](size_t, uint8_t *d) { put(d, &kFlagStrong, 8); });
  e.column("id", "int", n, 4ull * n);
  column_rows(e, n, 4, [&](size_t, uint8_t *d) { put(d, &o.id, 4); });
  e.column("intensity.sum.value", "double", n, 8ull * n);
  column_rows(e, n, 8, [&](size_t i, uint8_t *d) { put(d, &spots[i].intensity, 8); });
  e.column("intensity.sum.variance", "double", n, 8ull * n);
  column_rows(e, n, 8, [&](size_t i, uint8_t *d) {
    put(d, &spots[i].intensity_variance, 8);
  });
  e.column("n_signal", "int", n, 4ull * n);
  column_rows(e, n, 4, [&](size_t i, uint8_t *d) {
    const int32_t v = static_cast<int32_t>(spots[i].n_signal);
    put(d, &v, 4);
  });
  e.column("panel", "std::size_t", n, 8ull * n);
  column_rows(e, n, 8, [&](size_t, uint8_t *d) { put(d, &o.panel, 8); });

  if (o.shoeboxes) {
    uint64_t total = 0;
    for (size_t i = 0; i < n; i++) // panel, bbox, version, then 9 bytes a voxel
      total += 4 + 24 + 1 + 9 * box_elements(spots[i]);
    e.column("shoebox", "Shoebox<>", n, total);
    std::vector<float> data;
    std::vector<uint8_t> mask;
    for (size_t i = 0; i < n; i++) {
      const sf_spot &s = spots[i];
      const uint64_t nx = static_cast<uint64_t>(s.bbox[1] - s.bbox[0]);
      const uint64_t ny = static_cast<uint64_t>(s.bbox[3] - s.bbox[2]);
      const uint64_t voxels = box_elements(s);
      data.assign(voxels, 0.0f);
      mask.assign(voxels, 0);
      for (uint32_t k = 0; k < s.n_signal; k++) {
        const sf_pixel &p = pixels[s.first + k];
        const int64_t x = static_cast<int64_t>(p.index % width) - s.bbox[0];
        const int64_t y = static_cast<int64_t>(p.index / width) - s.bbox[2];
        const int64_t z = static_cast<int64_t>(p.frame) - s.bbox[4];
        if (x < 0 || y < 0 || z < 0 || static_cast<uint64_t>(x) >= nx ||
            static_cast<uint64_t>(y) >= ny ||
            z >= static_cast<int64_t>(s.bbox[5] - s.bbox[4]))
          throw std::runtime_error("refl: a spot's pixel lies outside its box");
        const uint64_t at = (static_cast<uint64_t>(z) * ny + static_cast<uint64_t>(y)) * nx +
                            static_cast<uint64_t>(x); // z slowest, x fastest
        data[at] = static_cast<float>(p.value);
        mask[at] = kMaskValid | kMaskForeground;
      }
      e.raw<uint32_t>(static_cast<uint32_t>(o.panel));
      e.bytes(s.bbox, 24);
      e.raw<uint8_t>(kShoeboxVersion);
      e.bytes(data.data(), voxels * sizeof(float));
      e.bytes(mask.data(), voxels);
      std::fill(data.begin(), data.end(), 0.0f); // the background: never filled
      e.bytes(data.data(), voxels * sizeof(float));
    }
  }

  e.column("xyzobs.px.value", "vec3<double>", n, 24ull * n);
  column_rows(e, n, 24, [&](size_t i, uint8_t *d) { put(d, spots[i].position, 24); });
  e.column("xyzobs.px.variance", "vec3<double>", n, 24ull * n);
  column_rows(e, n, 24, [&](size_t i, uint8_t *d) { put(d, spots[i].variance, 24); });
}

int failed(const std::exception &ex) {
  sfi::set_last_error(ex.what());
  return -1;
}

} // namespace

extern "C" {

void sf_refl_options_default(sf_refl_options *o) {
  if (o)
    *o = sf_refl_options{0, 0, 1, nullptr};
}

int sf_refl_write(const char *path, const sf_spot *spots, size_t nspots,
                  const sf_pixel *pixels, size_t npixels, size_t width,
                  const sf_refl_options *options) {
  if (!path || (nspots && !spots) || (npixels && !pixels)) {
    sfi::set_last_error("bad argument: path, spots or pixels");
    return -1;
  }
  try {
    sf_refl_options o;
    sf_refl_options_default(&o);
    if (options)
      o = *options;
    FileSink f(path);
    write_table(f, spots, nspots, pixels, npixels, width, o);
    f.finish();
    return 0;
  } catch (const std::exception &ex) {
    return failed(ex);
  }
}

int sf_refl_encode(const sf_spot *spots, size_t nspots, const sf_pixel *pixels,
                   size_t npixels, size_t width, const sf_refl_options *options,
                   uint8_t *out, size_t capacity, size_t *size) {
  if ((nspots && !spots) || (npixels && !pixels) || (capacity && !out)) {
    sfi::set_last_error("bad argument: spots, pixels or output");
    return -1;
  }
  try {
    sf_refl_options o;
    sf_refl_options_default(&o);
    if (options)
      o = *options;
    SpanSink m(out, capacity);
    write_table(m, spots, nspots, pixels, npixels, width, o);
    if (size)
      *size = m.size();
    return m.size() > capacity ? -2 : 0;
  } catch (const std::exception &ex) {
    return failed(ex);
  }
}

} // extern "C"
