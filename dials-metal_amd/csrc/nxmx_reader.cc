// The NXmx chunk reader behind include/spotfinder_nxmx.h: what the
// reference's series::nxmx does (src/spots/nxmx.cc), built as a table of
// blocks searched by frame number and a cache of opened source datasets, one
// entry per distinct (file, dataset), so that a frame costs one chunk-info
// query and one H5Dread_chunk.  Chunks come out byte for byte as HDF5 stored
// them; sf_nxmx_read_chunks lays a run of them out back to back with the
// decoder's descriptors, so that a batch of frames reaches the device in one
// sf_decoder_load.
//
// Kept from the reference, because callers and files depend on them: the
// dataset path /entry/data/data, the filter identifiers (32008 bitshuffle,
// 32004 lz4), one frame per chunk, a skipped filter (mask bit 0) meaning the
// chunk is stored raw, an unallocated chunk meaning a frame never written, a
// "." mapping meaning the master file itself, relative names resolved against
// the master's directory, an unlimited mapping clipped to the dataset's
// extent, and the checks' messages.
#include <hdf5.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "spotfinder_nxmx.h"

namespace {

constexpr const char *kImages = "/entry/data/data";
constexpr H5Z_filter_t kFilterBitshuffle = 32008;
constexpr H5Z_filter_t kFilterLz4 = 32004;

thread_local std::string t_error;

// HDF5 keeps process-wide state; every call into it is made under this lock,
// including the closes (recursive: a read holds it while a dataset opens).
std::recursive_mutex &hdf5_lock() {
  static std::recursive_mutex m;
  return m;
}
using Hold = std::lock_guard<std::recursive_mutex>;

struct hdf5_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

hid_t must(hid_t id, const std::string &what) {
  if (id < 0)
    throw hdf5_error("HDF5: " + what + " failed");
  return id;
}
void ok(herr_t status, const char *what) {
  if (status < 0)
    throw hdf5_error(std::string("HDF5: ") + what + " failed");
}

// An HDF5 identifier with the function that closes its kind.
class Id {
public:
  Id() = default;
  Id(hid_t id, herr_t (*close)(hid_t)) : id_(id), close_(close) {}
  Id(Id &&o) noexcept : id_(o.id_), close_(o.close_) { o.id_ = -1; }
  Id &operator=(Id &&o) noexcept {
    if (this != &o) {
      drop();
      id_ = o.id_;
      close_ = o.close_;
      o.id_ = -1;
    }
    return *this;
  }
  Id(const Id &) = delete;
  Id &operator=(const Id &) = delete;
  ~Id() { drop(); }
  hid_t get() const { return id_; }

private:
  void drop() {
    if (id_ >= 0 && close_) {
      const Hold hold(hdf5_lock());
      close_(id_);
    }
    id_ = -1;
  }
  hid_t id_ = -1;
  herr_t (*close_)(hid_t) = nullptr;
};

std::string mapping_name(hid_t dcpl, size_t i,
                         ssize_t (*get)(hid_t, size_t, char *, size_t)) {
  const ssize_t n = get(dcpl, i, nullptr, 0);
  if (n < 0)
    throw hdf5_error("HDF5: cannot size a virtual mapping name");
  std::vector<char> text(static_cast<size_t>(n) + 1, '\0');
  if (get(dcpl, i, text.data(), text.size()) < 0)
    throw hdf5_error("HDF5: cannot read a virtual mapping name");
  return std::string(text.data(), static_cast<size_t>(n));
}

// The HDF5 signature at the start of the file or after a user block (512
// bytes or any power of two beyond, up to 1 GiB).
bool has_hdf5_signature(const std::string &path) {
  static const unsigned char sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path.c_str(), "rb"), std::fclose);
  if (!f)
    return false;
  for (long at = 0; at <= (1L << 30); at = at ? 2 * at : 512) {
    unsigned char got[8];
    if (std::fseek(f.get(), at, SEEK_SET) != 0 || std::fread(got, 1, 8, f.get()) != 8)
      return false;
    if (std::memcmp(got, sig, 8) == 0)
      return true;
  }
  return false;
}

bool first_nonblank_is_brace(const std::string &path) {
  std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path.c_str(), "rb"), std::fclose);
  if (!f)
    return false;
  int c;
  while ((c = std::fgetc(f.get())) != EOF && std::isspace(c))
    ;
  return c == '{';
}

std::string directory_of(const std::string &path) {
  const size_t slash = path.find_last_of('/');
  return slash == std::string::npos ? std::string() : path.substr(0, slash);
}

struct Block {
  std::string file, dataset;
  uint64_t first = 0, last = 0, offset = 0;
  size_t source = 0; // index into sf_nxmx::sources
};

// One opened source dataset and what its creation properties say.
struct Source {
  std::string file, dataset;
  bool open = false;
  Id fid, did;
  uint64_t height = 0, width = 0;
  uint32_t bit_depth = 0;
  const char *algorithm = "";
};

} // namespace

struct sf_nxmx {
  std::string master;
  sf_nxmx_info info{};
  std::vector<Block> blocks; // ascending by first frame
  std::vector<Source> sources;

  const Block &block_of(uint64_t frame) const {
    // The last block starting at or before the frame.
    auto it = std::upper_bound(blocks.begin(), blocks.end(), frame,
                               [](uint64_t f, const Block &b) { return f < b.first; });
    if (it != blocks.begin()) {
      const Block &b = *(it - 1);
      if (frame >= b.first && frame <= b.last)
        return b;
    }
    throw std::invalid_argument("frame " + std::to_string(frame) +
                                " is outside every virtual mapping");
  }

  Source &source_of(const Block &b) {
    Source &s = sources[b.source];
    if (s.open)
      return s;
    const Hold hold(hdf5_lock());
    s.fid = Id(must(H5Fopen(s.file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT),
                    "opening " + s.file),
               H5Fclose);
    s.did = Id(must(H5Dopen2(s.fid.get(), s.dataset.c_str(), H5P_DEFAULT),
                    "opening " + s.dataset),
               H5Dclose);
    const Id dcpl(must(H5Dget_create_plist(s.did.get()), "H5Dget_create_plist"),
                  H5Pclose);
    if (H5Pget_layout(dcpl.get()) != H5D_CHUNKED)
      throw hdf5_error(s.file + " " + s.dataset +
                       " is not chunked, so it cannot be chunk-read");
    hsize_t chunk[3] = {0, 0, 0};
    if (H5Pget_chunk(dcpl.get(), 3, chunk) != 3)
      throw hdf5_error("expected a three-dimensional chunk in " + s.file);
    if (chunk[0] != 1)
      throw hdf5_error("chunks in " + s.file + " span " + std::to_string(chunk[0]) +
                       " frames; only one frame per chunk can be read this way");
    s.height = chunk[1];
    s.width = chunk[2];
    s.algorithm = "";
    const int nf = H5Pget_nfilters(dcpl.get());
    if (nf < 0)
      throw hdf5_error("HDF5: cannot count filters");
    for (int i = 0; i < nf; i++) {
      unsigned flags = 0;
      size_t nvalues = 0;
      const H5Z_filter_t id = H5Pget_filter2(dcpl.get(), static_cast<unsigned>(i),
                                             &flags, &nvalues, nullptr, 0, nullptr,
                                             nullptr);
      if (id == kFilterBitshuffle) {
        s.algorithm = "bslz4";
        break;
      }
      if (id == kFilterLz4) {
        s.algorithm = "lz4";
        break;
      }
    }
    const Id type(must(H5Dget_type(s.did.get()), "H5Dget_type"), H5Tclose);
    s.bit_depth = static_cast<uint32_t>(8 * H5Tget_size(type.get()));
    s.open = true;
    return s;
  }

  // The chunk of a frame: 0 and its description, or 1 when never written.
  int describe(uint64_t frame, sf_nxmx_frame *f, const Block **block, hsize_t *coord) {
    const Hold hold(hdf5_lock()); // also guards the source cache
    const Block &b = block_of(frame);
    Source &s = source_of(b);
    coord[0] = static_cast<hsize_t>(frame - b.first + b.offset);
    coord[1] = coord[2] = 0;
    unsigned mask = 0;
    haddr_t address = 0;
    hsize_t size = 0;
    ok(H5Dget_chunk_info_by_coord(s.did.get(), coord, &mask, &address, &size),
         "H5Dget_chunk_info_by_coord");
    *block = &b;
    f->number = frame;
    f->height = s.height;
    f->width = s.width;
    f->bit_depth = s.bit_depth;
    f->algorithm = (mask & 1u) ? "" : s.algorithm;
    f->size = static_cast<size_t>(size);
    return (address == HADDR_UNDEF || size == 0) ? 1 : 0;
  }

  void read_into(const Block &b, const hsize_t *coord, void *buf, sf_nxmx_frame *f) {
    const Hold hold(hdf5_lock());
    Source &s = source_of(b);
    uint32_t filters = 0;
    ok(H5Dread_chunk(s.did.get(), H5P_DEFAULT, coord, &filters, buf), "H5Dread_chunk");
    // Bit 0: the filter was skipped for this chunk, which is stored raw.
    f->algorithm = (filters & 1u) ? "" : s.algorithm;
  }
};

namespace {

template <typename F> int guarded(F &&body) {
  try {
    return body();
  } catch (const std::invalid_argument &e) {
    t_error = e.what();
    return -1;
  } catch (const std::bad_alloc &) {
    t_error = "out of host memory";
    return -3;
  } catch (const std::exception &e) {
    t_error = e.what();
    return -3;
  }
}

std::vector<Block> unpack_mapping(hid_t dcpl, uint64_t images, const std::string &master) {
  size_t count = 0;
  ok(H5Pget_virtual_count(dcpl, &count), "H5Pget_virtual_count");
  if (count == 0)
    throw hdf5_error(std::string(kImages) + " maps to nothing");
  const std::string dir = directory_of(master);
  std::vector<Block> out;
  for (size_t i = 0; i < count; i++) {
    Block b;
    const std::string name = mapping_name(dcpl, i, H5Pget_virtual_filename);
    b.dataset = mapping_name(dcpl, i, H5Pget_virtual_dsetname);
    if (name == ".")
      b.file = master; // the master itself, named as the caller named it
    else if (!name.empty() && name[0] != '/' && !dir.empty())
      b.file = dir + "/" + name;
    else
      b.file = name;
    const Id vspace(must(H5Pget_virtual_vspace(dcpl, i), "H5Pget_virtual_vspace"),
                    H5Sclose);
    const Id sspace(must(H5Pget_virtual_srcspace(dcpl, i), "H5Pget_virtual_srcspace"),
                    H5Sclose);
    hsize_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    ok(H5Sget_select_bounds(vspace.get(), lo, hi),
         "H5Sget_select_bounds on a virtual selection");
    b.first = lo[0];
    // An unlimited selection's bound is enormous: the extent decides.
    b.last = std::min<uint64_t>(hi[0], images ? images - 1 : 0);
    hsize_t slo[3] = {0, 0, 0}, shi[3] = {0, 0, 0};
    ok(H5Sget_select_bounds(sspace.get(), slo, shi),
         "H5Sget_select_bounds on a source selection");
    b.offset = slo[0];
    out.push_back(std::move(b));
  }
  std::sort(out.begin(), out.end(),
            [](const Block &x, const Block &y) { return x.first < y.first; });
  return out;
}

} // namespace

extern "C" {

const char *sf_nxmx_last_error(void) { return t_error.c_str(); }

int sf_nxmx_open(const char *master, sf_nxmx **out, sf_nxmx_info *info) {
  if (master == nullptr || out == nullptr) {
    t_error = "bad argument: null master name or result";
    return -1;
  }
  *out = nullptr;
  struct stat st;
  if (::stat(master, &st) != 0 || !S_ISREG(st.st_mode))
    return 1; // not there yet: a master file still being written may appear
  return guarded([&]() -> int {
    auto s = std::make_unique<sf_nxmx>();
    s->master = master;
    if (!has_hdf5_signature(s->master))
      throw hdf5_error(s->master + " is not an HDF5 file" +
                       (first_nonblank_is_brace(s->master)
                            ? std::string(": it looks like JSON, an experiment "
                                          "list perhaps, which is not the images")
                            : std::string()));
    const Hold hold(hdf5_lock());
    const Id file(must(H5Fopen(master, H5F_ACC_RDONLY, H5P_DEFAULT),
                       std::string("opening ") + master),
                  H5Fclose);
    const Id data(must(H5Dopen2(file.get(), kImages, H5P_DEFAULT),
                       std::string("opening ") + kImages),
                  H5Dclose);
    const Id space(must(H5Dget_space(data.get()), "H5Dget_space"), H5Sclose);
    if (H5Sget_simple_extent_ndims(space.get()) != 3)
      throw hdf5_error(std::string(kImages) + " is not a stack of images");
    hsize_t dims[3] = {0, 0, 0};
    ok(H5Sget_simple_extent_dims(space.get(), dims, nullptr),
         "H5Sget_simple_extent_dims");
    const Id dcpl(must(H5Dget_create_plist(data.get()), "H5Dget_create_plist"),
                  H5Pclose);
    if (H5Pget_layout(dcpl.get()) == H5D_VIRTUAL) {
      s->blocks = unpack_mapping(dcpl.get(), dims[0], s->master);
    } else {
      Block b;
      b.file = s->master;
      b.dataset = kImages;
      b.last = dims[0] ? dims[0] - 1 : 0;
      s->blocks.push_back(b);
    }
    // One source entry per distinct (file, dataset).
    std::map<std::pair<std::string, std::string>, size_t> seen;
    for (Block &b : s->blocks) {
      auto key = std::make_pair(b.file, b.dataset);
      auto it = seen.find(key);
      if (it == seen.end()) {
        it = seen.emplace(key, s->sources.size()).first;
        s->sources.emplace_back();
        s->sources.back().file = b.file;
        s->sources.back().dataset = b.dataset;
      }
      b.source = it->second;
    }
    s->info.images = dims[0];
    s->info.height = dims[1];
    s->info.width = dims[2];
    if (info)
      *info = s->info;
    *out = s.release();
    return 0;
  });
}

void sf_nxmx_close(sf_nxmx *s) { delete s; }

size_t sf_nxmx_block_count(const sf_nxmx *s) { return s ? s->blocks.size() : 0; }

int sf_nxmx_get_block(const sf_nxmx *s, size_t i, sf_nxmx_block *b) {
  if (s == nullptr || b == nullptr || i >= s->blocks.size()) {
    t_error = "bad argument: series, block index or result";
    return -1;
  }
  const Block &x = s->blocks[i];
  *b = sf_nxmx_block{x.file.c_str(), x.dataset.c_str(), x.first, x.last, x.offset};
  return 0;
}

int sf_nxmx_frame_info(sf_nxmx *s, uint64_t frame, sf_nxmx_frame *f) {
  if (s == nullptr || f == nullptr) {
    t_error = "bad argument: null series or result";
    return -1;
  }
  return guarded([&]() -> int {
    const Block *b = nullptr;
    hsize_t coord[3];
    return s->describe(frame, f, &b, coord);
  });
}

int sf_nxmx_read(sf_nxmx *s, uint64_t frame, void *buf, size_t capacity,
                 sf_nxmx_frame *f) {
  if (s == nullptr || f == nullptr || (capacity > 0 && buf == nullptr)) {
    t_error = "bad argument: null series, buffer or result";
    return -1;
  }
  return guarded([&]() -> int {
    const Block *b = nullptr;
    hsize_t coord[3];
    const int rc = s->describe(frame, f, &b, coord);
    if (rc != 0)
      return rc;
    if (f->size > capacity)
      return -2;
    s->read_into(*b, coord, buf, f);
    return 0;
  });
}

int sf_nxmx_read_chunks(sf_nxmx *s, uint64_t first, size_t n, uint8_t *buf,
                        size_t capacity, void *const *device_frames,
                        sf_chunk *chunks, size_t *loaded, size_t *used,
                        int *present) {
  if (loaded)
    *loaded = 0;
  if (used)
    *used = 0;
  if (s == nullptr || (n > 0 && (chunks == nullptr || device_frames == nullptr)) ||
      loaded == nullptr || used == nullptr || (capacity > 0 && buf == nullptr)) {
    t_error = "bad argument: null series, buffer, frames, descriptors or counts";
    return -1;
  }
  return guarded([&]() -> int {
    // Sizes first, so that a short buffer is refused before anything is read.
    std::vector<sf_nxmx_frame> f(n);
    std::vector<const Block *> b(n, nullptr);
    std::vector<hsize_t> coord(3 * n);
    std::vector<int> there(n, 0);
    size_t need = 0;
    for (size_t i = 0; i < n; i++) {
      there[i] = s->describe(first + i, &f[i], &b[i], &coord[3 * i]) == 0;
      if (there[i])
        need += f[i].size;
    }
    *used = need;
    if (need > capacity)
      return -2;
    size_t at = 0, k = 0;
    for (size_t i = 0; i < n; i++) {
      if (present)
        present[i] = there[i];
      if (!there[i])
        continue;
      s->read_into(*b[i], &coord[3 * i], buf + at, &f[i]);
      chunks[k] = sf_chunk{buf + at, f[i].size, f[i].algorithm, f[i].bit_depth,
                           static_cast<uint32_t>(f[i].height),
                           static_cast<uint32_t>(f[i].width), device_frames[i]};
      at += f[i].size;
      k++;
    }
    *loaded = k;
    return 0;
  });
}

} // extern "C"
