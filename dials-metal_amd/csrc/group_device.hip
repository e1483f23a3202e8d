// The grouping's pixels on the device.
//
// The host stitch (sfl::Labeller, after dials_spots::Labeller, src/spots/
// dials_spots.cc:57-360) joins components across frames and decides which
// spots close and in what order; that is a few thousand records a frame.  The
// pixels -- about a hundred times more -- stay in HBM:
//   * stage(): each labelled frame's ordered signal pixels are keyed by
//     (frame in the batch, component) and sorted stably by that key (a radix
//     sort), so one buffer holds every frame of the batch grouped by
//     component, each component ascending by index -- the layout the host
//     stitch's per-frame arenas have;
//   * gather(): each emitted node's pixels go to their place in the spot
//     pixel list; a spot's nodes on one frame form a group whose pixels are
//     merged by index, each pixel placed by its rank in its own node plus its
//     rank (a binary search) in each other node of the group -- the order
//     dials_spots.cc:286-327 builds with its sort by (frame, index).
// Buffers come from a pool and return to it when the stitch has released
// every frame in them.
#include "group_device.hh"

#include <hipcub/device/device_radix_sort.hpp>

#include <cstring>
#include <stdexcept>

#include "sf_device.hh"

namespace sfi {

namespace {

using sfd::gmem;

struct PackFrame {
  const sf_signal_pixel *list;
  const uint32_t *comp_of;
  uint64_t offset;
  uint32_t n;
  uint32_t key; // the frame's place in the batch, above the component bits
};

__global__ __launch_bounds__(256) void pack_pixels(
    const PackFrame *__restrict__ frames, uint32_t *__restrict__ keys,
    uint64_t *__restrict__ pixels) {
  const PackFrame f = frames[blockIdx.y];
  const SF_GLOBAL sf_signal_pixel *list = gmem(f.list);
  const SF_GLOBAL uint32_t *comp_of = gmem(f.comp_of);
  SF_GLOBAL uint32_t *k = gmem(keys) + f.offset;
  SF_GLOBAL uint64_t *p = gmem(pixels) + f.offset;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < f.n; i += stride) {
    k[i] = f.key | comp_of[i];
    p[i] = uint64_t(list[i].index) | (uint64_t(list[i].value) << 32);
  }
}

// One wave per task.  Pixels of different components of one frame have
// different indices, so the ranks within a group are distinct.
__global__ __launch_bounds__(256) void gather_spots(
    const sfl::GatherTask *__restrict__ tasks, uint32_t ntasks, uint64_t base,
    sf_pixel *__restrict__ out) {
  const uint32_t t =
      __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / 64);
  if (t >= ntasks)
    return;
  const sfl::GatherTask task = tasks[t];
  const SF_GLOBAL uint64_t *px =
      gmem(static_cast<const uint64_t *>(task.pixels));
  // sf_pixel records as three words: frame, index, value.
  SF_GLOBAL uint32_t *dst =
      reinterpret_cast<SF_GLOBAL uint32_t *>(gmem(out) + (task.out - base));
  for (uint32_t k = task.begin + (threadIdx.x & 63); k < task.end; k += 64) {
    const uint64_t v = px[k];
    const uint32_t index = static_cast<uint32_t>(v);
    uint64_t rank = k - task.begin;
    for (uint32_t o = task.group; o < task.group + task.count; o++) {
      if (o == t)
        continue;
      const sfl::GatherTask other = tasks[o];
      const SF_GLOBAL uint64_t *q =
          gmem(static_cast<const uint64_t *>(other.pixels));
      uint32_t lo = other.begin, hi = other.end;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (static_cast<uint32_t>(q[mid]) < index)
          lo = mid + 1;
        else
          hi = mid;
      }
      rank += lo - other.begin;
    }
    dst[3 * rank] = static_cast<uint32_t>(task.frame);
    dst[3 * rank + 1] = index;
    dst[3 * rank + 2] = static_cast<uint32_t>(v >> 32);
  }
}

int bits_for(size_t values) { // bits holding 0 .. values - 1
  int b = 0;
  while (values > (size_t(1) << b))
    b++;
  return b;
}

} // namespace

DevicePixels::DevicePixels() {
  check(hipStreamCreate(&stream_), "hipStreamCreate");
}

DevicePixels::~DevicePixels() {
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
  }
}

void DevicePixels::stage(Engine &e, const int *ids, size_t count,
                         const size_t *n, const size_t *ncomp, uint64_t *owner,
                         const void **pixels) {
  size_t total = 0, top = 1, most = 0;
  for (size_t i = 0; i < count; i++) {
    total += n[i];
    top = std::max(top, ncomp[i]);
    most = std::max(most, n[i]);
  }
  const int key_bits = bits_for(top) + bits_for(count);
  if (key_bits > 32 || total > 0x7fffffffu)
    throw std::runtime_error("dials_spots: a batch too large to regroup on "
                             "the device (" + std::to_string(total) +
                             " pixels)");
  Block block;
  const size_t bytes = total * sizeof(uint64_t);
  for (auto it = spare_.begin(); it != spare_.end(); ++it)
    if ((*it)->bytes >= bytes) {
      block.buf = std::move(*it);
      spare_.erase(it);
      break;
    }
  if (!block.buf && total) {
    spare_.clear(); // none was large enough
    block.buf = std::make_unique<DeviceBuffer>();
    block.buf->ensure(std::max<size_t>(bytes + bytes / 4, size_t(1) << 20),
                      stream_);
  }
  const hipStream_t s = e.stream();
  std::vector<size_t> offset(count);
  if (total) {
    keys_.ensure(total * sizeof(uint32_t), s);
    keys_sorted_.ensure(total * sizeof(uint32_t), s);
    packed_.ensure(bytes, s);
    // The engine's stream is waited for before the next stage(), so the
    // staging is free again by then.
    PackFrame *d = host_descs_.get<PackFrame>(count);
    size_t at = 0;
    for (size_t i = 0; i < count; i++) {
      const FrameSlot &slot = e.slot(ids[i]);
      offset[i] = at;
      d[i].list = static_cast<const sf_signal_pixel *>(slot.ordered.ptr);
      d[i].comp_of = static_cast<const uint32_t *>(slot.comp_of.ptr);
      d[i].offset = at;
      d[i].n = static_cast<uint32_t>(n[i]);
      d[i].key = count > 1 ? static_cast<uint32_t>(i) << bits_for(top) : 0;
      at += n[i];
    }
    descs_.ensure(count * sizeof(PackFrame), s);
    check(hipMemcpyAsync(descs_.ptr, d, count * sizeof(PackFrame),
                         hipMemcpyHostToDevice, s),
          "uploading the regrouping's frames");
    const unsigned gx = static_cast<unsigned>(
        std::min<size_t>((most + 255) / 256, 512));
    hipLaunchKernelGGL(pack_pixels, dim3(gx, static_cast<unsigned>(count)),
                       dim3(256), 0, s, static_cast<const PackFrame *>(descs_.ptr),
                       static_cast<uint32_t *>(keys_.ptr),
                       static_cast<uint64_t *>(packed_.ptr));
    check(hipGetLastError(), "launching pack_pixels");
    const int end_bit = std::max(1, key_bits);
    size_t temp = 0;
    check(hipcub::DeviceRadixSort::SortPairs(
              nullptr, temp, static_cast<const uint32_t *>(keys_.ptr),
              static_cast<uint32_t *>(keys_sorted_.ptr),
              static_cast<const uint64_t *>(packed_.ptr),
              static_cast<uint64_t *>(block.buf->ptr), static_cast<int>(total),
              0, end_bit, s),
          "sizing the regrouping's sort");
    sort_temp_.ensure(std::max<size_t>(temp, 256), s);
    check(hipcub::DeviceRadixSort::SortPairs(
              sort_temp_.ptr, temp, static_cast<const uint32_t *>(keys_.ptr),
              static_cast<uint32_t *>(keys_sorted_.ptr),
              static_cast<const uint64_t *>(packed_.ptr),
              static_cast<uint64_t *>(block.buf->ptr), static_cast<int>(total),
              0, end_bit, s),
          "regrouping the pixels by component");
  }
  for (size_t i = 0; i < count; i++)
    pixels[i] = total ? static_cast<const uint64_t *>(block.buf->ptr) + offset[i]
                      : nullptr;
  block.frames = count;
  *owner = next_owner_++;
  blocks_.emplace(*owner, std::move(block));
}

void DevicePixels::gather(const std::vector<sfl::GatherTask> &tasks,
                          uint64_t base, size_t count, sf_pixel *dest) {
  const size_t ntasks = tasks.size();
  if (ntasks == 0 || count == 0)
    return;
  if (ntasks > 0xFFFFFFFFu)
    throw std::runtime_error("dials_spots: too many nodes to gather at once");
  // Every write lands in [base, base + count): checked here, since the
  // kernel trusts the tasks.
  for (size_t t = 0; t < ntasks; t++) {
    const sfl::GatherTask &task = tasks[t];
    if (task.end <= task.begin || !task.pixels || task.count == 0 ||
        task.group > t || task.group + task.count > ntasks || task.out < base)
      throw std::runtime_error("dials_spots: a malformed gather task");
    if (task.group == t) {
      uint64_t held = 0;
      for (uint32_t o = task.group; o < task.group + task.count; o++) {
        if (tasks[o].out != task.out || tasks[o].frame != task.frame)
          throw std::runtime_error("dials_spots: a malformed gather group");
        held += tasks[o].end - tasks[o].begin;
      }
      if (task.out - base + held > count)
        throw std::runtime_error("dials_spots: a gather past the spot list");
    }
  }
  sfl::GatherTask *h = host_tasks_.get<sfl::GatherTask>(ntasks);
  std::memcpy(h, tasks.data(), ntasks * sizeof(sfl::GatherTask));
  tasks_.ensure(ntasks * sizeof(sfl::GatherTask), stream_);
  out_.ensure(count * sizeof(sf_pixel), stream_);
  check(hipMemcpyAsync(tasks_.ptr, h, ntasks * sizeof(sfl::GatherTask),
                       hipMemcpyHostToDevice, stream_),
        "uploading the gather tasks");
  hipLaunchKernelGGL(gather_spots, dim3(static_cast<unsigned>((ntasks + 3) / 4)),
                     dim3(256), 0, stream_,
                     static_cast<const sfl::GatherTask *>(tasks_.ptr),
                     static_cast<uint32_t>(ntasks), base,
                     static_cast<sf_pixel *>(out_.ptr));
  check(hipGetLastError(), "launching gather_spots");
  // Through pinned memory, which the copy engines write at the link's rate.
  sf_pixel *staged = host_out_.get<sf_pixel>(count);
  check(hipMemcpyAsync(staged, out_.ptr, count * sizeof(sf_pixel),
                       hipMemcpyDeviceToHost, stream_),
        "copying the spot pixels back");
  check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  std::memcpy(dest, staged, count * sizeof(sf_pixel));
}

void DevicePixels::release(uint64_t owner) {
  auto it = blocks_.find(owner);
  if (it == blocks_.end())
    throw std::runtime_error("dials_spots: releasing an unknown frame buffer");
  if (--it->second.frames == 0) {
    if (it->second.buf)
      spare_.push_back(std::move(it->second.buf));
    blocks_.erase(it);
  }
}

size_t DevicePixels::held_bytes() const {
  size_t b = 0;
  for (const auto &e : blocks_)
    b += e.second.buf ? e.second.buf->bytes : 0;
  return b;
}

size_t DevicePixels::pool_bytes() const {
  size_t b = held_bytes();
  for (const auto &s : spare_)
    b += s->bytes;
  return b;
}

} // namespace sfi
