// The grouping's pixels on the device: each labelled frame's signal pixels
// regrouped by component in HBM, and the spot pixel list written from there,
// so the host stitch handles components only (sf_labeller in device mode).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <vector>

#include "engine.hh"
#include "sf_internal.hh"
#include "sf_labeller.hh"

namespace sfi {

class DevicePixels {
public:
  DevicePixels();
  ~DevicePixels();
  DevicePixels(const DevicePixels &) = delete;
  DevicePixels &operator=(const DevicePixels &) = delete;

  // Regroups the signal pixels of engine slots ids[0..count) (n[i] pixels,
  // ncomp[i] components each, labelled) by component, stably, into one pooled
  // buffer, on the engine's stream: frame i's pixels start at pixels[i], as
  // {index, value} pairs.  The buffer is `*owner` until release() has been
  // called once per frame.
  void stage(Engine &e, const int *ids, size_t count, const size_t *n,
             const size_t *ncomp, uint64_t *owner, const void **pixels);
  // Runs the labeller's gather tasks into the spot pixel list's records
  // [base, base + count), copied to `dest`; waits.
  void gather(const std::vector<sfl::GatherTask> &tasks, uint64_t base,
              size_t count, sf_pixel *dest);
  void release(uint64_t owner);
  // Device bytes the staged frames hold, and the pool's total.
  size_t held_bytes() const;
  size_t pool_bytes() const;

private:
  struct Block {
    std::unique_ptr<DeviceBuffer> buf;
    size_t frames = 0; // frames not yet released
  };
  hipStream_t stream_ = nullptr;
  std::map<uint64_t, Block> blocks_;
  std::vector<std::unique_ptr<DeviceBuffer>> spare_;
  uint64_t next_owner_ = 1;
  DeviceBuffer keys_, keys_sorted_, packed_, sort_temp_, descs_, tasks_, out_;
  Pinned host_descs_, host_tasks_, host_out_;
};

} // namespace sfi
