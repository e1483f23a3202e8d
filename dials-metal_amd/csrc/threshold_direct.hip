// The direct-window form of the threshold: three kernels, one thread per
// pixel, every window summed straight from the frame.  It is the plain
// cross-check for the fused kernel, as the reference keeps its own direct
// kernels beside the tiled ones (dext_cuda.cu:56-100, 221-253, 404-443;
// docs/spots_notes.md:223-235).  Semantics: src/spots/dext.cc:53-222.
#include "sf_device.hh"
#include "sf_internal.hh"

namespace {

using namespace sfd;

constexpr int kBX = 64, kBY = 4; // one wave per block row: 64 columns

// Stage 0: 7 x 7 dispersion test -> scr0 = signal || masked.
template <typename T>
__global__ __launch_bounds__(kBX *kBY) void stage0_direct(
    const T *__restrict__ frame, int height, int width,
    uint8_t *__restrict__ scr0, const typename traits<T>::R *__restrict__ table) {
  using W = typename traits<T>::W;
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= width || y >= height)
    return;
  const T masked = traits<T>::masked;
  W m = 0, s = 0, s2 = 0;
  const int r0 = max(0, y - 3), r1 = min(height - 1, y + 3);
  const int c0 = max(0, x - 3), c1 = min(width - 1, x + 3);
  for (int r = r0; r <= r1; r++) {
    const T *row = frame + static_cast<size_t>(r) * width;
    for (int c = c0; c <= c1; c++) {
      const T v = row[c];
      if (v < masked) {
        const W p = static_cast<W>(v);
        m += 1;
        s += p;
        s2 += p * p;
      }
    }
  }
  const size_t k = static_cast<size_t>(y) * width + x;
  const bool signal = dispersion<W>(m, s, s2, table);
  scr0[k] = (signal || frame[k] >= masked) ? 1 : 0;
}

// Stage 1: 5 x 5 erode -> scr1 (masked pixels are always set).
template <typename T>
__global__ __launch_bounds__(kBX *kBY) void stage1_direct(
    const T *__restrict__ frame, int height, int width,
    const uint8_t *__restrict__ scr0, uint8_t *__restrict__ scr1) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= width || y >= height)
    return;
  const size_t k = static_cast<size_t>(y) * width + x;
  if (frame[k] >= traits<T>::masked) {
    scr1[k] = 1;
    return;
  }
  uint8_t keep = scr0[k];
  if (keep) {
    const int r0 = max(0, y - 2), r1 = min(height - 1, y + 2);
    const int c0 = max(0, x - 2), c1 = min(width - 1, x + 2);
    for (int r = r0; r <= r1 && keep; r++)
      for (int c = c0; c <= c1; c++)
        if (!scr0[static_cast<size_t>(r) * width + c]) {
          keep = 0;
          break;
        }
  }
  scr1[k] = keep;
}

// Stage 2: 11 x 11 Poisson test over valid && !scr1, emit, bitmask.
template <typename T>
__global__ __launch_bounds__(kBX *kBY) void stage2_direct(
    const T *__restrict__ frame, int height, int width,
    const uint8_t *__restrict__ scr1, sf_signal_pixel *__restrict__ out,
    uint32_t *__restrict__ counter, uint32_t capacity,
    uint32_t *__restrict__ mask, uint32_t wpr) {
  using W = typename traits<T>::W;
  using R = typename traits<T>::R;
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const bool inside = x < width && y < height;
  const T masked = traits<T>::masked;
  bool signal = false;
  sf_signal_pixel rec{};
  if (inside) {
    const size_t k = static_cast<size_t>(y) * width + x;
    const T v = frame[k];
    const uint32_t p = v >= masked ? 0u : static_cast<uint32_t>(v);
    if (p > 0 && scr1[k]) {
      W m = 0, s = 0;
      const int r0 = max(0, y - 5), r1 = min(height - 1, y + 5);
      const int c0 = max(0, x - 5), c1 = min(width - 1, x + 5);
      for (int r = r0; r <= r1; r++) {
        const size_t base = static_cast<size_t>(r) * width;
        for (int c = c0; c <= c1; c++) {
          const T u = frame[base + c];
          if (u < masked && !scr1[base + c]) {
            m += 1;
            s += static_cast<W>(u);
          }
        }
      }
      R mean;
      signal = poisson<W, R>(p, m, s, &mean);
      rec.index = static_cast<uint32_t>(k);
      rec.value = p;
      rec.background = static_cast<float>(mean);
      rec.population = static_cast<uint16_t>(m);
      rec.reserved = 0;
    }
  }
  emit_wave(signal, rec, out, counter, capacity);
  // The wave covers 64 columns of one row starting at a multiple of 64, so
  // its ballot is two whole words of the bitmask.
  const uint64_t bits = __ballot(signal);
  const uint32_t lane = lane_id();
  if (y < height && (lane == 0 || lane == 32)) {
    const int xb = blockIdx.x * kBX + static_cast<int>(lane);
    if (xb < width)
      mask[static_cast<size_t>(y) * wpr + xb / 32] =
          lane == 0 ? static_cast<uint32_t>(bits)
                    : static_cast<uint32_t>(bits >> 32);
  }
}

template <typename T>
void run(const sfi::FrameArgs &a, uint8_t *scr0, uint8_t *scr1,
         hipStream_t stream) {
  using R = typename traits<T>::R;
  const dim3 block(kBX, kBY);
  const dim3 grid((a.width + kBX - 1) / kBX, (a.height + kBY - 1) / kBY);
  const T *frame = static_cast<const T *>(a.frame);
  const int h = static_cast<int>(a.height), w = static_cast<int>(a.width);
  stage0_direct<T><<<grid, block, 0, stream>>>(frame, h, w, scr0,
                                               static_cast<const R *>(a.table));
  stage1_direct<T><<<grid, block, 0, stream>>>(frame, h, w, scr0, scr1);
  stage2_direct<T><<<grid, block, 0, stream>>>(frame, h, w, scr1, a.out,
                                               a.counter, a.capacity, a.mask,
                                               a.wpr);
  sfi::check(hipGetLastError(), "launching the direct threshold");
}

} // namespace

namespace sfi {

void launch_direct(const FrameArgs &a, uint8_t *scr0, uint8_t *scr1,
                   hipStream_t stream) {
  if (a.bits == 16)
    run<uint16_t>(a, scr0, scr1, stream);
  else
    run<uint32_t>(a, scr0, scr1, stream);
}

} // namespace sfi
