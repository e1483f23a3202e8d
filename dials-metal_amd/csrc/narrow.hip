// --gpu-force's narrowing of 32-bit frames to 16 bits (narrow_to_16,
// src/spots/find_spots.cc:162-182) and mxi_max's scan of a frame (scan<T>,
// apps/mxi_max.cc:45-68), on frames in HBM, any number of frames a launch.
//
// One pass over each frame: eight pixels a thread at a time (16-byte loads,
// and a 16-byte store of their low halves when narrowing), the largest count
// that is not a marker and the counts of each marker and of counts at or
// above 0xFFFD kept in registers, then reduced across the wave and added to
// the frame's totals with one atomic per wave.  Bound by HBM: 4 bytes read
// and 2 written a pixel when narrowing, 2 or 4 read when only scanning.

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "sf_device.hh"
#include "sf_internal.hh"

namespace {

using sfi::check;

constexpr uint32_t kThreads = 256;
constexpr uint32_t kPerThread = 8; // pixels a thread takes at a time
constexpr uint32_t kSixteenBitLimit = 0xFFFDu; // apps/mxi_max.cc:42

typedef uint32_t u32x4 __attribute__((ext_vector_type(4))); // 16-byte access

struct Job {
  uint64_t frame, out; // device addresses; out 0: scan only
  uint64_t pixels;
  uint32_t bits, pad;
};

struct Acc {
  uint32_t largest, pad;
  unsigned long long bad, joins, over;
};

struct Tally {
  uint32_t best = 0, dead = 0, join = 0, over = 0;
  template <typename T> __device__ void add(T v) {
    constexpr T kJoin = static_cast<T>(~T{0});
    constexpr T kDead = static_cast<T>(kJoin - 1);
    dead += v == kDead;
    join += v == kJoin;
    const uint32_t real = v >= kDead ? 0u : static_cast<uint32_t>(v);
    over += real >= kSixteenBitLimit;
    best = max(best, real);
  }
};

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o)));
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v += static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o));
  return v;
}

__global__ __launch_bounds__(kThreads) void narrow_scan(const Job *__restrict__ jobs,
                                                        Acc *acc) {
  const Job j = jobs[blockIdx.y];
  const uint64_t n = j.pixels;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kThreads * kPerThread;
  uint64_t i = (static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x) * kPerThread;
  Tally t;
  if (j.bits == 32) {
    const SF_GLOBAL uint32_t *in = sfd::gmem(reinterpret_cast<const uint32_t *>(j.frame));
    SF_GLOBAL uint16_t *out = sfd::gmem(reinterpret_cast<uint16_t *>(j.out));
    for (; i < n; i += stride) {
      if (i + kPerThread <= n) {
        const u32x4 a = *reinterpret_cast<const SF_GLOBAL u32x4 *>(in + i);
        const u32x4 b = *reinterpret_cast<const SF_GLOBAL u32x4 *>(in + i + 4);
        const uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; k++)
          t.add(v[k]);
        if (j.out) {
          u32x4 o;
          o.x = (v[0] & 0xFFFFu) | (v[1] << 16);
          o.y = (v[2] & 0xFFFFu) | (v[3] << 16);
          o.z = (v[4] & 0xFFFFu) | (v[5] << 16);
          o.w = (v[6] & 0xFFFFu) | (v[7] << 16);
          *reinterpret_cast<SF_GLOBAL u32x4 *>(out + i) = o;
        }
      } else {
        for (uint64_t k = i; k < n; k++) {
          const uint32_t v = in[k];
          t.add(v);
          if (j.out)
            out[k] = static_cast<uint16_t>(v);
        }
      }
    }
  } else {
    const SF_GLOBAL uint16_t *in = sfd::gmem(reinterpret_cast<const uint16_t *>(j.frame));
    for (; i < n; i += stride) {
      if (i + kPerThread <= n) {
        const u32x4 a = *reinterpret_cast<const SF_GLOBAL u32x4 *>(in + i);
        const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          t.add(static_cast<uint16_t>(w[k] & 0xFFFFu));
          t.add(static_cast<uint16_t>(w[k] >> 16));
        }
      } else {
        for (uint64_t k = i; k < n; k++)
          t.add(in[k]);
      }
    }
  }
  const uint32_t best = wave_max(t.best);
  const uint32_t dead = wave_sum(t.dead), join = wave_sum(t.join), over = wave_sum(t.over);
  if (threadIdx.x % 64 == 0) {
    Acc *a = acc + blockIdx.y;
    atomicMax(&a->largest, best);
    if (dead)
      atomicAdd(&a->bad, static_cast<unsigned long long>(dead));
    if (join)
      atomicAdd(&a->joins, static_cast<unsigned long long>(join));
    if (over)
      atomicAdd(&a->over, static_cast<unsigned long long>(over));
  }
}

__global__ void clear_acc(Acc *acc, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    acc[i] = Acc{0, 0, 0, 0, 0};
}

// A thread's device buffers and stream for the entry point, grow-only.
struct NarrowContext {
  hipStream_t stream = nullptr;
  sfi::DeviceBuffer jobs, acc;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  ~NarrowContext() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    for (hipEvent_t e : {t0, t1})
      if (e)
        (void)hipEventDestroy(e);
  }
};

NarrowContext &context() {
  static thread_local std::unique_ptr<NarrowContext> c;
  if (!c) {
    auto fresh = std::make_unique<NarrowContext>();
    check(hipStreamCreateWithFlags(&fresh->stream, hipStreamNonBlocking),
          "hipStreamCreate");
    check(hipEventCreate(&fresh->t0), "hipEventCreate");
    check(hipEventCreate(&fresh->t1), "hipEventCreate");
    c = std::move(fresh);
  }
  return *c;
}

} // namespace

extern "C" int sf_narrow_frames(const sf_narrow_job *jobs, size_t n,
                                sf_frame_scan *results, double *device_ms) {
  if ((n && (jobs == nullptr || results == nullptr)) || n > 65535) {
    sfi::set_last_error("bad argument: jobs, results or count");
    return -1;
  }
  for (size_t i = 0; i < n; i++) {
    const sf_narrow_job &j = jobs[i];
    const bool bits_ok = j.bits == 32 || (j.bits == 16 && j.out16 == nullptr);
    const uintptr_t align = reinterpret_cast<uintptr_t>(j.frame) |
                            reinterpret_cast<uintptr_t>(j.out16);
    if (!bits_ok || (j.pixels && j.frame == nullptr) || (align & 15) != 0 ||
        j.pixels > 0xFFFFFFFFull) {
      sfi::set_last_error("bad argument: a job's frame, output, bits, alignment or "
                          "size (16-byte aligned device frames below 2^32 pixels; "
                          "16-bit frames scan only)");
      return -1;
    }
  }
  try {
    if (n == 0)
      return 0;
    NarrowContext &c = context();
    std::vector<Job> host(n);
    uint64_t most = 0;
    for (size_t i = 0; i < n; i++) {
      host[i] = Job{reinterpret_cast<uint64_t>(jobs[i].frame),
                    reinterpret_cast<uint64_t>(jobs[i].out16), jobs[i].pixels,
                    jobs[i].bits, 0};
      most = jobs[i].pixels > most ? jobs[i].pixels : most;
    }
    c.jobs.ensure(n * sizeof(Job), c.stream);
    c.acc.ensure(n * sizeof(Acc), c.stream);
    check(hipMemcpyAsync(c.jobs.ptr, host.data(), n * sizeof(Job),
                         hipMemcpyHostToDevice, c.stream),
          "uploading the jobs");
    clear_acc<<<1, 256, 0, c.stream>>>(c.acc.as<Acc>(), static_cast<uint32_t>(n));
    // Enough blocks a frame to fill the device with one large frame, few for
    // small ones: up to 1024, each taking 2048 pixels a round.
    const uint64_t want = (most + kThreads * kPerThread - 1) / (kThreads * kPerThread);
    const uint32_t bx = static_cast<uint32_t>(want < 1 ? 1 : (want > 1024 ? 1024 : want));
    check(hipEventRecord(c.t0, c.stream), "hipEventRecord");
    narrow_scan<<<dim3(bx, static_cast<uint32_t>(n)), kThreads, 0, c.stream>>>(
        c.jobs.as<Job>(), c.acc.as<Acc>());
    check(hipGetLastError(), "launching the narrowing");
    check(hipEventRecord(c.t1, c.stream), "hipEventRecord");
    std::vector<Acc> acc(n);
    check(hipMemcpyAsync(acc.data(), c.acc.ptr, n * sizeof(Acc),
                         hipMemcpyDeviceToHost, c.stream),
          "reading the totals back");
    check(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
    for (size_t i = 0; i < n; i++) {
      results[i].largest = acc[i].largest;
      results[i].bad = acc[i].bad;
      results[i].joins = acc[i].joins;
      results[i].over = acc[i].over;
      results[i].fits = acc[i].largest < kSixteenBitLimit ? 1 : 0;
      results[i].reserved = 0;
    }
    if (device_ms) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, c.t0, c.t1), "hipEventElapsedTime");
      *device_ms = ms;
    }
    return 0;
  } catch (const std::exception &e) {
    sfi::set_last_error(e.what());
    return -3;
  }
}
