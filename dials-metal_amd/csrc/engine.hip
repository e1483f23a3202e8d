// The device engine: frame slots, descriptor upload, and the enqueue of a
// batch's threshold, ordering and labelling (see engine.hh).
#include <cmath>
#include <cstring>
#include <string>

#include "engine.hh"

namespace sfi {

namespace {

size_t align16(size_t b) { return (b + 15) & ~static_cast<size_t>(15); }

// The labelling's descriptors live in the second half of desc_, the
// threshold's in the first, so a captured threshold keeps its own.
constexpr size_t kDescHalf = 1 << 20;
constexpr size_t kLabelGraphs = 16; // captured frame counts kept at once

} // namespace

Engine::Engine(void *borrowed_stream) {
  int device = 0;
  check(hipGetDevice(&device), "hipGetDevice");
  if (borrowed_stream) {
    stream_ = static_cast<hipStream_t>(borrowed_stream);
  } else {
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
          "hipStreamCreate");
    own_stream_ = true;
  }
  for (hipEvent_t &e : staging_done_)
    check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  // R(sqrt(R(2) * (R(m) - 1))) for m = 0..127: the right-hand side of the
  // stage-0 test (dext.cc:105-106), by the host's correctly rounded sqrt.
  // Below two pixels the reference makes no test (dext.cc:99); +inf makes the
  // comparison false without a branch.
  float tf[128];
  double td[128];
  for (int m = 0; m < 128; m++) {
    tf[m] = m >= 2 ? std::sqrt(2.0f * (static_cast<float>(m) - 1.0f)) : HUGE_VALF;
    td[m] = m >= 2 ? std::sqrt(2.0 * (static_cast<double>(m) - 1.0)) : HUGE_VAL;
  }
  // Uploaded on the engine's own stream: a copy on the legacy stream would
  // be an implicit dependency on any stream another thread is capturing,
  // which invalidates that capture.
  table_.ensure(sizeof(tf) + sizeof(td), stream_);
  check(hipMemcpyAsync(table_.ptr, tf, sizeof(tf), hipMemcpyHostToDevice,
                       stream_),
        "uploading the root table");
  check(hipMemcpyAsync(table_.as<char>() + sizeof(tf), td, sizeof(td),
                       hipMemcpyHostToDevice, stream_),
        "uploading the root table");
  check(hipStreamSynchronize(stream_), "uploading the root table");
  desc_.ensure(2 * kDescHalf, stream_);
  queue_.ensure(kQueueBytes, stream_);
}

Engine::~Engine() {
  if (stream_)
    (void)hipStreamSynchronize(stream_);
  if (graph_exec_)
    (void)hipGraphExecDestroy(graph_exec_);
  if (graph_)
    (void)hipGraphDestroy(graph_);
  drop_label_graphs();
  for (hipEvent_t &e : marks_)
    if (e)
      (void)hipEventDestroy(e);
  for (hipEvent_t &e : staging_done_)
    if (e)
      (void)hipEventDestroy(e);
  if (host_counters_)
    (void)hipHostFree(host_counters_);
  for (char *s : staging_)
    if (s)
      (void)hipHostFree(s);
  slots_.clear();
  if (read_stream_) {
    (void)hipStreamSynchronize(read_stream_);
    (void)hipStreamDestroy(read_stream_);
  }
  if (own_stream_ && stream_)
    (void)hipStreamDestroy(stream_);
}

hipStream_t Engine::read_stream() {
  if (!read_stream_)
    check(hipStreamCreateWithFlags(&read_stream_, hipStreamNonBlocking),
          "hipStreamCreate");
  return read_stream_;
}

void Engine::mark(hipEvent_t e) { check(hipEventRecord(e, stream_), "hipEventRecord"); }

void Engine::sync_after(hipEvent_t done, const int *ids, size_t n, size_t *counts,
                        int *status, const int *later, const int *later_prev,
                        size_t nlater) {
  hipStream_t rs = read_stream();
  check(hipStreamWaitEvent(rs, done, 0), "hipStreamWaitEvent");
  check(hipMemcpyAsync(host_counters_, counters_.ptr,
                       kCounters * sizeof(uint32_t) * slots_.size(),
                       hipMemcpyDeviceToHost, rs),
        "reading the counters back");
  check(hipStreamSynchronize(rs), "hipStreamSynchronize");
  bool redo = false;
  for (size_t i = 0; i < n; i++) {
    const FrameSlot &s = slot(ids[i]);
    redo |= s.kernel == SF_KERNEL_FUSED && (host_counters_[kCounters * ids[i] + 1] != 0 ||
                                            fused_narrow(s.width, s.lane_columns));
  }
  if (redo) {
    sync(ids, n, counts, status);
    if (nlater) {
      label(later, later_prev, nlater);
      read_counters();
    }
    return;
  }
  for (size_t i = 0; i < n; i++) {
    const uint32_t c = host_counters_[kCounters * ids[i]];
    if (counts)
      counts[i] = c;
    if (status)
      status[i] = c > slot(ids[i]).capacity ? -2 : 0;
  }
}

void Engine::fetch_after(hipEvent_t done, int id, sf_signal_pixel *out, size_t n) {
  if (n == 0)
    return;
  hipStream_t rs = read_stream();
  check(hipStreamWaitEvent(rs, done, 0), "hipStreamWaitEvent");
  check(hipMemcpyAsync(out, slot(id).ordered.ptr, n * sizeof(sf_signal_pixel),
                       hipMemcpyDeviceToHost, rs),
        "copying the signal pixels back");
  check(hipStreamSynchronize(rs), "hipStreamSynchronize");
}

void Engine::grow(DeviceBuffer &b, size_t bytes) {
  if (b.ensure(bytes, stream_))
    epoch_++;
}

void Engine::ensure_slots(size_t n) {
  while (slots_.size() < n)
    slots_.push_back(std::make_unique<FrameSlot>());
  // The slots already in use keep their counts: a frame's labelling reads
  // the previous frame's.
  if (counters_.ensure_keep(kCounters * sizeof(uint32_t) * slots_.size(), stream_))
    epoch_++;
  if (host_counter_words_ < kCounters * slots_.size()) {
    check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    if (host_counters_)
      check(hipHostFree(host_counters_), "hipHostFree");
    host_counters_ = nullptr;
    host_counter_words_ = kCounters * slots_.size();
    check(hipHostMalloc(reinterpret_cast<void **>(&host_counters_),
                        host_counter_words_ * sizeof(uint32_t),
                        hipHostMallocDefault),
          "hipHostMalloc for the counters");
    std::memset(host_counters_, 0, host_counter_words_ * sizeof(uint32_t));
  }
}

char *Engine::stage(size_t bytes) {
  turn_ ^= 1;
  check(hipEventSynchronize(staging_done_[turn_]),
        "waiting for a descriptor staging buffer");
  if (bytes > staging_bytes_[turn_]) {
    if (staging_[turn_])
      check(hipHostFree(staging_[turn_]), "hipHostFree");
    staging_[turn_] = nullptr;
    check(hipHostMalloc(reinterpret_cast<void **>(&staging_[turn_]), bytes,
                        hipHostMallocDefault),
          "hipHostMalloc for the descriptors");
    staging_bytes_[turn_] = bytes;
  }
  return staging_[turn_];
}

void Engine::staged() {
  check(hipEventRecord(staging_done_[turn_], stream_), "hipEventRecord");
}

void Engine::shape(FrameSlot &s, int bits, uint32_t h, uint32_t w) {
  const size_t n = static_cast<size_t>(h) * w;
  s.wpr = mask_words_per_row(w);
  s.words = static_cast<size_t>(h) * s.wpr;
  // The reference's bound: a list the size of the frame's bytes
  // (dext_cuda.cu:736-741), one entry per 4 pixels at 32 bits, 8 at 16.
  const size_t cap = n * static_cast<size_t>(bits / 8) / sizeof(sf_signal_pixel);
  grow(s.mask, s.words * 4);
  grow(s.prefix, s.words * 4);
  grow(s.totals, scan_tmp_words(s.words) * 4);
  grow(s.unordered, (cap ? cap : 1) * sizeof(sf_signal_pixel));
  grow(s.ordered, (cap ? cap : 1) * sizeof(sf_signal_pixel));
  s.height = h;
  s.width = w;
  s.bits = bits;
  s.capacity = static_cast<uint32_t>(cap);
  s.thresholded = true;
  s.labelled = false;
  s.serial = ++serial_;
}

size_t Engine::slot_bytes(uint32_t height, uint32_t width, int bits) {
  const size_t n = static_cast<size_t>(height) * width;
  const size_t words = static_cast<size_t>(height) * mask_words_per_row(width);
  const size_t cap = n * static_cast<size_t>(bits / 8) / sizeof(sf_signal_pixel);
  // mask and prefix; two lists; parent and comp_of; components; links; the
  // root bits and scans; runs per row.
  return 8 * words + 32 * cap + 12 * (cap + 1) + sizeof(sf_component) * (cap + 1) +
         8 * (cap + 1) + 8 * (cap / 32 + 2) + 4 * (static_cast<size_t>(height) + 1) +
         (1u << 16);
}

size_t Engine::device_bytes() const {
  size_t total = counters_.bytes + table_.bytes + desc_.bytes + queue_.bytes +
                 scr0_.bytes + scr1_.bytes;
  for (const auto &p : slots_) {
    const FrameSlot &s = *p;
    for (const DeviceBuffer *b :
         {&s.mask, &s.prefix, &s.totals, &s.unordered, &s.ordered, &s.parent,
          &s.roots, &s.root_prefix, &s.comp_of, &s.comps, &s.links,
          &s.row_runs, &s.comp_runs})
      total += b->bytes;
  }
  return total;
}

void Engine::shape_label(FrameSlot &s) {
  const size_t cap = s.capacity;
  s.root_words = cap / 32 + 2;
  grow(s.parent, (cap + 1) * 4);
  grow(s.roots, s.root_words * 4);
  grow(s.root_prefix, s.root_words * 4);
  grow(s.row_runs, (static_cast<size_t>(s.height) + 1) * 4);
  grow(s.comp_runs, (cap + 2) * 4);
  grow(s.comp_of, (cap + 1) * 4);
  grow(s.comps, (cap + 1) * sizeof(sf_component));
  s.link_capacity = static_cast<uint32_t>(cap + 1);
  grow(s.links, static_cast<size_t>(s.link_capacity) * 8);
}

FrameDev Engine::dev(int id, int prev) {
  FrameSlot &s = slot(id);
  FrameDev f{};
  f.unordered = s.unordered.as<sf_signal_pixel>();
  f.ordered = s.ordered.as<sf_signal_pixel>();
  f.runs = s.unordered.as<uint4>();
  f.counter = counters_.as<uint32_t>() + kCounters * id;
  f.capacity = s.capacity;
  f.width = s.width;
  f.height = s.height;
  f.wpr = s.wpr;
  f.mask = s.mask.as<uint32_t>();
  f.prefix = s.prefix.as<uint32_t>();
  f.totals = s.totals.as<uint32_t>();
  f.parent = s.parent.as<uint32_t>();
  f.roots = s.roots.as<uint32_t>();
  f.root_prefix = s.root_prefix.as<uint32_t>();
  f.comp_of = s.comp_of.as<uint32_t>();
  f.row_runs = s.row_runs.as<uint32_t>();
  f.comp_runs = s.comp_runs.as<uint32_t>();
  f.comps = s.comps.as<sf_component>();
  f.links = s.links.as<uint32_t>();
  f.link_capacity = s.link_capacity;
  if (prev >= 0) {
    FrameSlot &p = slot(prev);
    f.prev_mask = p.mask.as<uint32_t>();
    f.prev_prefix = p.prefix.as<uint32_t>();
    f.prev_comp_of = p.comp_of.as<uint32_t>();
    f.prev_counter = counters_.as<uint32_t>() + kCounters * prev;
    f.prev_capacity = p.capacity;
  }
  return f;
}

ScanSeg Engine::mask_seg(int id) {
  FrameSlot &s = slot(id);
  return ScanSeg{s.mask.as<uint32_t>(), s.prefix.as<uint32_t>(),
                 s.totals.as<uint32_t>(), s.words, nullptr, 0, 1};
}

FrameArgs Engine::args(int id, const void *frame) {
  FrameSlot &s = slot(id);
  FrameArgs a;
  a.frame = frame;
  a.height = s.height;
  a.width = s.width;
  a.bits = s.bits;
  a.out = s.unordered.as<sf_signal_pixel>();
  a.counter = counters_.as<uint32_t>() + kCounters * id;
  a.capacity = s.capacity;
  a.mask = s.mask.as<uint32_t>();
  a.wpr = s.wpr;
  a.table = s.bits == 16
                ? table_.ptr
                : static_cast<const void *>(table_.as<char>() + 128 * sizeof(float));
  a.overflow = a.counter + 1;
  a.totals = s.totals.as<uint32_t>();
  return a;
}

// Fills and uploads the descriptors of a batch (the frame pointers among
// them) into the first half of desc_.
Engine::Prepared Engine::prepare(const FrameIn *frames, const int *ids,
                                 size_t n, int kernel) {
  uint64_t mix_rows = 0; // the batch's size, whatever shape plans it
  for (size_t i = 0; i < n; i++)
    mix_rows += static_cast<uint64_t>(slot(ids[i]).height) * strips_of(slot(ids[i]).width, 4);
  const Tuning t = tuning_for_mix(mix_rows);
  Prepared p;
  p.b_fused = align16(n * sizeof(FusedParams));
  p.b_frames = align16(n * sizeof(FrameDev));
  p.b_segs = align16(n * sizeof(ScanSeg));
  const size_t bytes = p.b_fused + p.b_frames + p.b_segs;
  if (bytes > kDescHalf)
    throw device_error("too many frames in one batch");
  char *host = stage(bytes);
  FusedParams *fp = reinterpret_cast<FusedParams *>(host);
  FrameDev *fd = reinterpret_cast<FrameDev *>(host + p.b_fused);
  ScanSeg *sg = reinterpret_cast<ScanSeg *>(host + p.b_fused + p.b_frames);
  p.lane_columns = t.lane_columns;
  p.waves_per_block = t.waves_per_block;
  uint64_t strip_rows = 0;
  for (size_t i = 0; i < n; i++)
    strip_rows += static_cast<uint64_t>(slot(ids[i]).height) *
                  strips_of(slot(ids[i]).width, p.lane_columns);
  const uint32_t seg = segment_rows(strip_rows, t);
  p.seg_rows = seg;
  for (size_t i = 0; i < n; i++) {
    FrameSlot &s = slot(ids[i]);
    s.frame = frames[i].frame;
    s.kernel = kernel;
    s.lane_columns = p.lane_columns;
    FusedPlan plan;
    const uint32_t k = plan_fused(plan, s.height, s.width, seg, p.lane_columns);
    fp[i] = to_params(args(ids[i], frames[i].frame), plan, p.items);
    p.items += k;
    fd[i] = dev(ids[i], -1);
    sg[i] = mask_seg(ids[i]);
    p.max_words = s.words > p.max_words ? s.words : p.max_words;
  }
  check(hipMemcpyAsync(desc_.ptr, host, bytes, hipMemcpyHostToDevice, stream_),
        "uploading frame descriptors");
  staged();
  return p;
}

// The batch's kernels: clear, threshold, scan, scatter.  Nothing here
// allocates or copies from the host, so the sequence can be captured.
void Engine::enqueue_kernels(const Prepared &p, const FrameIn *frames,
                             const int *ids, size_t n, int kernel, bool timing) {
  char *d = desc_.as<char>();
  const FusedParams *fp_d = reinterpret_cast<const FusedParams *>(d);
  const FrameDev *fd_d = reinterpret_cast<const FrameDev *>(d + p.b_fused);
  const ScanSeg *sg_d =
      reinterpret_cast<const ScanSeg *>(d + p.b_fused + p.b_frames);
  const uint32_t nf = static_cast<uint32_t>(n);
  const bool fused = kernel == SF_KERNEL_FUSED;
  launch_clear(fd_d, nf, fused, 0, stream_,
               fused ? queue_.as<uint32_t>() : nullptr);
  if (timing)
    check(hipEventRecord(marks_[0], stream_), "hipEventRecord");
  if (fused) {
    launch_fused_batch(fp_d, nf, p.items, table_.as<float>(),
                       reinterpret_cast<const double *>(table_.as<char>() +
                                                        128 * sizeof(float)),
                       queue_.as<uint32_t>(), p.lane_columns, p.waves_per_block, stream_);
  } else {
    for (size_t i = 0; i < n; i++)
      launch_direct(args(ids[i], frames[i].frame), scr0_.as<uint8_t>(),
                    scr1_.as<uint8_t>(), stream_);
  }
  if (timing)
    check(hipEventRecord(marks_[1], stream_), "hipEventRecord");
  // The fused kernel counts each block's signal pixels as it emits them.
  launch_popcount_scan(sg_d, nf, p.max_words, stream_, fused);
  launch_scatter(fd_d, nf, stream_);
  if (timing)
    check(hipEventRecord(marks_[2], stream_), "hipEventRecord");
}

void Engine::threshold(const FrameIn *frames, const int *ids, size_t n,
                       int kernel, bool graph) {
  if (n == 0)
    return;
  int most = 0;
  for (size_t i = 0; i < n; i++)
    most = ids[i] > most ? ids[i] : most;
  ensure_slots(static_cast<size_t>(most) + 1);
  size_t pixels = 0;
  for (size_t i = 0; i < n; i++) {
    shape(slot(ids[i]), frames[i].bits, frames[i].height, frames[i].width);
    const size_t px = static_cast<size_t>(frames[i].height) * frames[i].width;
    pixels = px > pixels ? px : pixels;
  }
  if (kernel == SF_KERNEL_DIRECT) {
    grow(scr0_, pixels);
    grow(scr1_, pixels);
  }
  const bool timing = timing_on_;
  timed_ = timing;
  if (timing && !marks_[0])
    for (hipEvent_t &e : marks_)
      check(hipEventCreate(&e), "hipEventCreate");
  // Only the fused sequence is captured: the direct kernels take their frame
  // and slot pointers as launch arguments, which a replay would repeat.
  if (!graph || timing || kernel != SF_KERNEL_FUSED) {
    const Prepared p = prepare(frames, ids, n, kernel);
    enqueue_kernels(p, frames, ids, n, kernel, timing);
    return;
  }
  // One graph per (shapes, tuning, buffer epoch): every frame and slot
  // pointer the fused sequence uses is read from the descriptors uploaded
  // just before each launch, outside the graph, and the engine buffers it
  // names change the epoch when they move; so one replay serves any batch of
  // these shapes, in either set of slots.
  const Prepared p = prepare(frames, ids, n, kernel);
  std::vector<uint64_t> key = {static_cast<uint64_t>(kernel), epoch_,
                               static_cast<uint64_t>(p.seg_rows),
                               static_cast<uint64_t>(p.waves_per_block),
                               static_cast<uint64_t>(p.lane_columns)};
  for (size_t i = 0; i < n; i++)
    key.push_back((static_cast<uint64_t>(frames[i].height) << 32) |
                  (static_cast<uint64_t>(frames[i].width) << 8) |
                  static_cast<uint64_t>(frames[i].bits));
  if (!(graph_exec_ && key == graph_key_)) {
    if (graph_exec_)
      check(hipGraphExecDestroy(graph_exec_), "hipGraphExecDestroy");
    if (graph_)
      check(hipGraphDestroy(graph_), "hipGraphDestroy");
    graph_exec_ = nullptr;
    graph_ = nullptr;
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
          "hipStreamBeginCapture");
    enqueue_kernels(p, frames, ids, n, kernel, false);
    check(hipStreamEndCapture(stream_, &graph_), "hipStreamEndCapture");
    check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0),
          "hipGraphInstantiate");
    graph_key_ = key;
  }
  check(hipGraphLaunch(graph_exec_, stream_), "hipGraphLaunch");
}

void Engine::read_counters() {
  check(hipMemcpyAsync(host_counters_, counters_.ptr,
                       kCounters * sizeof(uint32_t) * slots_.size(),
                       hipMemcpyDeviceToHost, stream_),
        "reading the counters back");
  check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Engine::sync(const int *ids, size_t n, size_t *counts, int *status) {
  read_counters();
  size_t first_redone = n;
  for (size_t i = 0; i < n; i++) {
    FrameSlot &s = slot(ids[i]);
    if (s.kernel == SF_KERNEL_FUSED && (host_counters_[kCounters * ids[i] + 1] != 0 ||
                                        fused_narrow(s.width, s.lane_columns))) {
      first_redone = first_redone < i ? first_redone : i;
      // A valid pixel beyond the fused kernel's exact range (or a frame it
      // left alone, narrower than a lane's columns): this frame is
      // thresholded again, in its own slot, by the direct kernels, whose sums
      // are in the reference's types.
      fallbacks_++;
      const FrameIn f{s.frame, s.height, s.width, s.bits};
      const size_t px = static_cast<size_t>(s.height) * s.width;
      grow(scr0_, px);
      grow(scr1_, px);
      const Prepared p = prepare(&f, &ids[i], 1, SF_KERNEL_DIRECT);
      enqueue_kernels(p, &f, &ids[i], 1, SF_KERNEL_DIRECT, false);
      read_counters();
    }
  }
  // Labels enqueued before this sync were taken from the lists the exact path
  // replaced: relabel the re-run frames and, since a frame's links read the
  // components of the frame before, every labelled frame after the first.
  bool relabelled = false;
  for (size_t i = first_redone; i < n; i++) {
    FrameSlot &s = slot(ids[i]);
    if (!s.labelled)
      continue;
    int prev = s.linked_slot;
    if (prev >= 0 && slot(prev).serial != s.linked_serial)
      prev = -1;
    label(&ids[i], &prev, 1);
    relabelled = true;
  }
  if (relabelled)
    read_counters();
  for (size_t i = 0; i < n; i++) {
    const uint32_t c = host_counters_[kCounters * ids[i]];
    if (counts)
      counts[i] = c;
    if (status)
      status[i] = c > slot(ids[i]).capacity ? -2 : 0;
  }
}

void Engine::fetch(int id, sf_signal_pixel *out, size_t n) {
  if (n == 0)
    return;
  check(hipMemcpyAsync(out, slot(id).ordered.ptr, n * sizeof(sf_signal_pixel),
                       hipMemcpyDeviceToHost, stream_),
        "copying the signal pixels back");
  check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Engine::load_signal(int id, const sf_signal_pixel *list, size_t n,
                         uint32_t height, uint32_t width) {
  ensure_slots(static_cast<size_t>(id) + 1);
  FrameSlot &s = slot(id);
  shape(s, 32, height, width);
  if (n > s.capacity) {
    // Labelling a list never needs the threshold's bound; size for it (the
    // unordered list's memory holds the labelling's runs).
    s.capacity = static_cast<uint32_t>(n);
    grow(s.ordered, n * sizeof(sf_signal_pixel));
    grow(s.unordered, n * sizeof(sf_signal_pixel));
  }
  s.kernel = SF_KERNEL_DIRECT;
  const size_t b_frames = align16(sizeof(FrameDev));
  const size_t b_segs = align16(sizeof(ScanSeg));
  const size_t bytes = b_frames + b_segs + 16;
  char *host = stage(bytes);
  *reinterpret_cast<FrameDev *>(host) = dev(id, -1);
  *reinterpret_cast<ScanSeg *>(host + b_frames) = mask_seg(id);
  *reinterpret_cast<uint32_t *>(host + b_frames + b_segs) =
      static_cast<uint32_t>(n);
  char *d = desc_.as<char>() + kDescHalf;
  check(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, stream_),
        "uploading frame descriptors");
  staged();
  const FrameDev *fd_d = reinterpret_cast<const FrameDev *>(d);
  launch_clear(fd_d, 1, true, 0, stream_);
  check(hipMemcpyAsync(counters_.as<uint32_t>() + kCounters * id, d + b_frames + b_segs,
                       sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_),
        "setting the signal count");
  if (n)
    check(hipMemcpyAsync(s.ordered.ptr, list, n * sizeof(sf_signal_pixel),
                         hipMemcpyHostToDevice, stream_),
          "uploading a signal list");
  launch_set_bits(fd_d, 1, stream_);
  launch_popcount_scan(reinterpret_cast<const ScanSeg *>(d + b_frames), 1,
                       s.words, stream_);
  check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Engine::label(const int *ids, const int *prev, size_t n) {
  if (n == 0)
    return;
  for (size_t i = 0; i < n; i++)
    shape_label(slot(ids[i]));
  const size_t bytes = align16(n * sizeof(FrameDev));
  if (bytes > kDescHalf)
    throw device_error("too many frames in one batch");
  char *host = stage(bytes);
  FrameDev *fd = reinterpret_cast<FrameDev *>(host);
  for (size_t i = 0; i < n; i++)
    fd[i] = dev(ids[i], prev ? prev[i] : -1);
  char *d = desc_.as<char>() + kDescHalf;
  check(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, stream_),
        "uploading label descriptors");
  staged();
  // Every buffer the labelling touches is named by the descriptors just
  // uploaded, so one capture per frame count serves any frames and slots.
  const uint32_t nframes = static_cast<uint32_t>(n);
  auto it = label_graphs_.find(nframes);
  if (it == label_graphs_.end()) {
    if (label_graphs_.size() >= kLabelGraphs)
      drop_label_graphs();
    LabelGraph g;
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
          "hipStreamBeginCapture");
    launch_label(reinterpret_cast<const FrameDev *>(d), nframes, stream_);
    check(hipStreamEndCapture(stream_, &g.graph), "hipStreamEndCapture");
    const hipError_t e =
        hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g.graph);
      check(e, "hipGraphInstantiate");
    }
    it = label_graphs_.emplace(nframes, g).first;
  }
  check(hipGraphLaunch(it->second.exec, stream_), "hipGraphLaunch");
  for (size_t i = 0; i < n; i++) {
    slot(ids[i]).labelled = true;
    slot(ids[i]).linked_serial =
        (prev && prev[i] >= 0) ? slot(prev[i]).serial : 0;
    slot(ids[i]).linked_slot = (prev && prev[i] >= 0) ? prev[i] : -1;
  }
}

void Engine::drop_label_graphs() {
  for (auto &entry : label_graphs_) {
    if (entry.second.exec)
      (void)hipGraphExecDestroy(entry.second.exec);
    if (entry.second.graph)
      (void)hipGraphDestroy(entry.second.graph);
  }
  label_graphs_.clear();
}

void Engine::label_counts(int id, size_t *n, size_t *ncomp, size_t *nlinks,
                          bool fresh) {
  if (fresh)
    read_counters();
  const uint32_t *c = host_counters_ + kCounters * id;
  if (c[3] > slot(id).link_capacity)
    throw device_error("more frame links than the link list holds");
  if (n)
    *n = c[0];
  if (ncomp)
    *ncomp = c[2];
  if (nlinks)
    *nlinks = c[3];
}

void Engine::fetch_labels(int id, uint32_t *comp_of, size_t n,
                          sf_component *comps, size_t ncomp, uint32_t *links,
                          size_t nlinks, sf_signal_pixel *list, bool wait) {
  FrameSlot &s = slot(id);
  if (list && n)
    check(hipMemcpyAsync(list, s.ordered.ptr, n * sizeof(sf_signal_pixel),
                         hipMemcpyDeviceToHost, stream_),
          "copying the signal pixels back");
  if (comp_of && n)
    check(hipMemcpyAsync(comp_of, s.comp_of.ptr, n * 4, hipMemcpyDeviceToHost,
                         stream_),
          "copying the labels back");
  if (comps && ncomp)
    check(hipMemcpyAsync(comps, s.comps.ptr, ncomp * sizeof(sf_component),
                         hipMemcpyDeviceToHost, stream_),
          "copying the components back");
  if (links && nlinks)
    check(hipMemcpyAsync(links, s.links.ptr, nlinks * 8, hipMemcpyDeviceToHost,
                         stream_),
          "copying the frame links back");
  if (wait)
    check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Engine::kernel_ms(double *threshold_ms, double *order_ms) {
  if (!timed_)
    throw device_error("no timed batch");
  check(hipEventSynchronize(marks_[2]), "hipEventSynchronize");
  float a = 0, b = 0;
  check(hipEventElapsedTime(&a, marks_[0], marks_[1]), "hipEventElapsedTime");
  check(hipEventElapsedTime(&b, marks_[1], marks_[2]), "hipEventElapsedTime");
  if (threshold_ms)
    *threshold_ms = a;
  if (order_ms)
    *order_ms = b;
}

} // namespace sfi
