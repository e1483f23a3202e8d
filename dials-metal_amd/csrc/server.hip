// The serving loop behind sf_server_*: frames from several concurrent sweeps,
// taken as they arrive.
//
// Submitters queue requests {sweep, device frame or compressed host chunk,
// optional host list} stamped with the time of submission.  One worker
// thread owns the engine and keeps two batches in flight: it takes every
// request that has arrived into the next batch -- at most ring - 1 frames a
// sweep counting the batch still on the device, so that the sweep's previous
// frame keeps its slot, and at most max_batch in all -- uploads that batch's
// chunks in one copy and decodes them into the sweep's frame buffers,
// thresholds the batch in one persistent fused launch and labels it (each
// frame linked to its own sweep's previous frame), all enqueued behind the
// batch before; then it waits for the batch before (an event, not the
// stream), reads its counts on a stream of its own, copies the requested
// lists out and completes its requests.  Requests arriving while two batches
// are out wait for the next: continuous batching, with batch sizes following
// the load, and the device kept busy across the host's turn-around.  Each
// sweep holds a ring of slots of its shape for its whole life, so its
// buffers and label scratch stay resident.  A sweep is admitted only while
// the device's free memory stays above the configured reserve after its
// slots.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "engine.hh"
#include "sf_internal.hh"

using namespace sfi;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id;
  int sweep;
  const void *frame; // a device frame, or null for a chunk
  sf_signal_pixel *host_out;
  size_t host_capacity;
  Clock::time_point arrived;
  const uint8_t *chunk = nullptr; // a compressed chunk in host memory
  size_t chunk_size = 0;
  char algorithm[8] = {};
};

struct Sweep {
  uint32_t height, width;
  int bits;
  int first_slot;     // slots first_slot .. first_slot + ring - 1
  uint64_t taken = 0; // frames taken into batches so far
  int last_slot = -1; // the slot of its last frame, once labelled
  // Frame buffers chunks decode into, one per slot (made on the first chunk).
  std::vector<std::unique_ptr<DeviceBuffer>> chunk_frames;
};

// A batch taken from the queue: its requests, where their frames are, and
// the event its work ends with.
struct Batch {
  std::vector<Request> reqs;
  std::vector<FrameIn> frames;
  std::vector<int> ids, prevs;
  std::vector<sf_chunk> chunks;
  std::vector<int> chunk_of; // each request's chunk, or -1
  std::vector<uint32_t> per_sweep;
  int decoder = -1; // which decoder holds its chunks
  hipEvent_t done = nullptr;
  std::string error; // set when enqueueing it failed
  bool empty() const { return reqs.empty(); }
  void clear(size_t sweeps) {
    reqs.clear();
    frames.clear();
    ids.clear();
    prevs.clear();
    chunks.clear();
    chunk_of.clear();
    per_sweep.assign(sweeps, 0);
    decoder = -1;
    error.clear();
  }
};

struct Outcome {
  sf_server_result result;
  std::string error;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

int bad(const char *what) {
  set_last_error(std::string("bad argument: ") + what);
  return -1;
}

} // namespace

struct sf_server {
  explicit sf_server(const sf_server_config &c) : cfg(c), engine(nullptr) {
    check(hipGetDevice(&device), "hipGetDevice");
    worker = std::thread([this] { run(); });
  }
  ~sf_server() {
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    work.notify_all();
    worker.join();
  }

  sf_server_config cfg;
  int device = 0;
  Engine engine; // the worker's alone once it runs
  sf_decoder *decoders[2] = {nullptr, nullptr}; // the worker's, used in turn
  int next_decoder = 0;

  std::mutex m;
  std::condition_variable work, finished;
  std::deque<Request> queue;
  std::vector<Sweep> sweeps;
  std::unordered_map<uint64_t, Outcome> outcomes; // until waited for
  std::unordered_set<uint64_t> pending;           // submitted, not waited for
  std::vector<double> latencies;                  // since the last reset
  uint64_t next_id = 1, batches = 0, batch_frames = 0;
  bool started = false;
  Clock::time_point first_submit, last_done;
  bool stop = false;
  std::thread worker;

  void run();
  // Takes what has arrived into `b` (under the lock), within the sweeps'
  // rings counting `out`, the batch on the device.
  void take(Batch &b, const Batch &out);
  // Enqueues a taken batch: chunks, threshold, labels, its event.
  void launch(Batch &b);
  // Waits for a launched batch and completes its requests; `behind` is the
  // batch launched after it.
  void complete(Batch &b, const Batch &behind);
};

void sf_server::take(Batch &b, const Batch &out) {
  b.clear(sweeps.size());
  // Everything that has arrived, in order, at most ring - 1 frames a sweep
  // with those on the device: once a sweep reaches that, its later frames
  // wait too, so each sweep's frames stay in order.
  for (auto it = queue.begin(); it != queue.end() && b.reqs.size() < cfg.max_batch;) {
    const size_t w = static_cast<size_t>(it->sweep);
    const uint32_t held = b.per_sweep[w] + (w < out.per_sweep.size() ? out.per_sweep[w] : 0);
    if (held + 1 >= cfg.ring) {
      ++it;
      continue;
    }
    b.per_sweep[w]++;
    b.reqs.push_back(*it);
    it = queue.erase(it);
  }
  const size_t n = b.reqs.size();
  b.frames.resize(n);
  b.ids.resize(n);
  b.prevs.resize(n);
  b.chunk_of.assign(n, -1);
  for (size_t i = 0; i < n; i++) {
    Sweep &s = sweeps[static_cast<size_t>(b.reqs[i].sweep)];
    const int k = static_cast<int>(s.taken % cfg.ring);
    const int slot = s.first_slot + k;
    s.taken++;
    b.frames[i] = FrameIn{b.reqs[i].frame, s.height, s.width, s.bits};
    b.ids[i] = slot;
    b.prevs[i] = cfg.label ? s.last_slot : -1;
    s.last_slot = slot;
  }
}

void sf_server::launch(Batch &b) {
  const size_t n = b.reqs.size();
  try {
    for (size_t i = 0; i < n; i++) {
      const Request &r = b.reqs[i];
      if (r.frame)
        continue; // a device frame
      Sweep &s = sweeps[static_cast<size_t>(r.sweep)];
      const size_t k = static_cast<size_t>(b.ids[i] - s.first_slot);
      if (s.chunk_frames.empty())
        s.chunk_frames.resize(cfg.ring);
      if (!s.chunk_frames[k])
        s.chunk_frames[k] = std::make_unique<DeviceBuffer>();
      DeviceBuffer &f = *s.chunk_frames[k];
      f.ensure(sf_frame_bytes(s.height, s.width, static_cast<unsigned>(s.bits)),
               engine.stream());
      sf_chunk c{};
      c.data = r.chunk;
      c.size = r.chunk_size;
      c.algorithm = r.algorithm;
      c.bit_depth = static_cast<uint32_t>(s.bits);
      c.height = s.height;
      c.width = s.width;
      c.frame = f.ptr;
      b.chunk_of[i] = static_cast<int>(b.chunks.size());
      b.chunks.push_back(c);
      b.frames[i].frame = f.ptr;
    }
    if (!b.chunks.empty()) {
      // The batch before may still be reading the other decoder's status.
      b.decoder = next_decoder;
      next_decoder ^= 1;
      sf_decoder *&d = decoders[b.decoder];
      if (!d && !(d = sf_decoder_create(engine.stream())))
        throw device_error("could not make a decoder: " + std::string(sf_last_error()));
      decoder_load(d, b.chunks.data(), b.chunks.size());
      decoder_run(d);
    }
    engine.threshold(b.frames.data(), b.ids.data(), n, SF_KERNEL_FUSED, false);
    if (cfg.label)
      engine.label(b.ids.data(), b.prevs.data(), n);
    engine.mark(b.done);
  } catch (const std::exception &e) {
    b.error = e.what();
  }
}

void sf_server::complete(Batch &b, const Batch &behind) {
  const size_t n = b.reqs.size();
  std::vector<size_t> counts(n, 0);
  std::vector<int> status(n, 0), chunk_status;
  std::vector<std::string> chunk_msgs;
  std::vector<Outcome> outs(n);
  std::string error = b.error;
  if (error.empty()) {
    try {
      const bool relabel = cfg.label && !behind.empty() && behind.error.empty();
      engine.sync_after(b.done, b.ids.data(), n, counts.data(), status.data(),
                        relabel ? behind.ids.data() : nullptr,
                        relabel ? behind.prevs.data() : nullptr,
                        relabel ? behind.ids.size() : 0);
      if (b.decoder >= 0) {
        chunk_status.assign(b.chunks.size(), 0);
        (void)decoder_status_after(decoders[b.decoder], engine.read_stream(), b.done,
                                   chunk_status.data(), &chunk_msgs);
      }
      for (size_t i = 0; i < n; i++) {
        sf_server_result &r = outs[i].result;
        const int c = b.chunk_of[i];
        if (c >= 0 && chunk_status[static_cast<size_t>(c)] != 0) {
          r.status = -1;
          outs[i].error = chunk_msgs[static_cast<size_t>(c)];
          continue;
        }
        r.signal_pixels = counts[i];
        r.status = status[i];
        if (cfg.label && status[i] == 0) {
          size_t np = 0, nc = 0, nl = 0;
          engine.label_counts(b.ids[i], &np, &nc, &nl, false);
          r.components = nc;
          r.links = nl;
        }
        if (b.reqs[i].host_out && status[i] == 0)
          engine.fetch_after(b.done, b.ids[i], b.reqs[i].host_out,
                             std::min(counts[i], b.reqs[i].host_capacity));
      }
    } catch (const std::exception &e) {
      error = e.what();
    }
  }
  const Clock::time_point done = Clock::now();

  {
    std::lock_guard<std::mutex> lk(m);
    for (size_t i = 0; i < n; i++) {
      Outcome &o = outs[i];
      o.result.request = b.reqs[i].id;
      o.result.sweep = b.reqs[i].sweep;
      o.result.latency_ms = ms_between(b.reqs[i].arrived, done);
      o.result.batch_frames = static_cast<uint32_t>(n);
      if (!error.empty()) {
        o.result.status = -3;
        o.error = error;
      } else if (o.result.status == 0 || o.result.status == -2) {
        latencies.push_back(o.result.latency_ms);
      }
      outcomes[b.reqs[i].id] = o;
    }
    batches++;
    batch_frames += n;
    last_done = done;
  }
  finished.notify_all();
}

void sf_server::run() {
  (void)hipSetDevice(device);
  Batch batches_[2];
  for (Batch &b : batches_)
    if (hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess)
      b.done = nullptr;
  Batch *out = &batches_[0], *next = &batches_[1]; // on the device; to take
  out->clear(0);
  next->clear(0);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(m);
      if (out->empty()) {
        work.wait(lk, [&] { return stop || !queue.empty(); });
        if (queue.empty())
          break; // stopping, nothing left
      } else {
        // A batch is on the device: take the next as soon as anything
        // arrives, or go and complete this one once it is done.
        while (queue.empty() && hipEventQuery(out->done) == hipErrorNotReady)
          work.wait_for(lk, std::chrono::microseconds(20));
      }
      take(*next, *out);
    }
    if (!next->empty())
      launch(*next);
    if (!out->empty())
      complete(*out, *next);
    std::swap(out, next);
    next->clear(0);
  }
  for (Batch &b : batches_)
    if (b.done)
      (void)hipEventDestroy(b.done);
  for (sf_decoder *&d : decoders) {
    sf_decoder_destroy(d);
    d = nullptr;
  }
}

namespace {
template <typename F> int guarded(F &&body) {
  try {
    return body();
  } catch (const device_error &e) {
    set_last_error(e.what());
    return -3;
  } catch (const std::bad_alloc &) {
    set_last_error("out of host memory");
    return -4;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -3;
  }
}
} // namespace

extern "C" {

void sf_server_config_default(sf_server_config *c) {
  if (!c)
    return;
  c->max_batch = 64;
  c->ring = 4;
  c->label = 1;
  c->reserve_bytes = 1ull << 30;
}

int sf_server_create(const sf_server_config *c, sf_server **out) {
  if (!out)
    return bad("no place for the server");
  *out = nullptr;
  sf_server_config cfg;
  sf_server_config_default(&cfg);
  if (c) {
    cfg = *c;
    if (cfg.max_batch == 0)
      cfg.max_batch = 64;
    if (cfg.ring == 0)
      cfg.ring = 4;
  }
  if (cfg.ring < 2)
    return bad("a sweep needs a ring of at least two slots");
  return guarded([&]() -> int {
    *out = new sf_server(cfg);
    return 0;
  });
}

void sf_server_destroy(sf_server *s) { delete s; }

int sf_server_open_sweep(sf_server *s, uint32_t height, uint32_t width,
                         int bits, int *sweep) {
  if (!s || !sweep || height == 0 || width == 0 ||
      static_cast<uint64_t>(height) * width > 0x7FFFFFFFull ||
      (bits != 16 && bits != 32))
    return bad("server, shape or pixel type");
  return guarded([&]() -> int {
    const size_t need = Engine::slot_bytes(height, width, bits) * s->cfg.ring;
    const size_t free_now = sf_memory_free();
    if (free_now < need + s->cfg.reserve_bytes) {
      set_last_error("not admitted: the sweep's slots would leave less free "
                     "device memory than the reserve");
      return -6;
    }
    std::lock_guard<std::mutex> lk(s->m);
    Sweep w{height, width, bits, static_cast<int>(s->sweeps.size() * s->cfg.ring)};
    s->sweeps.push_back(std::move(w));
    *sweep = static_cast<int>(s->sweeps.size() - 1);
    return 0;
  });
}

int sf_server_submit(sf_server *s, int sweep, const void *device_frame,
                     sf_signal_pixel *host_out, size_t host_capacity,
                     uint64_t *request) {
  if (!s || !device_frame || !request)
    return bad("server, frame or request");
  {
    std::lock_guard<std::mutex> lk(s->m);
    if (sweep < 0 || static_cast<size_t>(sweep) >= s->sweeps.size())
      return bad("no such sweep");
    const Clock::time_point now = Clock::now();
    if (!s->started) {
      s->started = true;
      s->first_submit = now;
    }
    *request = s->next_id++;
    s->pending.insert(*request);
    s->queue.push_back(
        Request{*request, sweep, device_frame, host_out, host_capacity, now});
  }
  s->work.notify_one();
  return 0;
}

int sf_server_submit_chunk(sf_server *s, int sweep, const uint8_t *data, size_t size,
                           const char *algorithm, sf_signal_pixel *host_out,
                           size_t host_capacity, uint64_t *request) {
  if (!s || !request || (size && !data))
    return bad("server, chunk or request");
  const std::string alg = algorithm ? algorithm : "";
  if (alg.size() >= sizeof(Request{}.algorithm)) {
    set_last_error("unknown compression algorithm " + alg);
    return -1;
  }
  Request r{};
  sf_chunk c{};
  {
    std::lock_guard<std::mutex> lk(s->m);
    if (sweep < 0 || static_cast<size_t>(sweep) >= s->sweeps.size())
      return bad("no such sweep");
    const Sweep &w = s->sweeps[static_cast<size_t>(sweep)];
    c.bit_depth = static_cast<uint32_t>(w.bits);
    c.height = w.height;
    c.width = w.width;
  }
  c.data = data;
  c.size = size;
  c.algorithm = alg.c_str();
  try {
    check_chunk(c); // the chunk describes its frame (decompress.cc:76-133)
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return -1;
  }
  {
    std::lock_guard<std::mutex> lk(s->m);
    const Clock::time_point now = Clock::now();
    if (!s->started) {
      s->started = true;
      s->first_submit = now;
    }
    *request = s->next_id++;
    s->pending.insert(*request);
    r.id = *request;
    r.sweep = sweep;
    r.frame = nullptr;
    r.host_out = host_out;
    r.host_capacity = host_capacity;
    r.arrived = now;
    r.chunk = data;
    r.chunk_size = size;
    std::memcpy(r.algorithm, alg.c_str(), alg.size() + 1);
    s->queue.push_back(r);
  }
  s->work.notify_one();
  return 0;
}

int sf_server_wait(sf_server *s, uint64_t request, sf_server_result *result) {
  if (!s || !result)
    return bad("server or result");
  std::unique_lock<std::mutex> lk(s->m);
  if (s->pending.count(request) == 0)
    return bad("no such request, or one already waited for");
  s->finished.wait(lk, [&] { return s->outcomes.count(request) != 0; });
  const Outcome o = s->outcomes[request];
  s->outcomes.erase(request);
  s->pending.erase(request);
  *result = o.result;
  if (o.result.status == -3 || o.result.status == -1) {
    set_last_error(o.error);
    return o.result.status;
  }
  return 0;
}

int sf_server_stats_get(sf_server *s, sf_server_stats *stats) {
  if (!s || !stats)
    return bad("server or stats");
  std::lock_guard<std::mutex> lk(s->m);
  *stats = sf_server_stats{};
  std::vector<double> l = s->latencies;
  stats->frames = l.size();
  stats->batches = s->batches;
  if (!l.empty()) {
    std::sort(l.begin(), l.end());
    auto at = [&](double q) {
      const size_t i = static_cast<size_t>(q * static_cast<double>(l.size() - 1) + 0.5);
      return l[std::min(i, l.size() - 1)];
    };
    double sum = 0;
    for (double v : l)
      sum += v;
    stats->p50_ms = at(0.50);
    stats->p99_ms = at(0.99);
    stats->max_ms = l.back();
    stats->mean_ms = sum / static_cast<double>(l.size());
    const double secs = ms_between(s->first_submit, s->last_done) / 1e3;
    stats->frames_per_s = secs > 0 ? static_cast<double>(l.size()) / secs : 0.0;
    stats->mean_batch = s->batches
                            ? static_cast<double>(s->batch_frames) /
                                  static_cast<double>(s->batches)
                            : 0.0;
  }
  return 0;
}

void sf_server_stats_reset(sf_server *s) {
  if (!s)
    return;
  std::lock_guard<std::mutex> lk(s->m);
  s->latencies.clear();
  s->batches = 0;
  s->batch_frames = 0;
  s->started = false;
}

extern "C++" {
namespace {

// Submits item i at arrival_s[i] seconds after the call (sleeping to within
// 200 us of the arrival, then spinning, so arrivals land within microseconds
// of their schedule), then waits for every request.
template <typename Submit>
int replay(sf_server *s, const double *arrival_s, size_t n, sf_server_result *results,
           Submit &&submit) {
  for (size_t i = 1; i < n; i++)
    if (arrival_s[i] < arrival_s[i - 1])
      return bad("arrivals out of order");
  std::vector<uint64_t> ids(n, 0);
  const Clock::time_point t0 = Clock::now();
  for (size_t i = 0; i < n; i++) {
    const Clock::time_point at =
        t0 + std::chrono::duration_cast<Clock::duration>(
                 std::chrono::duration<double>(arrival_s[i]));
    const Clock::time_point wake = at - std::chrono::microseconds(200);
    if (Clock::now() < wake)
      std::this_thread::sleep_until(wake);
    while (Clock::now() < at) {
    }
    const int rc = submit(i, &ids[i]);
    if (rc != 0)
      return rc;
  }
  int worst = 0;
  for (size_t i = 0; i < n; i++) {
    const int rc = sf_server_wait(s, ids[i], &results[i]);
    if (rc != 0 && worst == 0)
      worst = rc;
  }
  return worst;
}

} // namespace
} // extern "C++"

int sf_server_replay(sf_server *s, const int *sweeps, const void *const *frames,
                     const double *arrival_s, size_t n,
                     sf_server_result *results) {
  if (!s || (n && (!sweeps || !frames || !arrival_s || !results)))
    return bad("server or schedule");
  return replay(s, arrival_s, n, results, [&](size_t i, uint64_t *id) {
    return sf_server_submit(s, sweeps[i], frames[i], nullptr, 0, id);
  });
}

int sf_server_replay_chunks(sf_server *s, const int *sweeps, const sf_chunk *chunks,
                            const double *arrival_s, size_t n,
                            sf_server_result *results) {
  if (!s || (n && (!sweeps || !chunks || !arrival_s || !results)))
    return bad("server or schedule");
  return replay(s, arrival_s, n, results, [&](size_t i, uint64_t *id) {
    return sf_server_submit_chunk(s, sweeps[i], chunks[i].data, chunks[i].size,
                                  chunks[i].algorithm, nullptr, 0, id);
  });
}

} // extern "C"
