// The serving loop behind sf_server_*: frames from several concurrent sweeps,
// taken as they arrive.
//
// Submitters queue requests {sweep, device frame, optional host list} stamped
// with the time of submission.  One worker thread owns the engine: it takes
// every request that has arrived -- at most ring - 1 per sweep, so that a
// sweep's previous frame keeps its slot, and at most max_batch in all -- into
// one batch, thresholds the batch in one persistent fused launch, labels it
// (each frame linked to its own sweep's previous frame), waits, copies the
// requested lists out and completes the requests.  Requests arriving while a
// batch is on the device wait for the next: continuous batching, with batch
// sizes following the load.  Each sweep holds a ring of slots of its shape
// for its whole life, so its buffers and label scratch stay resident.  A
// sweep is admitted only while the device's free memory stays above the
// configured reserve after its slots.
#include <algorithm>
#include <chrono>
#include <condition_variable>
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
  const void *frame;
  sf_signal_pixel *host_out;
  size_t host_capacity;
  Clock::time_point arrived;
};

struct Sweep {
  uint32_t height, width;
  int bits;
  int first_slot;     // slots first_slot .. first_slot + ring - 1
  uint64_t taken = 0; // frames taken into batches so far
  int last_slot = -1; // the slot of its last frame, once labelled
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
  Engine engine; // the worker's alone once it runs

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
};

void sf_server::run() {
  std::vector<Request> batch;
  std::vector<FrameIn> frames;
  std::vector<int> ids, prevs, status;
  std::vector<size_t> counts;
  std::vector<Outcome> outs;
  for (;;) {
    batch.clear();
    {
      std::unique_lock<std::mutex> lk(m);
      work.wait(lk, [&] { return stop || !queue.empty(); });
      if (queue.empty())
        return; // stopping
      // Everything that has arrived, in order, at most ring - 1 frames a
      // sweep: once a sweep reaches that, its later frames wait too, so each
      // sweep's frames stay in order.
      std::vector<uint32_t> per_sweep(sweeps.size(), 0);
      for (auto it = queue.begin();
           it != queue.end() && batch.size() < cfg.max_batch;) {
        uint32_t &k = per_sweep[static_cast<size_t>(it->sweep)];
        if (k + 1 >= cfg.ring) {
          ++it;
          continue;
        }
        k++;
        batch.push_back(*it);
        it = queue.erase(it);
      }
      const size_t n = batch.size();
      frames.resize(n);
      ids.resize(n);
      prevs.resize(n);
      for (size_t i = 0; i < n; i++) {
        Sweep &s = sweeps[static_cast<size_t>(batch[i].sweep)];
        const int slot =
            s.first_slot + static_cast<int>(s.taken % cfg.ring);
        s.taken++;
        frames[i] = FrameIn{batch[i].frame, s.height, s.width, s.bits};
        ids[i] = slot;
        prevs[i] = cfg.label ? s.last_slot : -1;
        s.last_slot = slot;
      }
    }

    const size_t n = batch.size();
    counts.assign(n, 0);
    status.assign(n, 0);
    outs.assign(n, Outcome{});
    std::string error;
    try {
      engine.threshold(frames.data(), ids.data(), n, SF_KERNEL_FUSED, false);
      if (cfg.label)
        engine.label(ids.data(), prevs.data(), n);
      engine.sync(ids.data(), n, counts.data(), status.data());
      for (size_t i = 0; i < n; i++) {
        sf_server_result &r = outs[i].result;
        r.signal_pixels = counts[i];
        r.status = status[i];
        if (cfg.label && status[i] == 0) {
          size_t np = 0, nc = 0, nl = 0;
          engine.label_counts(ids[i], &np, &nc, &nl, false);
          r.components = nc;
          r.links = nl;
        }
        if (batch[i].host_out && status[i] == 0)
          engine.fetch(ids[i], batch[i].host_out,
                       std::min(counts[i], batch[i].host_capacity));
      }
    } catch (const std::exception &e) {
      error = e.what();
    }
    const Clock::time_point done = Clock::now();

    {
      std::lock_guard<std::mutex> lk(m);
      for (size_t i = 0; i < n; i++) {
        Outcome &o = outs[i];
        o.result.request = batch[i].id;
        o.result.sweep = batch[i].sweep;
        o.result.latency_ms = ms_between(batch[i].arrived, done);
        o.result.batch_frames = static_cast<uint32_t>(n);
        if (!error.empty()) {
          o.result.status = -3;
          o.error = error;
        } else {
          latencies.push_back(o.result.latency_ms);
        }
        outcomes[batch[i].id] = o;
      }
      batches++;
      batch_frames += n;
      last_done = done;
    }
    finished.notify_all();
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
    s->sweeps.push_back(w);
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
  if (o.result.status == -3) {
    set_last_error(o.error);
    return -3;
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

int sf_server_replay(sf_server *s, const int *sweeps, const void *const *frames,
                     const double *arrival_s, size_t n,
                     sf_server_result *results) {
  if (!s || (n && (!sweeps || !frames || !arrival_s || !results)))
    return bad("server or schedule");
  for (size_t i = 1; i < n; i++)
    if (arrival_s[i] < arrival_s[i - 1])
      return bad("arrivals out of order");
  std::vector<uint64_t> ids(n, 0);
  const Clock::time_point t0 = Clock::now();
  for (size_t i = 0; i < n; i++) {
    const Clock::time_point at =
        t0 + std::chrono::duration_cast<Clock::duration>(
                 std::chrono::duration<double>(arrival_s[i]));
    // Sleep to within 200 us of the arrival, then spin, so that arrivals
    // land within microseconds of their schedule.
    const Clock::time_point wake = at - std::chrono::microseconds(200);
    if (Clock::now() < wake)
      std::this_thread::sleep_until(wake);
    while (Clock::now() < at) {
    }
    const int rc = sf_server_submit(s, sweeps[i], frames[i], nullptr, 0, &ids[i]);
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

} // extern "C"
