// Host memory copies on a pool of threads kept for the life of the process:
// copies into pinned staging, or into memory touched for the first time, run
// at several times one thread's rate when the bytes are spread over cores.
#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "host_copy.hh"

namespace sfi {

namespace {

class CopyPool {
public:
  static CopyPool &instance() {
    static CopyPool p;
    return p;
  }

  void copy(const CopyPiece *pieces, size_t n) {
    size_t total = 0;
    for (size_t i = 0; i < n; i++)
      total += pieces[i].bytes;
    const size_t parts = std::min<size_t>(threads_.size() + 1, total / kPerThread);
    if (parts <= 1) {
      for (size_t i = 0; i < n; i++)
        if (pieces[i].bytes)
          std::memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
      return;
    }
    std::lock_guard<std::mutex> one(job_); // one job at a time
    {
      std::lock_guard<std::mutex> lk(m_);
      pieces_ = pieces;
      npieces_ = n;
      total_ = total;
      parts_ = parts;
      next_part_ = 1; // part 0 is the caller's
      running_ = parts - 1;
      generation_++;
    }
    wake_.notify_all();
    run_part(0);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [&] { return running_ == 0; });
  }

private:
  static constexpr size_t kPerThread = size_t(2) << 20;

  CopyPool() {
    const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
    const unsigned n = std::min(7u, hw / 2);
    for (unsigned i = 0; i < n; i++)
      threads_.emplace_back([this] { serve(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    wake_.notify_all();
    for (std::thread &t : threads_)
      t.join();
  }

  // Bytes [part * total / parts, (part + 1) * total / parts) of the pieces
  // laid end to end.
  void run_part(size_t part) {
    const size_t lo = part * total_ / parts_, hi = (part + 1) * total_ / parts_;
    size_t at = 0;
    for (size_t i = 0; i < npieces_ && at < hi; i++) {
      const CopyPiece &p = pieces_[i];
      const size_t a = std::max(lo, at), b = std::min(hi, at + p.bytes);
      if (a < b)
        std::memcpy(static_cast<char *>(p.dst) + (a - at),
                    static_cast<const char *>(p.src) + (a - at), b - a);
      at += p.bytes;
    }
  }

  void serve() {
    uint64_t seen = 0;
    for (;;) {
      size_t part;
      {
        std::unique_lock<std::mutex> lk(m_);
        wake_.wait(lk, [&] { return stop_ || (generation_ != seen && next_part_ < parts_); });
        if (stop_)
          return;
        part = next_part_++;
        if (next_part_ >= parts_)
          seen = generation_;
      }
      run_part(part);
      std::lock_guard<std::mutex> lk(m_);
      if (--running_ == 0)
        done_.notify_all();
    }
  }

  std::mutex job_, m_;
  std::condition_variable wake_, done_;
  std::vector<std::thread> threads_;
  const CopyPiece *pieces_ = nullptr;
  size_t npieces_ = 0, total_ = 0, parts_ = 0, next_part_ = 0, running_ = 0;
  uint64_t generation_ = 0;
  bool stop_ = false;
};

} // namespace

void host_copy(const CopyPiece *pieces, size_t n) { CopyPool::instance().copy(pieces, n); }

} // namespace sfi
