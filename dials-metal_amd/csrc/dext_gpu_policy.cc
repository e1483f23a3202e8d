// The backend-independent part of namespace gpu, for a build of this library
// on its own: the window choice, the fused switch, the profiling and
// reporting switches, the per-thread stage times and the stderr line that
// names the configuration when it changes.  In a build inside the reference
// the reference's own src/spots/dext_gpu.cc takes this file's place and
// dext_gpu_hip.cc links against that instead; both meet it through
// dext_gpu_internal.hh.
//
// Interface kept from the reference (dext_gpu.hh:50-134, dext_gpu.cc:56-96,
// 146-166): the variables SPOTFINDER_GPU_STAGE0 / _STAGE2 ("tile" or
// "direct") and SPOTFINDER_GPU_FUSED ("1" or "0"), read once, before the first
// read or write of a choice, so that an explicit setter call is never undone
// by the environment; a value that is neither word is a warning and leaves the
// default; the announcement text, printed only while reporting is on and only
// when the configuration differs from the last one announced.
//
// Design: the whole configuration is one atomic word of flag bits, built once
// by a function-local static from the backend's defaults and a table of
// environment switches.  Getters and setters are single atomic loads and
// read-modify-writes on that word, so a switch in one thread is seen by the
// next find() in any thread, and nothing here takes a lock after start-up.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "dext_gpu_internal.hh"

namespace gpu {
namespace {

enum Flag : unsigned {
  StageZeroTile = 1u << 0, // stage 0 sums its window from a tile
  StageTwoTile = 1u << 1,  // stage 2 likewise
  Fused = 1u << 2,         // the three stages as one kernel
  Profiling = 1u << 3,     // find() records the stage split
  Quiet = 1u << 4,         // report_windows(false)
};

// One environment switch: the variable, the flag it drives and the words
// that set and clear it.  The warning names the two words in this order.
struct Switch {
  const char *variable;
  Flag flag;
  const char *set_word;
  const char *clear_word;
};

constexpr Switch kSwitches[] = {
    {"SPOTFINDER_GPU_STAGE0", StageZeroTile, "tile", "direct"},
    {"SPOTFINDER_GPU_STAGE2", StageTwoTile, "tile", "direct"},
    {"SPOTFINDER_GPU_FUSED", Fused, "1", "0"},
};

unsigned apply_environment(unsigned word) {
  for (const Switch &s : kSwitches) {
    const char *value = std::getenv(s.variable);
    if (value == nullptr)
      continue;
    if (std::strcmp(value, s.set_word) == 0)
      word |= s.flag;
    else if (std::strcmp(value, s.clear_word) == 0)
      word &= ~static_cast<unsigned>(s.flag);
    else
      std::fprintf(stderr,
                   "warning: %s=%s is not '%s' or '%s'; leaving it alone\n",
                   s.variable, value, s.set_word, s.clear_word);
  }
  return word;
}

// Configurations as the announcement sees them: the window pair in the low
// two bits, or the fused kernel; Shown marks a value as announced, so that
// "nothing yet" (zero) differs from every configuration.
constexpr unsigned kShown = 1u << 8;
constexpr unsigned kFusedConfig = 1u << 2;

class Policy {
public:
  static Policy &instance() {
    static Policy policy; // initialised once, thread-safely, on first use
    return policy;
  }

  bool has(Flag f) const {
    return (word_.load(std::memory_order_relaxed) & f) != 0;
  }
  void put(Flag f, bool on) {
    if (on)
      word_.fetch_or(f, std::memory_order_relaxed);
    else
      word_.fetch_and(~static_cast<unsigned>(f), std::memory_order_relaxed);
  }

  // True when `config` should be announced now: reporting is on and it is
  // not the configuration announced last.  Two threads switching at the same
  // moment may both announce, which repeats a line but never misstates one.
  bool should_announce(unsigned config) {
    if (has(Quiet))
      return false;
    const unsigned mark = config | kShown;
    return last_shown_.exchange(mark, std::memory_order_relaxed) != mark;
  }

private:
  Policy() {
    unsigned word = 0;
    if (internal::default_stage0_window() == Window::Tile)
      word |= StageZeroTile;
    if (internal::default_stage2_window() == Window::Tile)
      word |= StageTwoTile;
    word_.store(apply_environment(word), std::memory_order_relaxed);
  }

  std::atomic<unsigned> word_{0};
  std::atomic<unsigned> last_shown_{0};
};

Window window_of(Flag f) {
  return Policy::instance().has(f) ? Window::Tile : Window::Direct;
}

thread_local StageTimes t_last_split;

} // namespace

const char *name(Window window) {
  return window == Window::Tile ? "tile" : "direct";
}

Window stage0_window() { return window_of(StageZeroTile); }
Window stage2_window() { return window_of(StageTwoTile); }
void stage0_window(Window w) {
  Policy::instance().put(StageZeroTile, w == Window::Tile);
}
void stage2_window(Window w) {
  Policy::instance().put(StageTwoTile, w == Window::Tile);
}
bool fused() { return Policy::instance().has(Fused); }
void fused(bool on) { Policy::instance().put(Fused, on); }

void profile_stages(bool on) { Policy::instance().put(Profiling, on); }
bool profile_stages() { return Policy::instance().has(Profiling); }
StageTimes last_stage_times() { return t_last_split; }
void report_windows(bool on) { Policy::instance().put(Quiet, !on); }

namespace internal {

void set_stage_times(const StageTimes &times) { t_last_split = times; }

void announce_windows(Window stage0, Window stage2) {
  const unsigned config = (stage0 == Window::Tile ? 2u : 0u) |
                          (stage2 == Window::Tile ? 1u : 0u);
  if (Policy::instance().should_announce(config))
    std::fprintf(stderr, "%s: stage0 %s window, stage2 %s window\n", backend(),
                 name(stage0), name(stage2));
}

void announce_fused() {
  if (Policy::instance().should_announce(kFusedConfig))
    std::fprintf(stderr, "%s: the three stages fused, one kernel\n", backend());
}

} // namespace internal
} // namespace gpu
