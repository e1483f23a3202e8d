// Host memory copies spread over a pool of host threads.  Not part of the
// public ABI.
#pragma once

#include <cstddef>

namespace sfi {

// Copies pieces laid end to end, spread over a pool of host threads kept for
// the process when they total a few MB or more (copies into pinned staging,
// or into memory touched for the first time, are limited by one core's rate
// otherwise); smaller totals are copied on the calling thread.
struct CopyPiece {
  void *dst;
  const void *src;
  size_t bytes;
};
void host_copy(const CopyPiece *pieces, size_t n);

} // namespace sfi
