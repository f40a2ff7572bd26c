// Device-to-host copies into pageable host memory (host_copy.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace ibfs {

// One contiguous device range and its host destination.
struct D2HPiece {
  void *dst;
  const void *src;
  size_t bytes;
};

// Copies the pieces from device memory to pageable host memory in stream
// order on `st` (after the work already enqueued there) and returns when
// they are on the host.  Large copies stage through two pinned chunks: the
// DMA engine fills one while host threads copy the other out.  Returns
// IBFS_OK or an error code with ibfs_last_error set; waits end at the
// library's wait limit.
int copy_to_host(const D2HPiece *pieces, size_t n, hipStream_t st);

}  // namespace ibfs
