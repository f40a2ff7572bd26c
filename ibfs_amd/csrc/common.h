// Internal declarations shared by the host sources of libibfs.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ibfs.h"

namespace ibfs {

// Thread-local message of the last failure, returned by ibfs_last_error().
void set_error(const char *fmt, ...);
const std::string &last_error();

// Returns `code` after recording a message.
int fail(int code, const char *fmt, ...);

struct HostCsr {
  int64_t V = 0;
  int64_t E = 0;
  std::vector<int64_t> beg;  // V+1
  std::vector<int32_t> adj;  // E
};

// Parallel helper: runs fn(t) on nthreads std::threads.
void parallel_for_threads(int nthreads, void (*fn)(int, void *), void *ctx);

}  // namespace ibfs

struct ibfs_hostcsr_s {
  ibfs::HostCsr csr;
};

// Device-resident graph (read-only after upload).
struct ibfs_graph_s {
  int device = 0;
  int64_t V = 0, E = 0;
  int64_t *d_beg = nullptr;
  int32_t *d_adj = nullptr;   // adjacency as given (top-down, GroupBy)
  int32_t *d_tadj = nullptr;  // traversal adjacency: by descending degree, hubs encoded
  int32_t *d_deg = nullptr;
  int32_t *d_hubs = nullptr;  // the H highest-degree vertices
  int64_t H = 0;
  int64_t max_deg = 0;
};
