// Internal declarations shared by the host sources of libibfs.
#pragma once

#include <algorithm>
#include <cmath>
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

// inspector chunking override of ibfs_config (0 = by vertex count)
inline bool valid_inspect_vpt(int v) {
  return v == 0 || v == 1 || v == 2 || v == 4 || v == 8 || v == 16;
}
// inspector workgroup size of ibfs_config
inline bool valid_wg_size(int v) { return v == 256 || v == 512 || v == 1024; }

// Direction settings of a config: policy, alpha/beta, hybrid_gamma.
inline int check_direction_config(const ibfs_config &c) {
  if (c.policy < IBFS_POLICY_FIXED || c.policy > IBFS_POLICY_HYBRID)
    return fail(IBFS_ERR_ARG, "unknown policy %d", c.policy);
  if (c.policy == IBFS_POLICY_ALPHA_BETA && (c.alpha < 1 || c.beta < 1))
    return fail(IBFS_ERR_ARG, "alpha and beta must be >= 1");
  if (c.hybrid_gamma < 0 || c.hybrid_gamma > 4096)
    return fail(IBFS_ERR_ARG, "hybrid_gamma must be in [0, 4096] (got %d)", c.hybrid_gamma);
  if (c.hybrid_hubs < 0) return fail(IBFS_ERR_ARG, "hybrid_hubs must be >= 0");
  if (c.split_cov < 0 || c.split_cov > 100)
    return fail(IBFS_ERR_ARG, "split_cov must be in [0, 100] (got %d)", c.split_cov);
  if (c.phase_offsets < 0 || c.phase_offsets > 3)
    return fail(IBFS_ERR_ARG, "phase_offsets must be in [0, 3] (got %d)", c.phase_offsets);
  if (c.tail_div < 0) return fail(IBFS_ERR_ARG, "tail_div must be >= 0 (got %d)", c.tail_div);
  return IBFS_OK;
}

// Start level of each lane of a group from its source's two-hop size
// (phase offsets p = 1..3, ibfs_config::phase_offsets): a lane whose key is
// below the group's median by f^lag, with f = d^(p/2) the frontier growth
// per level taken for a graph of average degree d (lag rounded, at most 2),
// is `lag` levels behind the median lanes and starts max_lag - lag levels
// earlier than they do; p = 0 or fewer than two lanes: every lane at 0.
inline void phase_starts(const std::vector<double> &two_hop, double avg_deg, int p,
                         std::vector<int> &start) {
  const size_t k = two_hop.size();
  start.assign(k, 0);
  if (p <= 0 || k < 2) return;
  std::vector<double> lk(k);
  for (size_t i = 0; i < k; ++i) lk[i] = std::log(two_hop[i] + 1.0);
  std::vector<double> srt = lk;
  std::nth_element(srt.begin(), srt.begin() + (long)(k / 2), srt.end());
  const double med = srt[k / 2];
  const double ld = std::log(avg_deg > 2.0 ? avg_deg : 2.0) * 0.5 * p;
  int lag_max = 0;
  std::vector<int> lag(k);
  for (size_t i = 0; i < k; ++i) {
    const int l = (int)std::floor((med - lk[i]) / ld + 0.5);
    lag[i] = l < 0 ? 0 : (l > 2 ? 2 : l);
    lag_max = lag[i] > lag_max ? lag[i] : lag_max;
  }
  for (size_t i = 0; i < k; ++i) start[i] = lag_max - lag[i];
}

// Parallel helper: runs fn(t) on nthreads std::threads.
void parallel_for_threads(int nthreads, void (*fn)(int, void *), void *ctx);

// Limit of every host wait of the library on device work (seconds,
// process-wide, ibfs_set_wait_timeout; default 120).  A wait that exceeds it
// returns IBFS_ERR_STATE instead of spinning on a stream that never drains.
double wait_timeout_s();

// Deadline of one wait, started at construction.
class WaitClock {
 public:
  WaitClock();
  // true once the limit has passed (reads the clock every 16 calls)
  bool expired();
  void restart();

 private:
  double t0_;
  unsigned n_ = 0;
};

}  // namespace ibfs

struct ibfs_hostcsr_s {
  ibfs::HostCsr csr;
};

// Device-resident graph (read-only after upload).
//
// Traversal vertex space: the Vt vertices with at least one edge (out or
// in), numbered in their original order (cid[v] = traversal id of original
// vertex v, -1 when v is isolated).  Status words, queues and the CSR below
// use traversal ids; an isolated vertex can only ever be reached as its own
// source, so it needs no status word.  Depth arrays and the ABI use original
// ids (V of them).  An isolated vertex has no out-edges, so the CSR offsets
// of the traversal vertices are the original offsets.
struct ibfs_graph_s {
  int device = 0;
  int64_t V = 0, E = 0;
  int64_t Vt = 0;
  int64_t *d_beg = nullptr;   // Vt+1, traversal ids
  int32_t *d_adj = nullptr;   // adjacency in input order (top-down, GroupBy)
  int32_t *d_tadj = nullptr;  // traversal adjacency: by descending degree, hubs encoded
  int32_t *d_deg = nullptr;   // Vt
  int32_t *d_hubs = nullptr;  // the H highest-degree vertices
  int32_t *d_cid = nullptr;   // V: original id -> traversal id or -1
  int32_t *d_vid = nullptr;   // Vt: traversal id -> original id
  std::vector<int32_t> h_cid;
  int64_t H = 0;
  int64_t max_deg = 0;
  // per hub slot (descending degree, ties by id): degree and traversal id;
  // hub_entries[i] = traversal-adjacency entries that point to slots < i
  std::vector<int32_t> h_hub_deg, h_hub_id;
  std::vector<int64_t> h_hub_entries;
};

// Pull cut of hybrid levels at the top r hubs (r <= 0 or r > H: all H hubs):
// the degree and id of the last pulled hub (deg INT32_MAX, tie -1 when there
// is none), the number of pulled hubs, and the fraction of adjacency entries
// that point to them (x1024).
struct HubCut {
  int32_t deg = INT32_MAX, tie = -1, cut = 0, rho_q10 = 0;
};
HubCut hub_cut(const ibfs_graph_s *g, int64_t r);
// Hubs the split rule of device-driven levels samples: the pulled hubs, or
// the top 4,096 when no hub is pulled, at most the hub cache; 0 when split
// levels are off (the decision kernel then runs as one wave).
inline int split_hubs(const ibfs_graph_s *g, const HubCut &hc, const ibfs_config &c) {
  if (c.split_cov <= 0) return 0;
  const int64_t r = hc.cut > 0 ? hc.cut : 4096;
  return (int)(r < g->H ? r : g->H);
}
