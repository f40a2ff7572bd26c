// Batched launcher: an arbitrary source set packed into lane groups, several
// traversal contexts in flight on their own HIP streams, each level a
// captured hipGraph of three kernels (expand, inspect, direction decision):
// the direction is decided on the device, so the host only polls once per
// burst of levels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "common.h"
#include "host_copy.h"
#include "kernels.h"

using namespace ibfs;

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(IBFS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                  __FILE__, __LINE__);                                                 \
  } while (0)

namespace {

enum CtxState { IDLE = 0, RUNNING = 1 };

struct Ctx {
  hipStream_t st = nullptr;
  u64 *ring = nullptr;
  LevelCtl *ctl = nullptr;
  Queues q{};
  u64 *hs = nullptr;
  uint8_t *flags = nullptr;  // changed-at-this-level byte per traversal vertex
  // a group's set-up data (level masks, sources, lane table, plan) lives in
  // one pinned block and one device block of the same layout, so a group
  // starts with one copy; the pointers below point into them
  char *d_setup = nullptr, *h_setup = nullptr;
  size_t setup_bytes = 0;
  Plan *d_plan = nullptr, *h_plan = nullptr;
  u64 *d_masks = nullptr, *h_masks = nullptr;  // [0,W) lanes_all, [W,2W) sel, then (R+1)W level masks
  int32_t *d_src = nullptr, *h_src = nullptr;
  LaneOut *d_lanes = nullptr, *h_lanes = nullptr;
  int32_t *d_stage = nullptr;  // host-placed output staging, K*V
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // a full burst's levels in one graph: consecutive graph launches leave
  // ≈8 µs between one level's last kernel and the next level's first
  hipGraph_t graph_burst = nullptr;
  hipGraphExec_t exec_burst = nullptr;
  hipEvent_t ev = nullptr;
  int state = IDLE;
  int64_t g0 = 0;
  int k = 0;
  int base = 0;
  int level = 0;  // level the device will expand next (host view)
  int32_t *out_host = nullptr;    // host-placed output (rows of the caller's order)
  const int64_t *rows = nullptr;  // caller's index of each lane of the group
  // lanes that start after level 0 (phase offsets): each lane's start level,
  // the latest one, and the admission list staged for one level
  std::vector<int> start;
  int a_max = 0;
  AdmitList *d_al = nullptr, *h_al = nullptr;
  // sparse tail log, and the group's arguments for a rerun without the tail
  // after the log overflowed
  TailLog tl{};
  u64 tail_cap = 0;
  bool no_tail = false;
  const int32_t *src = nullptr;
  int32_t *dev_out = nullptr, *stage = nullptr;
};

template <class T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}
template <class T>
void hfree(T *&p) {
  if (p) (void)hipHostFree((void *)p);
  p = nullptr;
}

}  // namespace

struct ibfs_batch_s {
  ibfs_graph_s *g = nullptr;
  ibfs_config cfg{};
  int K = 128, W = 2, R = 16, C = 2, burst = 8, chunk_shift = 12, grid = 2048;
  HubCut hc;  // hubs a hybrid level pulls through
  std::vector<Ctx> ctx;
  // grouping order of the current run: order[p] = caller's index of the
  // source at position p (groups are consecutive positions); psrc = sources
  // in that order; two-hop keys staged on the device
  std::vector<int64_t> order;
  std::vector<int32_t> psrc;
  std::vector<u64> h_keys;
  int32_t *d_keysrc = nullptr;
  u64 *d_keys = nullptr;
  int64_t key_cap = 0;
  u64 *d_edges = nullptr, *h_edges = nullptr;
  // closeness runs: per-source depth sums and reach counts instead of depths
  bool closeness = false;
  // a run failed with work possibly still enqueued on the contexts' streams
  // (a wait past the limit): the next run drains them first, since its
  // set-up rewrites the pinned blocks that work may still copy into
  bool unsettled = false;
  u64 *d_sum = nullptr, *d_reach = nullptr;
  int64_t sum_cap = 0;
};

namespace {

size_t masks_words(const ibfs_batch_s *b) { return (size_t)(2 + (b->R + 1)) * b->W; }

u64 *slot(const ibfs_batch_s *b, const Ctx &c, int L) {
  return c.ring + (size_t)(L % b->R) * (size_t)b->g->Vt * (size_t)b->W;
}

void release_ctx(Ctx &c) {
  if (c.st) (void)hipStreamSynchronize(c.st);
  if (c.exec) (void)hipGraphExecDestroy(c.exec);
  if (c.graph) (void)hipGraphDestroy(c.graph);
  if (c.exec_burst) (void)hipGraphExecDestroy(c.exec_burst);
  if (c.graph_burst) (void)hipGraphDestroy(c.graph_burst);
  c.exec = c.exec_burst = nullptr;
  c.graph = c.graph_burst = nullptr;
  dfree(c.ring);
  dfree(c.ctl);
  for (int i = 0; i < 6; ++i) dfree(c.q.q[i]);
  dfree(c.q.chunks);
  dfree(c.hs);
  dfree(c.flags);
  dfree(c.d_setup);
  hfree(c.h_setup);
  c.d_plan = c.h_plan = nullptr;
  c.d_masks = c.h_masks = nullptr;
  c.d_src = c.h_src = nullptr;
  c.d_lanes = c.h_lanes = nullptr;
  dfree(c.d_stage);
  dfree(c.d_al);
  hfree(c.h_al);
  dfree(c.tl.v);
  dfree(c.tl.lev);
  dfree(c.tl.bits);
  if (c.ev) (void)hipEventDestroy(c.ev);
  c.ev = nullptr;
  if (c.st) (void)hipStreamDestroy(c.st);
  c.st = nullptr;
}

void release_batch(ibfs_batch_s *b) {
  if (b->g) (void)hipSetDevice(b->g->device);
  for (auto &c : b->ctx) release_ctx(c);
  b->ctx.clear();
  dfree(b->d_edges);
  dfree(b->d_keysrc);
  dfree(b->d_keys);
  hfree(b->h_edges);
  dfree(b->d_sum);
  dfree(b->d_reach);
  b->sum_cap = 0;
}

// Output of the snapshot window [lo, hi] of context c: depth arrays, or the
// closeness sums of the group's sources.
// edges (a final depth window only, or null): the emission also adds the
// TEPS numerator.
hipError_t window_out(ibfs_batch_s *b, Ctx &c, int lo, int hi, u64 *edges = nullptr) {
  const EmitSpace sp{b->g->d_cid, b->g->V, b->g->Vt, b->g->d_deg};
  if (b->closeness)
    return launch_close(b->W, c.ring, b->R, sp, lo, hi, c.d_lanes, c.d_masks + 2 * b->W,
                        c.d_masks + b->W, c.k, b->d_sum + c.g0, b->d_reach + c.g0, c.st);
  return launch_emit(b->W, b->cfg.emit_mode, 0, c.ring, b->R, sp, lo, hi, c.d_lanes,
                     c.d_masks + 2 * b->W, c.d_masks + b->W, c.k, edges, c.st);
}

InspectArgs dd_inspect_args(const ibfs_batch_s *b, const Ctx &c) {
  InspectArgs a{};
  a.deg = b->g->d_deg;
  a.V = b->g->Vt;
  a.q = c.q;
  a.sml = b->cfg.sml_shed;
  a.lrg = b->cfg.lrg_shed;
  a.chunk_shift = b->chunk_shift;
  a.hubs = b->g->d_hubs;
  a.H = b->g->H;
  a.hs = c.hs;
  a.plan = c.d_plan;
  a.ring = c.ring;
  a.R = b->R;
  a.ctl_ring = c.ctl;
  a.lanes_all = c.d_masks;
  a.flags = c.flags;
  a.vpt = b->cfg.inspect_vpt;
  a.nw = b->cfg.wg_size / 64;
  a.hsplit = HubSplit{b->hc.deg, b->hc.tie, b->hc.cut};
  return a;
}

int alloc_ctx(ibfs_batch_s *b, Ctx &c) {
  const int64_t Vt = b->g->Vt, Vq = std::max<int64_t>(Vt, 1);
  const size_t slot_bytes = (size_t)Vq * (size_t)b->W * sizeof(u64);
  hipError_t err = hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking);
  auto A = [&](void **p, size_t bytes) {
    if (err == hipSuccess) err = hipMalloc(p, bytes ? bytes : 1);
  };
  auto H = [&](void **p, size_t bytes) {
    if (err == hipSuccess) err = hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault);
  };
  A((void **)&c.ring, (size_t)b->R * slot_bytes);
  A((void **)&c.ctl, sizeof(LevelCtl) * (size_t)b->R);
  for (int i = 0; i < 6; ++i)
    if (i != Q_TD_L) A((void **)&c.q.q[i], sizeof(int32_t) * (size_t)Vq);
  A((void **)&c.q.chunks, sizeof(int2) * ((size_t)Vq + (size_t)(b->g->E >> b->chunk_shift) + 2));
  A((void **)&c.hs, sizeof(u64) * (size_t)b->W * (size_t)std::max<int64_t>(b->g->H, 1));
  A((void **)&c.flags, sizeof(u64) * (size_t)((Vq + 7) / 8));
  {  // set-up block: masks, sources, lane table, plan (256-byte aligned parts)
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_src = al(sizeof(u64) * masks_words(b));
    const size_t o_lanes = o_src + al(sizeof(int32_t) * (size_t)b->K);
    const size_t o_plan = o_lanes + al(sizeof(LaneOut) * (size_t)b->K);
    c.setup_bytes = o_plan + al(sizeof(Plan));
    A((void **)&c.d_setup, c.setup_bytes);
    H((void **)&c.h_setup, c.setup_bytes);
    if (err == hipSuccess) {
      c.d_masks = reinterpret_cast<u64 *>(c.d_setup);
      c.h_masks = reinterpret_cast<u64 *>(c.h_setup);
      c.d_src = reinterpret_cast<int32_t *>(c.d_setup + o_src);
      c.h_src = reinterpret_cast<int32_t *>(c.h_setup + o_src);
      c.d_lanes = reinterpret_cast<LaneOut *>(c.d_setup + o_lanes);
      c.h_lanes = reinterpret_cast<LaneOut *>(c.h_setup + o_lanes);
      c.d_plan = reinterpret_cast<Plan *>(c.d_setup + o_plan);
      c.h_plan = reinterpret_cast<Plan *>(c.h_setup + o_plan);
      std::memset(c.h_setup, 0, c.setup_bytes);
    }
  }
  A((void **)&c.d_al, sizeof(AdmitList));
  H((void **)&c.h_al, sizeof(AdmitList));
  if (b->cfg.tail_div > 0) {
    c.tail_cap = (u64)Vq + 65536;
    A((void **)&c.tl.v, sizeof(int32_t) * (size_t)c.tail_cap);
    A((void **)&c.tl.lev, sizeof(int32_t) * (size_t)c.tail_cap);
    A((void **)&c.tl.bits, sizeof(u64) * (size_t)b->W * (size_t)c.tail_cap);
  }
  if (err == hipSuccess) err = hipEventCreateWithFlags(&c.ev, hipEventDisableTiming);
  if (err != hipSuccess)
    return fail(err == hipErrorOutOfMemory ? IBFS_ERR_NOMEM : IBFS_ERR_HIP,
                "batch context allocation failed: %s", hipGetErrorString(err));
  // capture the level body (device-driven expansion + inspection + plan),
  // once and, for full bursts, `burst` times in a row: every kernel reads
  // its level from the plan, so repeated bodies run consecutive levels
  for (int pass = 0; pass < 2; ++pass) {
    const int reps = pass == 0 ? 1 : b->burst;
    if (pass == 1 && reps < 2) break;  // bursts of one level need no second graph
    clear_hip_error();
    HIP_TRY(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = hipSuccess, e2 = hipSuccess;
    for (int r = 0; r < reps && e1 == hipSuccess && e2 == hipSuccess; ++r) {
      e1 = launch_expand_dd(b->W, b->g->d_beg, b->g->d_adj, b->g->d_tadj, c.ring, c.flags, b->R,
                            Vt, c.ctl, c.d_plan, c.hs, c.q, b->chunk_shift,
                            HubSplit{b->hc.deg, b->hc.tie, b->hc.cut}, c.tl, b->grid, c.st);
      e2 = launch_inspect(b->W, dd_inspect_args(b, c), c.st);
      if (e2 == hipSuccess)
        e2 = launch_plan_decide(c.d_plan, c.ctl, b->R, c.hs, b->W,
                                split_hubs(b->g, b->hc, b->cfg), c.st);
    }
    hipGraph_t &gr = pass == 0 ? c.graph : c.graph_burst;
    hipError_t e3 = hipStreamEndCapture(c.st, &gr);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
      return fail(IBFS_ERR_HIP, "level graph capture failed: %s / %s / %s",
                  hipGetErrorString(e1), hipGetErrorString(e2), hipGetErrorString(e3));
    HIP_TRY(hipGraphInstantiate(pass == 0 ? &c.exec : &c.exec_burst, gr, nullptr, nullptr, 0));
  }
  return IBFS_OK;
}

// Starts group [g0, g0+k) of src on context c: lane table, masks, S_0 and the
// level-0 inspection (whose last workgroup plans level 1).
// src: sources in grouping order; rows[i] = caller's index of lane i (its
// depth row in dev_out); stage: host-placed output staging (lane i at row i).
// Start level of each lane of a group (phase offsets, the rule of
// ibfs::phase_starts over the sources' two-hop sizes).
void group_starts(const ibfs_batch_s *b, const int64_t *rows, int k, std::vector<int> &start) {
  if (!b->cfg.phase_offsets || b->h_keys.empty()) {
    start.assign((size_t)k, 0);
    return;
  }
  std::vector<double> key((size_t)k);
  for (int i = 0; i < k; ++i) key[(size_t)i] = (double)b->h_keys[(size_t)rows[i]];
  const double avg_deg = b->g->Vt > 0 ? (double)b->g->E / (double)b->g->Vt : 2.0;
  phase_starts(key, avg_deg, b->cfg.phase_offsets, start);
}

int start_group(ibfs_batch_s *b, Ctx &c, const int32_t *src, int64_t g0, int k, int32_t *dev_out,
                int32_t *stage, const int64_t *rows) {
  const int W = b->W;
  const int64_t V = b->g->V;
  group_starts(b, rows, k, c.start);
  c.a_max = 0;
  for (int i = 0; i < k; ++i) c.a_max = std::max(c.a_max, c.start[(size_t)i]);
  std::memset(c.h_masks, 0, sizeof(u64) * masks_words(b));
  for (int i = 0; i < k; ++i) {
    const int a = c.start[(size_t)i];
    if (a == 0) c.h_masks[i >> 6] |= 1ull << (i & 63);  // lanes of level 0
    c.h_masks[W + (i >> 6)] |= 1ull << (i & 63);         // every lane of the group
    // lanes starting later are set by the admission of their level
    c.h_src[i] = a == 0 ? b->g->h_cid[src[g0 + i]] : -1;
    c.h_lanes[i].out = dev_out ? dev_out + (size_t)rows[i] * (size_t)V
                               : (stage ? stage + (size_t)i * (size_t)V : nullptr);
    c.h_lanes[i].admit = a;
    c.h_lanes[i].src = src[g0 + i];
  }
  c.rows = rows;
  // level masks of the first window (base 0): lanes started at or before L
  for (int L = 0; L <= b->R; ++L)
    for (int w = 0; w < W; ++w) {
      u64 m = ~0ull;
      if (L < c.a_max) {
        m = 0;
        for (int j = 0; j < 64 && w * 64 + j < k; ++j)
          if (c.start[(size_t)(w * 64 + j)] <= L) m |= 1ull << j;
      }
      c.h_masks[2 * W + (size_t)L * W + w] = m;
    }
  plan_init(c.h_plan, b->cfg, b->hc.rho_q10, (u64)V);
  c.h_plan->hold = c.a_max;
  if (c.tail_cap > 0 && !c.no_tail) {
    c.h_plan->tail_div = b->cfg.tail_div;
    c.h_plan->tail_cap = c.tail_cap;
    c.h_plan->E = (u64)b->g->E;
  }
  // masks, sources, lane table and plan in one copy (lane entries past k are
  // stale and never read: every kernel takes the group's k); then S_0, the
  // flags and the control blocks zeroed in one launch
  HIP_TRY(hipMemcpyAsync(c.d_setup, c.h_setup, c.setup_bytes, hipMemcpyHostToDevice, c.st));
  HIP_TRY(launch_zero3_u64(slot(b, c, 0), b->g->Vt * W, reinterpret_cast<u64 *>(c.flags),
                           (b->g->Vt + 7) / 8, reinterpret_cast<u64 *>(c.ctl),
                           (int64_t)b->R * CTL_WORDS, c.st));
  HIP_TRY(launch_init_sources(W, c.d_src, k, 0, slot(b, c, 0), c.flags, c.st));
  HIP_TRY(launch_inspect(W, dd_inspect_args(b, c), c.st));
  HIP_TRY(launch_plan_decide(c.d_plan, c.ctl, b->R, c.hs, W, split_hubs(b->g, b->hc, b->cfg), c.st));
  c.g0 = g0;
  c.k = k;
  c.base = 0;
  c.level = 1;
  c.src = src;
  c.dev_out = dev_out;
  c.stage = stage;
  return IBFS_OK;
}

// Enqueues up to `burst` level bodies (bounded so that no snapshot of the
// current emission window is overwritten), then a copy of the plan and the
// context's event.
int launch_burst(ibfs_batch_s *b, Ctx &c, ibfs_stats *stats) {
  // sparse tail levels write no new snapshot: no window to flush
  int B = c.h_plan->tail_start ? b->burst : std::min(b->burst, b->R - 1 - (c.level - c.base));
  if (B < 1) {  // flush the window of complete snapshots [base, level-1]
    HIP_TRY(window_out(b, c, c.base, c.level - 1));
    if (stats) stats->emit_windows++;
    c.base = c.level - 1;
    B = std::min(b->burst, b->R - 1 - (c.level - c.base));
    if (c.a_max > 0) {  // every lane has started before the new base
      const int W = b->W;
      for (size_t i = 0; i < (size_t)(b->R + 1) * W; ++i) c.h_masks[2 * W + i] = ~0ull;
      HIP_TRY(hipMemcpyAsync(c.d_masks + 2 * W, c.h_masks + 2 * W,
                             sizeof(u64) * (size_t)(b->R + 1) * W, hipMemcpyHostToDevice, c.st));
    }
  }
  if (c.level <= c.a_max) {
    // one level at a time while lanes still start: those of this level join
    // between its expansion and its inspection
    int n = 0;
    for (int i = 0; i < c.k; ++i)
      if (c.start[(size_t)i] == c.level) {
        c.h_al->src[n] = b->g->h_cid[c.h_lanes[i].src];
        c.h_al->lane[n] = i;
        ++n;
      }
    if (n > 0) {
      c.h_al->level = c.level;
      c.h_al->n = n;
      HIP_TRY(hipMemcpyAsync(c.d_al, c.h_al, sizeof(AdmitList), hipMemcpyHostToDevice, c.st));
      const HubSplit hsp{b->hc.deg, b->hc.tie, b->hc.cut};
      HIP_TRY(launch_expand_dd(b->W, b->g->d_beg, b->g->d_adj, b->g->d_tadj, c.ring, c.flags,
                               b->R, b->g->Vt, c.ctl, c.d_plan, c.hs, c.q, b->chunk_shift, hsp,
                               c.tl, b->grid, c.st));
      HIP_TRY(launch_admit_dd(b->W, c.d_al, c.d_plan, c.ring, b->R, b->g->Vt, c.ctl, c.flags,
                              c.st));
      HIP_TRY(launch_inspect(b->W, dd_inspect_args(b, c), c.st));
      HIP_TRY(launch_plan_decide(c.d_plan, c.ctl, b->R, c.hs, b->W,
                                 split_hubs(b->g, b->hc, b->cfg), c.st));
    } else {
      HIP_TRY(hipGraphLaunch(c.exec, c.st));
    }
    B = 0;
  }
  if (B == b->burst && c.exec_burst) HIP_TRY(hipGraphLaunch(c.exec_burst, c.st));
  else
    for (int i = 0; i < B; ++i) HIP_TRY(hipGraphLaunch(c.exec, c.st));
  HIP_TRY(hipMemcpyAsync(c.h_plan, c.d_plan, sizeof(Plan), hipMemcpyDeviceToHost, c.st));
  HIP_TRY(hipEventRecord(c.ev, c.st));
  return IBFS_OK;
}

int finish_group(ibfs_batch_s *b, Ctx &c, int last, ibfs_stats *stats) {
  // levels of a sparse tail after its first level T share T's slot, which
  // holds the final words: the window ends at T and the tail's log then
  // corrects the depths of the bits that arrived after T
  const int T = c.h_plan->tail_start;
  const int e = (T > 0 && last > T) ? T : last;
  // a final depth window counts the TEPS numerator during its pass
  const bool fused = stats && !b->closeness && e >= c.base;
  if (e >= c.base) {
    HIP_TRY(window_out(b, c, c.base, e, fused ? b->d_edges : nullptr));
    if (stats) stats->emit_windows++;
  }
  if (e < last)
    HIP_TRY(launch_tail_patch(b->W, c.tl, (int64_t)std::min(c.h_plan->tail_count, c.tail_cap), T,
                              b->g->d_vid, c.d_lanes, c.d_masks + b->W, c.k,
                              b->closeness ? b->d_sum + c.g0 : nullptr, c.st));
  if (stats) {
    stats->levels = std::max(stats->levels, last + 1);
    // one device-driven expansion per level, by the direction the plan logged
    const int dense_last = T > 0 ? std::min(last, T - 1) : last;
    for (int l = 1; l <= dense_last && l < 256; ++l) {
      const int d = c.h_plan->dirlog[l];
      (d == 0 ? stats->n_td_launches : d == 1 ? stats->n_bu_launches
               : d == 2 ? stats->n_hy_launches : stats->n_split_levels)++;
    }
    // tail expansions: T..last and the one that found nothing
    if (T > 0) stats->n_tail_levels += last - T + 2;
    if (!fused)  // every lane of the group (lanes that started later included)
      HIP_TRY(launch_count_edges(b->W, slot(b, c, e), b->g->d_deg, b->g->Vt,
                                 c.d_masks + b->W, b->d_edges, c.st));
  }
  if (c.out_host) {  // rows of the caller's order: one piece when consecutive
    const size_t V = (size_t)b->g->V;
    bool consecutive = true;
    for (int i = 1; i < c.k && consecutive; ++i) consecutive = c.rows[i] == c.rows[0] + i;
    std::vector<D2HPiece> pcs;
    if (consecutive)
      pcs.push_back({c.out_host + (size_t)c.rows[0] * V, c.d_stage,
                     sizeof(int32_t) * (size_t)c.k * V});
    else
      for (int i = 0; i < c.k; ++i)
        pcs.push_back({c.out_host + (size_t)c.rows[i] * V, c.d_stage + (size_t)i * V,
                       sizeof(int32_t) * V});
    const int rc = copy_to_host(pcs.data(), pcs.size(), c.st);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(c.ev, c.st));
  return IBFS_OK;
}

}  // namespace

extern "C" {

int ibfs_batch_create(ibfs_graph_t g, const ibfs_config *cfg_in, int32_t contexts, int32_t burst,
                      ibfs_batch_t *out) {
  if (!g || !out) return fail(IBFS_ERR_ARG, "ibfs_batch_create: null argument");
  if (contexts < 1 || contexts > 16) return fail(IBFS_ERR_ARG, "contexts must be in [1, 16]");
  if (burst < 1 || burst > 64) return fail(IBFS_ERR_ARG, "burst must be in [1, 64]");
  ibfs_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else ibfs_config_default(&cfg);
  if (cfg.lane_bits != 64 && cfg.lane_bits != 128 && cfg.lane_bits != 256 && cfg.lane_bits != 512)
    return fail(IBFS_ERR_ARG, "lane_bits must be 64, 128, 256 or 512 (got %d)", cfg.lane_bits);
  if (cfg.sml_shed < 1 || cfg.lrg_shed < cfg.sml_shed)
    return fail(IBFS_ERR_ARG, "need 1 <= sml_shed <= lrg_shed");
  if (int rc = check_direction_config(cfg)) return rc;
  if (cfg.td_chunk < 256 || (cfg.td_chunk & (cfg.td_chunk - 1)))
    return fail(IBFS_ERR_ARG, "td_chunk must be a power of two >= 256");
  if (cfg.ring_levels < 4 || cfg.ring_levels > 255)
    return fail(IBFS_ERR_ARG, "ring_levels must be in [4, 255]");
  if (cfg.emit_mode < 0 || cfg.emit_mode > 3) return fail(IBFS_ERR_ARG, "bad emit_mode");
  if (!valid_inspect_vpt(cfg.inspect_vpt)) return fail(IBFS_ERR_ARG, "bad inspect_vpt");
  if (!valid_wg_size(cfg.wg_size)) return fail(IBFS_ERR_ARG, "wg_size must be 256, 512 or 1024");
  if (cfg.device != g->device) return fail(IBFS_ERR_ARG, "config device differs from the graph's");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(IBFS_ERR_NODEVICE, "no HIP device available");
  HIP_TRY(hipSetDevice(g->device));
  auto *b = new ibfs_batch_s;
  b->g = g;
  b->cfg = cfg;
  b->K = cfg.lane_bits;
  b->W = cfg.lane_bits / 64;
  b->C = contexts;
  b->burst = burst;
  b->chunk_shift = __builtin_ctz((unsigned)cfg.td_chunk);
  const size_t slot_bytes = (size_t)std::max<int64_t>(g->Vt, 1) * (size_t)b->W * sizeof(u64);
  int R = cfg.ring_levels;
  const size_t budget = ((size_t)16 << 30) / (size_t)contexts;
  while (R > 4 && (size_t)R * slot_bytes > budget) --R;
  b->R = R;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device) != hipSuccess ||
      cus <= 0)
    cus = 256;
  // 32 workgroups per CU: a fixed grid strides over the queues, and fewer
  // workgroups leave too few gathers in flight (measured on R-MAT s22)
  b->grid = cfg.dd_grid > 0 ? cfg.dd_grid : cus * 32;
  b->hc = hub_cut(b->g, cfg.hybrid_hubs);
  b->ctx.resize(contexts);
  hipError_t err = hipMalloc((void **)&b->d_edges, sizeof(u64));
  if (err == hipSuccess) err = hipHostMalloc((void **)&b->h_edges, sizeof(u64), hipHostMallocDefault);
  if (err != hipSuccess) {
    release_batch(b);
    delete b;
    return fail(IBFS_ERR_NOMEM, "batch allocation failed");
  }
  for (auto &c : b->ctx) {
    int rc = alloc_ctx(b, c);
    if (rc) {
      release_batch(b);
      delete b;
      return rc;
    }
  }
  *out = b;
  return IBFS_OK;
}

int ibfs_batch_destroy(ibfs_batch_t b) {
  if (!b) return IBFS_OK;
  release_batch(b);
  delete b;
  return IBFS_OK;
}

}  // extern "C"

// Waits, within the wait limit, for the work enqueued so far on every
// context's stream.
static int wait_contexts(ibfs_batch_s *b, const char *what) {
  for (auto &c : b->ctx) HIP_TRY(hipEventRecord(c.ev, c.st));
  for (auto &c : b->ctx) {
    WaitClock clk;
    for (;;) {
      const hipError_t q = hipEventQuery(c.ev);
      if (q == hipSuccess) break;
      if (q != hipErrorNotReady) return fail(IBFS_ERR_HIP, "context event: %s", hipGetErrorString(q));
      if (clk.expired())
        return fail(IBFS_ERR_STATE, "%s did not finish within the wait limit (%.1f s)", what,
                    wait_timeout_s());
      std::this_thread::yield();
    }
  }
  return IBFS_OK;
}

// Scheduler shared by depth runs and closeness runs (b->closeness).
static int batch_run_impl(ibfs_batch_s *b, const int32_t *src, int64_t n, int32_t *depth_out,
                          int32_t out_place, ibfs_stats *stats) {
  const int64_t V = b->g->V;
  for (int64_t i = 0; i < n; ++i)
    if (src[i] < 0 || src[i] >= V)
      return fail(IBFS_ERR_ARG, "source %lld = %d out of range", (long long)i, src[i]);
  if (stats) std::memset(stats, 0, sizeof *stats);
  if (n == 0) return IBFS_OK;
  HIP_TRY(hipSetDevice(b->g->device));
  clear_hip_error();
  if (b->unsettled) {
    if (int rc = wait_contexts(b, "the work of an earlier failed run")) return rc;
    b->unsettled = false;
  }
  if (out_place == IBFS_OUT_HOST) {
    for (auto &c : b->ctx)
      if (!c.d_stage) HIP_TRY(hipMalloc((void **)&c.d_stage, sizeof(int32_t) * (size_t)b->K * (size_t)V));
  }
  // grouping order: the caller's, or by descending two-hop size so that the
  // lanes of a group progress alike (results are per lane and unchanged)
  b->order.resize((size_t)n);
  std::iota(b->order.begin(), b->order.end(), (int64_t)0);
  // set-up work goes on the first context's stream and is waited for there
  // (no device-wide synchronisation: other engines, batches and servers on
  // the device keep running); the contexts start after it
  hipStream_t st0 = b->ctx[0].st;
  const bool pack = b->cfg.pack_sources && n > b->K;
  b->h_keys.clear();
  if (pack || b->cfg.phase_offsets) {
    if (b->key_cap < n) {
      dfree(b->d_keysrc);
      dfree(b->d_keys);
      b->key_cap = 0;
      HIP_TRY(hipMalloc((void **)&b->d_keysrc, sizeof(int32_t) * (size_t)n));
      HIP_TRY(hipMalloc((void **)&b->d_keys, sizeof(u64) * (size_t)n));
      b->key_cap = n;
    }
    HIP_TRY(hipMemcpyAsync(b->d_keysrc, src, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice,
                           st0));
    HIP_TRY(launch_two_hop(b->g->d_beg, b->g->d_adj, b->g->d_deg, b->g->d_cid, b->d_keysrc, n,
                           b->d_keys, st0));
    b->h_keys.resize((size_t)n);
    HIP_TRY(hipMemcpyAsync(b->h_keys.data(), b->d_keys, sizeof(u64) * (size_t)n,
                           hipMemcpyDeviceToHost, st0));
    HIP_TRY(hipStreamSynchronize(st0));
    if (pack)
      std::stable_sort(b->order.begin(), b->order.end(),
                       [&](int64_t x, int64_t y) { return b->h_keys[x] > b->h_keys[y]; });
  }
  b->psrc.resize((size_t)n);
  for (int64_t p = 0; p < n; ++p) b->psrc[p] = src[b->order[p]];
  HIP_TRY(hipMemsetAsync(b->d_edges, 0, sizeof(u64), st0));
  HIP_TRY(hipStreamSynchronize(st0));
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t groups = (n + b->K - 1) / b->K;
  int64_t next = 0, finished = 0;
  int rc = IBFS_OK;
  for (auto &c : b->ctx) c.state = IDLE;
  WaitClock clk;  // restarted whenever a context makes progress
  while (finished < groups && rc == IBFS_OK) {
    bool progressed = false;
    for (auto &c : b->ctx) {
      if (c.state == IDLE) {
        if (next >= groups) continue;
        const int64_t g0 = next * b->K;
        const int k = (int)std::min<int64_t>(b->K, n - g0);
        int32_t *dev_out = nullptr, *stage = nullptr;
        c.out_host = nullptr;
        if (!b->closeness) {
          if (out_place == IBFS_OUT_DEVICE) dev_out = depth_out;
          else {
            stage = c.d_stage;
            c.out_host = depth_out;
          }
        }
        ++next;
        c.no_tail = false;
        rc = start_group(b, c, b->psrc.data(), g0, k, dev_out, stage, b->order.data() + g0);
        if (!rc) rc = launch_burst(b, c, stats);
        c.state = RUNNING;
        progressed = true;
      } else {
        hipError_t q = hipEventQuery(c.ev);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) {
          rc = fail(IBFS_ERR_HIP, "context event: %s", hipGetErrorString(q));
          break;
        }
        progressed = true;
        if (c.state == RUNNING) {
          if (c.h_plan->done && c.h_plan->tail_start && c.h_plan->tail_count > c.tail_cap) {
            // the sparse tail's log overflowed: the group again, dense (a
            // closeness window flushed before the tail is counted again, so
            // the group's sums restart from zero)
            c.no_tail = true;
            if (b->closeness) {
              rc = launch_fill_u64(b->d_sum + c.g0, c.k, 0ull, c.st) == hipSuccess &&
                           launch_fill_u64(b->d_reach + c.g0, c.k, 0ull, c.st) == hipSuccess
                       ? IBFS_OK
                       : fail(IBFS_ERR_HIP, "closeness sums reset failed");
            }
            if (!rc) rc = start_group(b, c, c.src, c.g0, c.k, c.dev_out, c.stage, c.rows);
            if (!rc) rc = launch_burst(b, c, stats);
            if (stats) stats->n_tail_reruns++;
          } else if (c.h_plan->done) {
            // the group's emission is enqueued; the context's next group can
            // be enqueued behind it at once (stream order: its set-up copy
            // and fills start after the emission has read the ring, the lane
            // table and the tail log), and the final stream synchronisation
            // waits for the last one
            rc = finish_group(b, c, c.h_plan->last, stats);
            c.state = IDLE;
            ++finished;
          } else if (c.h_plan->level > V + 2) {
            rc = fail(IBFS_ERR_STATE, "traversal exceeded V+2 levels");
          } else {
            c.level = c.h_plan->level;
            rc = launch_burst(b, c, stats);
          }
        }
      }
      if (rc) break;
    }
    if (progressed) {
      clk.restart();
    } else {
      if (clk.expired()) {
        rc = fail(IBFS_ERR_STATE, "no batch context progressed within the wait limit (%.1f s)",
                  wait_timeout_s());
        break;
      }
      std::this_thread::yield();
    }
  }
  if (rc) {
    // a context that exceeded the wait limit is not waited for here; the
    // next run drains it
    if (rc != IBFS_ERR_STATE)
      for (auto &c : b->ctx) (void)hipStreamSynchronize(c.st);
    else
      b->unsettled = true;
    return rc;
  }
  // the last groups' emissions (enqueued behind their levels) end within the
  // wait limit like every other wait of the scheduler
  rc = wait_contexts(b, "emission");
  if (rc) {
    b->unsettled = true;
    return rc;
  }
  if (stats) {
    stats->total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    HIP_TRY(hipMemcpy(b->h_edges, b->d_edges, sizeof(u64), hipMemcpyDeviceToHost));
    stats->traversed_edges = (int64_t)*b->h_edges;
  }
  return IBFS_OK;
}

extern "C" {

int ibfs_batch_run(ibfs_batch_t b, const int32_t *src, int64_t n, int32_t *depth_out,
                   int32_t out_place, ibfs_stats *stats) {
  if (!b || n < 0 || (n > 0 && (!src || !depth_out)))
    return fail(IBFS_ERR_ARG, "ibfs_batch_run: bad arguments");
  if (out_place != IBFS_OUT_HOST && out_place != IBFS_OUT_DEVICE)
    return fail(IBFS_ERR_ARG, "bad out_place %d", out_place);
  b->closeness = false;
  return batch_run_impl(b, src, n, depth_out, out_place, stats);
}

int ibfs_batch_closeness(ibfs_batch_t b, const int32_t *src, int64_t n, int64_t *depth_sum,
                         int64_t *reach, ibfs_stats *stats) {
  if (!b || n < 0 || (n > 0 && (!src || !depth_sum || !reach)))
    return fail(IBFS_ERR_ARG, "ibfs_batch_closeness: bad arguments");
  if (n == 0) {
    if (stats) std::memset(stats, 0, sizeof *stats);
    return IBFS_OK;
  }
  HIP_TRY(hipSetDevice(b->g->device));
  if (b->sum_cap < n) {
    dfree(b->d_sum);
    dfree(b->d_reach);
    b->sum_cap = 0;
    HIP_TRY(hipMalloc((void **)&b->d_sum, sizeof(u64) * (size_t)n));
    HIP_TRY(hipMalloc((void **)&b->d_reach, sizeof(u64) * (size_t)n));
    b->sum_cap = n;
  }
  // cleared on the first context's stream, which the run synchronises
  // before any context starts
  HIP_TRY(hipMemsetAsync(b->d_sum, 0, sizeof(u64) * (size_t)n, b->ctx[0].st));
  HIP_TRY(hipMemsetAsync(b->d_reach, 0, sizeof(u64) * (size_t)n, b->ctx[0].st));
  b->closeness = true;
  int rc = batch_run_impl(b, src, n, nullptr, IBFS_OUT_DEVICE, stats);
  b->closeness = false;
  if (rc) return rc;
  // the device sums are in grouping order; back to the caller's
  std::vector<u64> hs((size_t)n), hr((size_t)n);
  HIP_TRY(hipMemcpy(hs.data(), b->d_sum, sizeof(u64) * (size_t)n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(hr.data(), b->d_reach, sizeof(u64) * (size_t)n, hipMemcpyDeviceToHost));
  for (int64_t p = 0; p < n; ++p) {
    depth_sum[b->order[p]] = (int64_t)hs[p];
    reach[b->order[p]] = (int64_t)hr[p];
  }
  // an isolated source reaches itself (depth 0) and has no status word
  for (int64_t i = 0; i < n; ++i)
    if (b->g->h_cid[src[i]] < 0) reach[i] += 1;
  return IBFS_OK;
}

}  // extern "C"
