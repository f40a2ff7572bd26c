// Serving loop: continuous batching of BFS queries (1-256 sources each)
// against a device-resident graph.
//
// Every source of a query occupies one status lane of one traversal context
// from its admission level a until its BFS ends; its depths are L - a.  A
// worker thread owns the GPU.  With host-driven levels, at each level
// boundary of a context it (1) emits and frees the lanes whose BFS ended (no
// vertex gained the lane at this level), (2) issues the next expansion, (3)
// admits waiting sources into clean lanes, (4) issues the inspection that
// also clears the freed lanes in the next snapshot.  Lanes are independent bits, so a query's depths do not
// depend on what shares the context with it.  Contexts are multiplexed with
// events, so one context's level boundary overlaps another's kernels.
//
// Admission is by cohort: a context takes waiting sources when all its lanes
// have finished (a new cohort) and, within admit_window levels of that
// start, into its remaining clean lanes.  A level's direction and cost are
// shared by all lanes of a context, so lanes far apart in BFS phase would
// make every level as expensive as the densest phase among them.  By default
// (host_levels 0) a cohort runs on the device, at the narrowest status-word
// width that holds its first sources: captured level graphs replayed in
// bursts with one host poll per burst; inside the admission window the host
// takes every level boundary and the sources waiting there join the running
// cohort through a device-side admission list; all its lanes are emitted
// once it ends.  With host_levels 1 every level boundary is on the host and
// lanes are emitted, freed and reused one by one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "kernels.h"

using namespace ibfs;

#define HIP_CHK(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(IBFS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                  __FILE__, __LINE__);                                                 \
  } while (0)

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Query {
  int64_t id = 0;
  int n = 0;
  const int32_t *src = nullptr;  // owned copy in `srcbuf`
  std::vector<int32_t> srcbuf;
  int32_t *out = nullptr;        // device [n][V]
  int remaining = 0;             // sources not finished
  double t_submit = 0, t_admit = -1, t_done = -1;
  bool done = false;
};

struct PendingSrc {
  int64_t qid;
  int idx;
};

// CLEAN: free; ACTIVE: its BFS runs; FINISHED: its BFS ended, depths not yet
// emitted (its bits stay in the ring); DIRTY: freed, bits being cleared.
enum LaneState { CLEAN = 0, ACTIVE = 1, FINISHED = 2, DIRTY = 3 };

struct Lane {
  int state = CLEAN;
  int64_t qid = -1;
  int admit = 0;
  int freed_at = 0;
};

struct SCtx {
  hipStream_t st = nullptr;
  u64 *ring = nullptr;
  LevelCtl *ctl = nullptr;
  LevelCtl *h_ctl = nullptr;
  Queues q{};
  u64 *hs = nullptr;
  uint8_t *flags = nullptr;  // changed-at-this-level byte per traversal vertex
  u64 *d_mask = nullptr, *h_mask = nullptr;
  int32_t *d_adm = nullptr, *h_adm = nullptr;  // [0,K) sources, [K,2K) lanes
  LaneOut *d_lanes = nullptr, *h_lanes = nullptr;
  hipEvent_t ev = nullptr;
  std::vector<Lane> lanes;
  bool running = false;
  int L = 0, base = 0, dir_prev = -1;
  int cohort = 0;  // level at which the current cohort's first sources were admitted
  u64 nf_prev = ~0ull;
  std::vector<int64_t> done_pending;
  // Device-driven cohorts (admission at cohort starts only): a captured level
  // graph (expand, inspect, direction decision) replayed in bursts.
  Plan *d_plan = nullptr, *h_plan = nullptr;
  u64 *d_all = nullptr, *h_all = nullptr;  // lanes of the cohort (lam of level 0)
  AdmitList *d_al = nullptr, *h_al = nullptr;  // sources joining a running cohort
  // one captured level graph per status-word width (1, 2, 4, 8 words), so a
  // cohort runs at the narrowest width that holds its sources
  hipGraph_t graph[4] = {};
  hipGraphExec_t exec[4] = {};
  // per width, kCohortBurst levels in one graph for full bursts (consecutive
  // graph launches leave ≈8 µs between levels)
  hipGraph_t graph_burst[4] = {};
  hipGraphExec_t exec_burst[4] = {};
  int W = 0, K = 0;  // current width: words per vertex and lanes (<= the server's)
  int coh = 0;     // 0: host-driven levels, 1: cohort levels on the device, 2: cohort emitting
  int dlevel = 0;  // cohort level the device expands next (host view)
  int last_admit = 0;  // latest level at which sources joined the cohort
  // phase offsets: lanes of the cohort's first sources that start after
  // level 0 (lane, start level), and the latest start
  std::vector<std::pair<int, int>> deferred;
  int a_max = 0;
  WaitClock clk;   // since the context's last completed event (wait limit)
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

struct ibfs_server_s {
  ibfs_graph_s *g = nullptr;
  ibfs_config cfg{};
  int K = 256, W = 4, R = 32, C = 1, chunk_shift = 12;
  std::vector<SCtx> ctx;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int error = IBFS_OK;
  std::string error_msg;
  int64_t next_id = 1;
  std::unordered_map<int64_t, Query> queries;
  std::deque<PendingSrc> incoming;  // submitted, not yet seen by the worker
  std::deque<PendingSrc> pending;   // worker-owned FIFO of sources to admit
  std::atomic<int64_t> levels_issued{0}, lanes_admitted{0};
  // sources join a running context only within this many levels of its
  // cohort's start, so that the lanes of a context stay close in phase (set
  // from any thread, read by the worker)
  std::atomic<int> admit_window{0};
  // 1: host-driven levels (every level boundary on the host; lanes finish,
  // are emitted and reused one by one); 0: device-driven cohorts whose
  // levels run as captured graphs, joined by waiting sources at the level
  // boundaries inside the admission window (set from any thread)
  std::atomic<int> host_levels{0};
  // sources that joined a running cohort (after its first level)
  std::atomic<int64_t> lanes_admitted_in_flight{0};
  int grid = 2048;  // workgroups of the device-driven expander
  HubCut hc;        // hubs a hybrid level pulls through
  std::vector<u64> h_key;  // two-hop size of every vertex (cohort composition)
};

namespace {
constexpr int kCohortBurst = 4;  // level graphs per host poll in cohort mode
}

namespace {

// mask staging (words): E1 [sel W | levels (R+1)W], E2 same, I [filter W | keep W | nand W]
size_t emit_words(const ibfs_server_s *s) { return (size_t)(s->R + 2) * s->W; }
size_t mask_words(const ibfs_server_s *s) { return 2 * emit_words(s) + 3 * (size_t)s->W; }

u64 *slot(const ibfs_server_s *s, const SCtx &c, int L) {
  return c.ring + (size_t)(L % s->R) * (size_t)s->g->Vt * (size_t)c.W;
}

void release_ctx(SCtx &c) {
  if (c.st) (void)hipStreamSynchronize(c.st);
  for (int i = 0; i < 4; ++i) {
    if (c.exec[i]) (void)hipGraphExecDestroy(c.exec[i]);
    if (c.graph[i]) (void)hipGraphDestroy(c.graph[i]);
    if (c.exec_burst[i]) (void)hipGraphExecDestroy(c.exec_burst[i]);
    if (c.graph_burst[i]) (void)hipGraphDestroy(c.graph_burst[i]);
    c.exec[i] = c.exec_burst[i] = nullptr;
    c.graph[i] = c.graph_burst[i] = nullptr;
  }
  dfree(c.d_plan);
  hfree(c.h_plan);
  dfree(c.d_all);
  hfree(c.h_all);
  dfree(c.d_al);
  hfree(c.h_al);
  dfree(c.ring);
  dfree(c.ctl);
  hfree(c.h_ctl);
  for (int i = 0; i < 6; ++i) dfree(c.q.q[i]);
  dfree(c.q.chunks);
  dfree(c.hs);
  dfree(c.flags);
  dfree(c.d_mask);
  hfree(c.h_mask);
  dfree(c.d_adm);
  hfree(c.h_adm);
  dfree(c.d_lanes);
  hfree(c.h_lanes);
  if (c.ev) (void)hipEventDestroy(c.ev);
  c.ev = nullptr;
  if (c.st) (void)hipStreamDestroy(c.st);
  c.st = nullptr;
}

// Inspector arguments of a cohort level (plan mode: the level, its slots and
// control blocks come from the plan on the device).
InspectArgs plan_args(const ibfs_server_s *s, const SCtx &c) {
  InspectArgs a{};
  a.deg = s->g->d_deg;
  a.V = s->g->Vt;
  a.q = c.q;
  a.sml = s->cfg.sml_shed;
  a.lrg = s->cfg.lrg_shed;
  a.chunk_shift = s->chunk_shift;
  a.hubs = s->g->d_hubs;
  a.H = s->g->H;
  a.hs = c.hs;
  a.plan = c.d_plan;
  a.ring = c.ring;
  a.R = s->R;
  a.ctl_ring = c.ctl;
  a.lanes_all = c.d_all;
  a.flags = c.flags;
  a.vpt = s->cfg.inspect_vpt;
  a.nw = s->cfg.wg_size / 64;
  a.hsplit = HubSplit{s->hc.deg, s->hc.tie, s->hc.cut};
  return a;
}

int alloc_ctx(ibfs_server_s *s, SCtx &c) {
  const int64_t V = std::max<int64_t>(s->g->Vt, 1);  // status words and queues: traversal ids
  hipError_t err = hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking);
  auto A = [&](void **p, size_t bytes) {
    if (err == hipSuccess) err = hipMalloc(p, bytes ? bytes : 1);
  };
  auto H = [&](void **p, size_t bytes) {
    if (err == hipSuccess) err = hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault);
  };
  A((void **)&c.ring, (size_t)s->R * (size_t)V * (size_t)s->W * sizeof(u64));
  A((void **)&c.ctl, sizeof(LevelCtl) * (size_t)s->R);
  H((void **)&c.h_ctl, sizeof(LevelCtl));
  for (int i = 0; i < 6; ++i)
    if (i != Q_TD_L) A((void **)&c.q.q[i], sizeof(int32_t) * (size_t)V);
  A((void **)&c.q.chunks, sizeof(int2) * ((size_t)V + (size_t)(s->g->E >> s->chunk_shift) + 2));
  A((void **)&c.hs, sizeof(u64) * (size_t)s->W * (size_t)std::max<int64_t>(s->g->H, 1));
  A((void **)&c.flags, sizeof(u64) * (size_t)((V + 7) / 8));
  A((void **)&c.d_mask, sizeof(u64) * mask_words(s));
  H((void **)&c.h_mask, sizeof(u64) * mask_words(s));
  A((void **)&c.d_adm, sizeof(int32_t) * 2 * (size_t)s->K);
  H((void **)&c.h_adm, sizeof(int32_t) * 2 * (size_t)s->K);
  A((void **)&c.d_lanes, sizeof(LaneOut) * (size_t)s->K);
  H((void **)&c.h_lanes, sizeof(LaneOut) * (size_t)s->K);
  A((void **)&c.d_plan, sizeof(Plan));
  H((void **)&c.h_plan, sizeof(Plan));
  A((void **)&c.d_all, sizeof(u64) * (size_t)s->W);
  H((void **)&c.h_all, sizeof(u64) * (size_t)s->W);
  A((void **)&c.d_al, sizeof(AdmitList));
  H((void **)&c.h_al, sizeof(AdmitList));
  if (err == hipSuccess) err = hipEventCreateWithFlags(&c.ev, hipEventDisableTiming);
  if (err != hipSuccess)
    return fail(err == hipErrorOutOfMemory ? IBFS_ERR_NOMEM : IBFS_ERR_HIP,
                "server context allocation failed: %s", hipGetErrorString(err));
  std::memset(c.h_lanes, 0, sizeof(LaneOut) * (size_t)s->K);
  c.lanes.assign(s->K, Lane{});
  // one cohort level per width: device-driven expansion, inspection,
  // direction decision (every kernel derives its slot stride from its width)
  c.W = s->W;
  c.K = s->K;
  // (and kCohortBurst of them in a row: every kernel reads its level from
  // the plan, so repeated bodies run consecutive levels)
  for (int i = 0; i < 4 && (1 << i) <= s->W; ++i) {
    const int Wc = 1 << i;
    for (int pass = 0; pass < 2; ++pass) {
      const int reps = pass == 0 ? 1 : kCohortBurst;
      clear_hip_error();
      HIP_CHK(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
      hipError_t e1 = hipSuccess, e2 = hipSuccess;
      for (int r = 0; r < reps && e1 == hipSuccess && e2 == hipSuccess; ++r) {
        e1 = launch_expand_dd(Wc, s->g->d_beg, s->g->d_adj, s->g->d_tadj, c.ring, c.flags, s->R,
                              s->g->Vt, c.ctl, c.d_plan, c.hs, c.q, s->chunk_shift,
                              HubSplit{s->hc.deg, s->hc.tie, s->hc.cut}, TailLog{}, s->grid,
                              c.st);
        e2 = launch_inspect(Wc, plan_args(s, c), c.st);
        if (e2 == hipSuccess)
          e2 = launch_plan_decide(c.d_plan, c.ctl, s->R, c.hs, Wc,
                                  split_hubs(s->g, s->hc, s->cfg), c.st);
      }
      hipGraph_t &gr = pass == 0 ? c.graph[i] : c.graph_burst[i];
      hipError_t e3 = hipStreamEndCapture(c.st, &gr);
      if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
        return fail(IBFS_ERR_HIP, "cohort level capture failed: %s / %s / %s",
                    hipGetErrorString(e1), hipGetErrorString(e2), hipGetErrorString(e3));
      HIP_CHK(hipGraphInstantiate(pass == 0 ? &c.exec[i] : &c.exec_burst[i], gr, nullptr,
                                  nullptr, 0));
    }
  }
  return IBFS_OK;
}

void set_bit(u64 *m, int k) { m[k >> 6] |= 1ull << (k & 63); }
bool has_bit(const u64 *m, int k) { return (m[k >> 6] >> (k & 63)) & 1ull; }

// Emission of lanes `sel` over [b, e] using staging slot `which` (0/1).
int emit_lanes(ibfs_server_s *s, SCtx &c, const std::vector<int> &lanes, int b, int e, int which) {
  if (lanes.empty() || e < b) return IBFS_OK;
  const int W = c.W;
  u64 *h = c.h_mask + (size_t)which * emit_words(s);
  u64 *d = c.d_mask + (size_t)which * emit_words(s);
  std::memset(h, 0, sizeof(u64) * emit_words(s));
  for (int k : lanes) set_bit(h, k);
  for (int L = b; L <= e; ++L)
    for (int k : lanes)
      if (c.lanes[k].admit <= L) set_bit(h + W + (size_t)(L - b) * W, k);
  HIP_CHK(hipMemcpyAsync(d, h, sizeof(u64) * emit_words(s), hipMemcpyHostToDevice, c.st));
  const EmitSpace sp{s->g->d_cid, s->g->V, s->g->Vt, s->g->d_deg};
  HIP_CHK(launch_emit(W, s->cfg.emit_mode, 0, c.ring, s->R, sp, b, e, c.d_lanes, d + W, d, c.K,
                      nullptr, c.st));
  return IBFS_OK;
}

// Admits up to the free clean lanes' worth of pending sources at level L
// (their bits go into slot L); returns the admitted lane mask in A.
int admit(ibfs_server_s *s, SCtx &c, int L, u64 *A, int *n_adm) {
  int n = 0;
  for (int k = 0; k < c.K && !s->pending.empty(); ++k) {
    if (c.lanes[k].state != CLEAN) continue;
    const PendingSrc p = s->pending.front();
    s->pending.pop_front();
    Query &q = s->queries[p.qid];
    c.lanes[k].state = ACTIVE;
    c.lanes[k].qid = p.qid;
    c.lanes[k].admit = L;
    c.h_adm[n] = s->g->h_cid[q.src[p.idx]];
    c.h_adm[s->K + n] = k;
    c.h_lanes[k].out = q.out + (size_t)p.idx * (size_t)s->g->V;
    c.h_lanes[k].admit = L;
    c.h_lanes[k].src = q.src[p.idx];
    if (q.t_admit < 0) q.t_admit = now_ms();
    set_bit(A, k);
    ++n;
  }
  *n_adm = n;
  if (n == 0) return IBFS_OK;
  s->lanes_admitted += n;
  HIP_CHK(hipMemcpyAsync(c.d_adm, c.h_adm, sizeof(int32_t) * 2 * (size_t)s->K,
                         hipMemcpyHostToDevice, c.st));
  HIP_CHK(hipMemcpyAsync(c.d_lanes, c.h_lanes, sizeof(LaneOut) * (size_t)s->K,
                         hipMemcpyHostToDevice, c.st));
  HIP_CHK(launch_admit(c.W, c.d_adm, c.d_adm + s->K, n, slot(s, c, L), c.flags, c.st));
  return IBFS_OK;
}

// Inspection of level L (masks in staging slot I) + control-block copy + event.
int inspect_level(ibfs_server_s *s, SCtx &c, int L, const u64 *filter, const u64 *keep,
                  const u64 *nand, int full_diff) {
  const int W = c.W;
  u64 *h = c.h_mask + 2 * emit_words(s);
  u64 *d = c.d_mask + 2 * emit_words(s);
  for (int w = 0; w < W; ++w) {
    h[w] = filter[w];
    h[W + w] = keep[w];
    h[2 * W + w] = nand[w];
  }
  HIP_CHK(hipMemcpyAsync(d, h, sizeof(u64) * 3 * W, hipMemcpyHostToDevice, c.st));
  InspectArgs a{};
  a.prev = L > 0 ? slot(s, c, L - 1) : nullptr;
  a.cur = slot(s, c, L);
  a.next = slot(s, c, L + 1);
  a.lam_prev = d;
  a.prev_keep = d + W;
  a.next_and = d + 2 * W;
  a.ctl = c.ctl + L % s->R;
  a.ctl_zero = c.ctl + (L + 1) % s->R;
  a.deg = s->g->d_deg;
  a.V = s->g->Vt;
  a.q = c.q;
  a.sml = s->cfg.sml_shed;
  a.lrg = s->cfg.lrg_shed;
  a.chunk_shift = s->chunk_shift;
  a.hubs = s->g->d_hubs;
  a.H = s->g->H;
  a.hs = c.hs;
  a.flags = c.flags;
  a.vpt = s->cfg.inspect_vpt;
  a.nw = s->cfg.wg_size / 64;
  a.hsplit = HubSplit{s->hc.deg, s->hc.tie, s->hc.cut};
  a.full_diff = full_diff;
  HIP_CHK(launch_inspect(W, a, c.st));
  HIP_CHK(hipMemcpyAsync(c.h_ctl, c.ctl + L % s->R, sizeof(LevelCtl), hipMemcpyDeviceToHost, c.st));
  HIP_CHK(hipEventRecord(c.ev, c.st));
  c.L = L;
  ++s->levels_issued;
  return IBFS_OK;
}

int start_ctx(ibfs_server_s *s, SCtx &c) {
  c.W = s->W;  // admission mid-flight: the full width
  c.K = s->K;
  const int W = c.W;
  for (auto &l : c.lanes) l = Lane{};
  c.base = 0;
  c.cohort = 0;
  c.dir_prev = -1;
  c.nf_prev = ~0ull;
  // S_0, the changed flags and the control blocks zeroed in one launch
  HIP_CHK(launch_zero3_u64(slot(s, c, 0), s->g->Vt * W, reinterpret_cast<u64 *>(c.flags),
                           (s->g->Vt + 7) / 8, reinterpret_cast<u64 *>(c.ctl),
                           (int64_t)s->R * CTL_WORDS, c.st));
  std::vector<u64> A(W, 0), keep(W, ~0ull), nand(W, ~0ull);
  int n = 0;
  int rc = admit(s, c, 0, A.data(), &n);
  if (rc) return rc;
  c.running = true;
  return inspect_level(s, c, 0, A.data(), keep.data(), nand.data(), 1);
}

int decide(const ibfs_server_s *s, int Ln, const LevelCtl &lc, int dir_prev, u64 nf_prev) {
  const ibfs_config &k = s->cfg;
  return choose_direction(k.policy, k.sw_level, k.alpha, k.beta, k.hybrid_gamma, s->hc.rho_q10,
                          (u64)s->g->V, Ln, dir_prev, nf_prev, lc.n_front, lc.m_front, lc.m_cand,
                          lc.m_front_nh);
}

void mark_done(ibfs_server_s *s, SCtx &c) {
  if (c.done_pending.empty()) return;
  const double t = now_ms();
  for (int64_t id : c.done_pending) {
    auto it = s->queries.find(id);
    if (it != s->queries.end()) {
      it->second.done = true;
      it->second.t_done = t;
    }
  }
  c.done_pending.clear();
  s->cv_done.notify_all();
}

// Level L of context c has completed (its control block is in h_ctl).
int process_level(ibfs_server_s *s, SCtx &c) {
  const int W = c.W;
  const int L = c.L;
  const LevelCtl lc = *c.h_ctl;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    mark_done(s, c);  // emissions issued before this level's event are complete
  }
  // 1. lanes whose BFS ended finish.  With admission at cohort starts only,
  //    their depths are emitted once no lane of the cohort is still active
  //    (one pass over the ring window per cohort instead of one per level at
  //    which lanes finish; a finished lane's bits no longer change, so a
  //    longer window gives the same depths).  Otherwise they are emitted now.
  //    Emission starts at the earliest admission among them (the window base
  //    if one is older); emitted lanes are freed.
  bool active_left = false;
  for (int k = 0; k < c.K; ++k) {
    Lane &l = c.lanes[k];
    if (l.state == ACTIVE && l.admit < L && !has_bit(lc.lam, k)) l.state = FINISHED;
    active_left |= (l.state == ACTIVE);
  }
  std::vector<int> fin;
  std::vector<u64> finmask(W, 0);
  int fin_from = L;
  if (!(s->admit_window == 0 && active_left)) {
    for (int k = 0; k < c.K; ++k) {
      Lane &l = c.lanes[k];
      if (l.state == FINISHED) {
        fin.push_back(k);
        set_bit(finmask.data(), k);
        fin_from = std::min(fin_from, l.admit);
      }
    }
  }
  int rc = emit_lanes(s, c, fin, std::max(c.base, fin_from), L - 1, 0);
  if (rc) return rc;
  bool any_active = false;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (int k : fin) {
      Lane &l = c.lanes[k];
      Query &q = s->queries[l.qid];
      if (--q.remaining == 0) c.done_pending.push_back(l.qid);
      l.state = DIRTY;
      l.freed_at = L;
      l.qid = -1;
    }
    for (auto &l : c.lanes) {
      if (l.state == DIRTY && l.freed_at <= L - 1) l.state = CLEAN;  // cleared in S_{L+1}
      any_active |= (l.state == ACTIVE || l.state == FINISHED);
    }
    if (!any_active && s->pending.empty()) {
      c.running = false;
      HIP_CHK(hipStreamSynchronize(c.st));
      mark_done(s, c);
      return IBFS_OK;
    }
  }
  // 2. expansion of L+1
  const int Ln = L + 1;
  const int dir = decide(s, Ln, lc, c.dir_prev, c.nf_prev);
  const u64 *lamp = reinterpret_cast<const u64 *>(c.ctl + L % s->R);
  if (dir == 0)
    HIP_CHK(launch_expand_td(W, s->g->d_beg, s->g->d_adj, slot(s, c, L), slot(s, c, Ln), lamp,
                             c.q, lc.q[Q_TD_S], lc.q[Q_TD_M], lc.q[Q_TD_L], s->chunk_shift,
                             nullptr, c.st));
  else if (dir == 2)
    HIP_CHK(launch_expand_hy(W, s->g->d_beg, s->g->d_adj, s->g->d_tadj, slot(s, c, L), c.hs,
                             slot(s, c, Ln), c.ctl + L % s->R, c.q, s->chunk_shift,
                             HubSplit{s->hc.deg, s->hc.tie, s->hc.cut}, s->grid, c.st));
  else
    HIP_CHK(launch_expand_bu(W, s->g->d_beg, s->g->d_tadj, slot(s, c, L), c.hs, slot(s, c, Ln),
                             c.flags, lamp, c.q, lc.q[Q_BU_S], lc.q[Q_BU_M], lc.q[Q_BU_L], c.st));
  c.dir_prev = dir;
  c.nf_prev = lc.n_front;
  // 3. snapshots older than every unemitted lane's admission are no longer
  //    needed: the window base follows the oldest active or finished lane.
  //    If the next copy (slot Ln+1) would still overwrite the base, flush all
  //    of them over [base, Ln].
  {
    int oldest = Ln;
    for (int k = 0; k < c.K; ++k)
      if (c.lanes[k].state == ACTIVE || c.lanes[k].state == FINISHED)
        oldest = std::min(oldest, c.lanes[k].admit);
    c.base = std::max(c.base, oldest);
  }
  if ((Ln + 1) - c.base >= s->R) {
    std::vector<int> act;
    for (int k = 0; k < c.K; ++k)
      if (c.lanes[k].state == ACTIVE || c.lanes[k].state == FINISHED) act.push_back(k);
    rc = emit_lanes(s, c, act, c.base, Ln, 1);
    if (rc) return rc;
    c.base = Ln;
  }
  // 4. admission at Ln: into a context whose lanes have all finished (a new
  //    cohort starts), or while the current cohort is young enough
  std::vector<u64> A(W, 0), filter(W), keep(W), nand(W);
  int n_adm = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    const bool fresh = !any_active;
    if (fresh || Ln - c.cohort <= s->admit_window) {
      rc = admit(s, c, Ln, A.data(), &n_adm);
      if (fresh && n_adm > 0) c.cohort = Ln;
      if (!fresh) s->lanes_admitted_in_flight += n_adm;
    }
  }
  if (rc) return rc;
  for (int w = 0; w < W; ++w) {
    filter[w] = lc.lam[w] | A[w];
    keep[w] = ~A[w];
    nand[w] = ~finmask[w];
  }
  // 5. inspection of Ln (clears the lanes freed at L in slot Ln+1)
  return inspect_level(s, c, Ln, filter.data(), keep.data(), nand.data(), dir != 1 ? 1 : 0);
}

// Waiting sources take the clean lanes of a running cohort at level L
// (caller holds the lock): entries n0, n0+1, ... of the admission list, the
// lane table, host bookkeeping.  *n_new = sources admitted.
void take_pending(ibfs_server_s *s, SCtx &c, int L, int n0, int *n_new) {
  int n = n0;
  for (int k = 0; k < c.K && !s->pending.empty(); ++k) {
    if (c.lanes[k].state != CLEAN) continue;
    const PendingSrc p = s->pending.front();
    s->pending.pop_front();
    Query &q = s->queries[p.qid];
    c.lanes[k].state = ACTIVE;
    c.lanes[k].qid = p.qid;
    c.lanes[k].admit = L;
    c.h_al->src[n] = s->g->h_cid[q.src[p.idx]];
    c.h_al->lane[n] = k;
    c.h_lanes[k].out = q.out + (size_t)p.idx * (size_t)s->g->V;
    c.h_lanes[k].admit = L;
    c.h_lanes[k].src = q.src[p.idx];
    if (q.t_admit < 0) q.t_admit = now_ms();
    ++n;
  }
  *n_new = n - n0;
  s->lanes_admitted += n - n0;
  s->lanes_admitted_in_flight += n - n0;
}

// Launches the cohort's next levels, then a copy of the plan and the
// context's event.  Inside the admission window (levels 1..admit_window of
// the cohort) one level at a time, each joined by the sources waiting at its
// boundary: expansion, admission, inspection and decision launched
// directly; after the window up to kCohortBurst captured level graphs
// (bounded so that no snapshot of the emission window is overwritten; the
// window is flushed first when it is full).  locked: the caller holds the
// server lock.
int cohort_burst(ibfs_server_s *s, SCtx &c, bool locked) {
  int B = std::min(kCohortBurst, s->R - 1 - (c.dlevel - c.base));
  if (B < 1) {
    std::vector<int> act;
    for (int k = 0; k < c.K; ++k)
      if (c.lanes[k].state == ACTIVE) act.push_back(k);
    int rc = emit_lanes(s, c, act, c.base, c.dlevel - 1, 1);
    if (rc) return rc;
    c.base = c.dlevel - 1;
    B = std::min(kCohortBurst, s->R - 1 - (c.dlevel - c.base));
  }
  hipGraphExec_t x = c.exec[__builtin_ctz((unsigned)c.W)];
  const int win = s->admit_window;
  const bool in_window = win > 0 && c.dlevel <= win;
  if (in_window || c.dlevel <= c.a_max) {
    // lanes of the cohort's first sources that start at this level, then
    // (inside the admission window) sources waiting now
    int n = 0, fresh = 0;
    for (const auto &d : c.deferred)
      if (d.second == c.dlevel) {
        c.h_al->src[n] = s->g->h_cid[c.h_lanes[d.first].src];
        c.h_al->lane[n] = d.first;
        ++n;
      }
    if (in_window) {
      if (locked) {
        take_pending(s, c, c.dlevel, n, &fresh);
      } else {
        std::lock_guard<std::mutex> lk(s->mu);
        take_pending(s, c, c.dlevel, n, &fresh);
      }
    }
    n += fresh;
    if (n > 0) {
      c.h_al->level = c.dlevel;
      c.h_al->n = n;
      c.last_admit = c.dlevel;
      HIP_CHK(hipMemcpyAsync(c.d_al, c.h_al, sizeof(AdmitList), hipMemcpyHostToDevice, c.st));
      if (fresh > 0)
        HIP_CHK(hipMemcpyAsync(c.d_lanes, c.h_lanes, sizeof(LaneOut) * (size_t)c.K,
                               hipMemcpyHostToDevice, c.st));
      const HubSplit hsp{s->hc.deg, s->hc.tie, s->hc.cut};
      HIP_CHK(launch_expand_dd(c.W, s->g->d_beg, s->g->d_adj, s->g->d_tadj, c.ring, c.flags, s->R,
                               s->g->Vt, c.ctl, c.d_plan, c.hs, c.q, s->chunk_shift, hsp,
                               TailLog{}, s->grid, c.st));
      HIP_CHK(launch_admit_dd(c.W, c.d_al, c.d_plan, c.ring, s->R, s->g->Vt, c.ctl, c.flags, c.st));
      HIP_CHK(launch_inspect(c.W, plan_args(s, c), c.st));
      HIP_CHK(launch_plan_decide(c.d_plan, c.ctl, s->R, c.hs, c.W,
                                 split_hubs(s->g, s->hc, s->cfg), c.st));
    } else {
      HIP_CHK(hipGraphLaunch(x, c.st));
    }
  } else if (B == kCohortBurst) {
    HIP_CHK(hipGraphLaunch(c.exec_burst[__builtin_ctz((unsigned)c.W)], c.st));
  } else {
    for (int i = 0; i < B; ++i) HIP_CHK(hipGraphLaunch(x, c.st));
  }
  HIP_CHK(hipMemcpyAsync(c.h_plan, c.d_plan, offsetof(Plan, dirlog), hipMemcpyDeviceToHost, c.st));
  HIP_CHK(hipEventRecord(c.ev, c.st));
  ++s->levels_issued;
  return IBFS_OK;
}

// Cohort composition when more sources wait than a cohort holds (caller
// holds the lock): of the oldest 8K waiting sources, the K nearest in two-hop
// size to the oldest one move to the front, the oldest included, so that the
// cohort's lanes progress alike; everything else keeps its arrival order.
void pick_cohort(ibfs_server_s *s, int K) {
  const size_t n = s->pending.size();
  if (n <= (size_t)K || s->h_key.empty()) return;
  const size_t win = std::min(n, (size_t)K * 8);
  std::vector<u64> key(win);
  for (size_t i = 0; i < win; ++i) {
    const PendingSrc &p = s->pending[i];
    key[i] = s->h_key[(size_t)s->queries[p.qid].src[p.idx]];
  }
  std::vector<size_t> idx(win);
  std::iota(idx.begin(), idx.end(), (size_t)0);
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return key[a] > key[b]; });
  const size_t pos = (size_t)(std::find(idx.begin(), idx.end(), (size_t)0) - idx.begin());
  size_t start = pos >= (size_t)K / 2 ? pos - (size_t)K / 2 : 0;
  if (start + (size_t)K > win) start = win - (size_t)K;
  std::vector<char> take(win, 0);
  for (size_t i = start; i < start + (size_t)K; ++i) take[idx[i]] = 1;
  std::deque<PendingSrc> next;
  for (size_t i = 0; i < win; ++i)
    if (take[i]) next.push_back(s->pending[i]);
  for (size_t i = 0; i < win; ++i)
    if (!take[i]) next.push_back(s->pending[i]);
  for (size_t i = win; i < n; ++i) next.push_back(s->pending[i]);
  s->pending.swap(next);
}

// Starts a cohort on an idle context (caller holds the lock): level 0 holds
// up to c.K waiting sources; its levels then run on the device.
int start_cohort(ibfs_server_s *s, SCtx &c) {
  // the narrowest width (from 64 lanes) that holds the waiting sources:
  // narrower words make every pass of a small cohort cheaper
  c.K = 64;
  while (c.K < s->K && (size_t)c.K < s->pending.size()) c.K *= 2;
  c.W = c.K / 64;
  pick_cohort(s, c.K);
  const int W = c.W;
  for (auto &l : c.lanes) l = Lane{};
  // S_0, the changed flags and the control blocks zeroed in one launch
  HIP_CHK(launch_zero3_u64(slot(s, c, 0), s->g->Vt * W, reinterpret_cast<u64 *>(c.flags),
                           (s->g->Vt + 7) / 8, reinterpret_cast<u64 *>(c.ctl),
                           (int64_t)s->R * CTL_WORDS, c.st));
  // the cohort's first sources take lanes 0, 1, ...; with phase offsets a
  // lane whose source's two-hop size says it runs ahead of the others starts
  // at a later level (admitted then through the admission list)
  std::memset(c.h_all, 0, sizeof(u64) * W);
  int n = 0;
  for (int k = 0; k < c.K && !s->pending.empty(); ++k) {
    const PendingSrc p = s->pending.front();
    s->pending.pop_front();
    Query &q = s->queries[p.qid];
    c.lanes[k].state = ACTIVE;
    c.lanes[k].qid = p.qid;
    c.lanes[k].admit = 0;
    c.h_adm[n] = s->g->h_cid[q.src[p.idx]];
    c.h_adm[s->K + n] = k;
    c.h_lanes[k].out = q.out + (size_t)p.idx * (size_t)s->g->V;
    c.h_lanes[k].admit = 0;
    c.h_lanes[k].src = q.src[p.idx];
    if (q.t_admit < 0) q.t_admit = now_ms();
    ++n;
  }
  s->lanes_admitted += n;
  c.deferred.clear();
  c.a_max = 0;
  if (s->cfg.phase_offsets > 0 && n > 1 && !s->h_key.empty()) {
    std::vector<double> key((size_t)n);
    for (int i = 0; i < n; ++i) key[(size_t)i] = (double)s->h_key[(size_t)c.h_lanes[i].src];
    std::vector<int> start;
    const double avg_deg = s->g->Vt > 0 ? (double)s->g->E / (double)s->g->Vt : 2.0;
    phase_starts(key, avg_deg, s->cfg.phase_offsets, start);
    for (int i = 0; i < n; ++i) {
      const int a = start[(size_t)i];
      if (a == 0) continue;
      c.lanes[i].admit = a;
      c.h_lanes[i].admit = a;
      c.h_adm[i] = -1;  // not set at level 0
      c.deferred.emplace_back(i, a);
      c.a_max = std::max(c.a_max, a);
    }
  }
  for (int i = 0; i < n; ++i)
    if (c.lanes[i].admit == 0) set_bit(c.h_all, i);
  if (n > 0) {
    HIP_CHK(hipMemcpyAsync(c.d_adm, c.h_adm, sizeof(int32_t) * 2 * (size_t)s->K,
                           hipMemcpyHostToDevice, c.st));
    HIP_CHK(hipMemcpyAsync(c.d_lanes, c.h_lanes, sizeof(LaneOut) * (size_t)s->K,
                           hipMemcpyHostToDevice, c.st));
    HIP_CHK(launch_admit(c.W, c.d_adm, c.d_adm + s->K, n, slot(s, c, 0), c.flags, c.st));
  }
  HIP_CHK(hipMemcpyAsync(c.d_all, c.h_all, sizeof(u64) * W, hipMemcpyHostToDevice, c.st));
  plan_init(c.h_plan, s->cfg, s->hc.rho_q10, (u64)s->g->V);
  c.h_plan->hold = c.a_max;
  HIP_CHK(hipMemcpyAsync(c.d_plan, c.h_plan, sizeof(Plan), hipMemcpyHostToDevice, c.st));
  HIP_CHK(launch_inspect(W, plan_args(s, c), c.st));
  HIP_CHK(launch_plan_decide(c.d_plan, c.ctl, s->R, c.hs, W, split_hubs(s->g, s->hc, s->cfg), c.st));
  c.coh = 1;
  c.dlevel = 1;
  c.last_admit = 0;
  c.base = 0;
  c.running = true;
  return cohort_burst(s, c, true);
}

// A cohort's event completed: more levels, the emission of all its lanes once
// the plan reports that no vertex changed, or (after the emission) completion
// of its queries.
int cohort_step(ibfs_server_s *s, SCtx &c) {
  if (c.coh == 1) {
    if (c.h_plan->done) {
      std::vector<int> act;
      for (int k = 0; k < c.K; ++k)
        if (c.lanes[k].state == ACTIVE) act.push_back(k);
      int rc = emit_lanes(s, c, act, c.base, c.h_plan->last, 0);
      if (rc) return rc;
      HIP_CHK(hipEventRecord(c.ev, c.st));
      c.coh = 2;
      return IBFS_OK;
    }
    c.dlevel = c.h_plan->level;
    // a BFS admitted at level a ends by level a + Vt + 1
    if (c.dlevel > c.last_admit + s->g->Vt + 2)
      return fail(IBFS_ERR_STATE, "cohort exceeded V+2 levels after its last admission");
    return cohort_burst(s, c, false);
  }
  std::lock_guard<std::mutex> lk(s->mu);
  for (auto &l : c.lanes) {
    if (l.state != ACTIVE) continue;
    Query &q = s->queries[l.qid];
    if (--q.remaining == 0) c.done_pending.push_back(l.qid);
    l = Lane{};
  }
  mark_done(s, c);
  c.coh = 0;
  c.running = false;
  return IBFS_OK;
}

void worker_main(ibfs_server_s *s) {
  (void)hipSetDevice(s->g->device);
  while (true) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // pending sources start an idle context at once (a cohort that ended
      // leaves its context idle while sources still wait)
      bool busy = !s->pending.empty();
      for (auto &c : s->ctx) busy |= c.running;
      if (!busy)
        s->cv_work.wait(lk, [&] { return s->stop || !s->incoming.empty(); });
      if (s->stop) break;
      while (!s->incoming.empty()) {
        s->pending.push_back(s->incoming.front());
        s->incoming.pop_front();
      }
    }
    bool progressed = false;
    for (auto &c : s->ctx) {
      int rc = IBFS_OK;
      if (!c.running) {
        bool have;
        {
          std::lock_guard<std::mutex> lk(s->mu);
          have = !s->pending.empty();
          if (have) {
            c.clk.restart();
            rc = s->host_levels ? start_ctx(s, c) : start_cohort(s, c);
          }
        }
        progressed |= have;
      } else {
        hipError_t q = hipEventQuery(c.ev);
        if (q == hipErrorNotReady) {
          if (!c.clk.expired()) continue;
          rc = fail(IBFS_ERR_STATE, "server context did not progress within the wait limit (%.1f s)",
                    wait_timeout_s());
        } else if (q != hipSuccess) {
          rc = fail(IBFS_ERR_HIP, "server event: %s", hipGetErrorString(q));
        } else {
          c.clk.restart();
          rc = c.coh ? cohort_step(s, c) : process_level(s, c);
        }
        progressed = true;
      }
      if (rc) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->error = rc;
        s->error_msg = last_error();
        s->stop = true;
        s->cv_done.notify_all();
        return;
      }
    }
    if (!progressed) std::this_thread::yield();
  }
}

}  // namespace

extern "C" {

int ibfs_server_create(ibfs_graph_t g, const ibfs_config *cfg_in, int64_t status_bytes,
                       ibfs_server_t *out) {
  if (!g || !out) return fail(IBFS_ERR_ARG, "ibfs_server_create: null argument");
  ibfs_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else ibfs_config_default(&cfg);
  if (cfg.lane_bits != 64 && cfg.lane_bits != 128 && cfg.lane_bits != 256 && cfg.lane_bits != 512)
    return fail(IBFS_ERR_ARG, "lane_bits must be 64, 128, 256 or 512");
  if (cfg.ring_levels < 8 || cfg.ring_levels > 255)
    return fail(IBFS_ERR_ARG, "serving needs ring_levels in [8, 255]");
  if (cfg.td_chunk < 256 || (cfg.td_chunk & (cfg.td_chunk - 1)))
    return fail(IBFS_ERR_ARG, "td_chunk must be a power of two >= 256");
  if (cfg.emit_mode < 0 || cfg.emit_mode > 3) return fail(IBFS_ERR_ARG, "bad emit_mode");
  if (!valid_inspect_vpt(cfg.inspect_vpt)) return fail(IBFS_ERR_ARG, "bad inspect_vpt");
  if (!valid_wg_size(cfg.wg_size)) return fail(IBFS_ERR_ARG, "wg_size must be 256, 512 or 1024");
  if (int rc = check_direction_config(cfg)) return rc;
  if (cfg.device != g->device) return fail(IBFS_ERR_ARG, "config device differs from the graph's");
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0)
    return fail(IBFS_ERR_NODEVICE, "no HIP device available");
  HIP_CHK(hipSetDevice(g->device));
  // admission capacity from the status-array budget: one context holds
  // lane_bits lanes in ring_levels snapshots of Vt * lane_bits / 8 bytes
  const size_t per_ctx =
      (size_t)cfg.ring_levels * (size_t)std::max<int64_t>(g->Vt, 1) * (size_t)(cfg.lane_bits / 8);
  int C = status_bytes > 0 ? (int)((size_t)status_bytes / per_ctx) : 1;
  if (C < 1)
    return fail(IBFS_ERR_NOMEM, "status budget %lld below one context (%zu bytes)",
                (long long)status_bytes, per_ctx);
  C = std::min(C, 16);
  auto *s = new ibfs_server_s;
  s->g = g;
  s->cfg = cfg;
  s->K = cfg.lane_bits;
  s->W = cfg.lane_bits / 64;
  s->R = cfg.ring_levels;
  s->C = C;
  s->chunk_shift = __builtin_ctz((unsigned)cfg.td_chunk);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device) != hipSuccess ||
      cus <= 0)
    cus = 256;
  s->grid = cfg.dd_grid > 0 ? cfg.dd_grid : cus * 32;
  s->hc = hub_cut(s->g, cfg.hybrid_hubs);
  {  // two-hop size of every vertex, for cohort composition
    const int64_t V = g->V;
    std::vector<int32_t> all((size_t)V);
    std::iota(all.begin(), all.end(), 0);
    int32_t *d_all = nullptr;
    u64 *d_key = nullptr;
    hipError_t err = hipMalloc((void **)&d_all, sizeof(int32_t) * (size_t)std::max<int64_t>(V, 1));
    if (err == hipSuccess) err = hipMalloc((void **)&d_key, sizeof(u64) * (size_t)std::max<int64_t>(V, 1));
    if (err == hipSuccess) err = hipMemcpy(d_all, all.data(), sizeof(int32_t) * (size_t)V, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = launch_two_hop(g->d_beg, g->d_adj, g->d_deg, g->d_cid, d_all, V, d_key, nullptr);
    if (err == hipSuccess) {
      s->h_key.resize((size_t)V);
      err = hipMemcpy(s->h_key.data(), d_key, sizeof(u64) * (size_t)V, hipMemcpyDeviceToHost);
    }
    if (d_all) (void)hipFree(d_all);
    if (d_key) (void)hipFree(d_key);
    if (err != hipSuccess) {
      delete s;
      return fail(IBFS_ERR_HIP, "server key pass failed: %s", hipGetErrorString(err));
    }
  }
  s->ctx.resize(C);
  for (auto &c : s->ctx) {
    int rc = alloc_ctx(s, c);
    if (rc) {
      for (auto &x : s->ctx) release_ctx(x);
      delete s;
      return rc;
    }
  }
  s->worker = std::thread(worker_main, s);
  *out = s;
  return IBFS_OK;
}

int ibfs_server_set_admission(ibfs_server_t s, int32_t window_levels) {
  if (!s) return fail(IBFS_ERR_ARG, "null server");
  if (window_levels < 0) return fail(IBFS_ERR_ARG, "window_levels must be >= 0");
  std::lock_guard<std::mutex> lk(s->mu);
  s->admit_window = window_levels;
  return IBFS_OK;
}

int ibfs_server_set_levels(ibfs_server_t s, int32_t host_levels) {
  if (!s) return fail(IBFS_ERR_ARG, "null server");
  if (host_levels != 0 && host_levels != 1) return fail(IBFS_ERR_ARG, "host_levels must be 0 or 1");
  std::lock_guard<std::mutex> lk(s->mu);
  s->host_levels = host_levels;
  return IBFS_OK;
}

int ibfs_server_counters(ibfs_server_t s, int64_t *levels, int64_t *admitted,
                         int64_t *admitted_in_flight) {
  if (!s) return fail(IBFS_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (levels) *levels = s->levels_issued;
  if (admitted) *admitted = s->lanes_admitted;
  if (admitted_in_flight) *admitted_in_flight = s->lanes_admitted_in_flight;
  return IBFS_OK;
}

int ibfs_server_contexts(ibfs_server_t s, int32_t *contexts, int32_t *lanes) {
  if (!s) return fail(IBFS_ERR_ARG, "null server");
  if (contexts) *contexts = s->C;
  if (lanes) *lanes = s->C * s->K;
  return IBFS_OK;
}

int ibfs_server_submit(ibfs_server_t s, const int32_t *src, int32_t n, int32_t *depth_out,
                       int64_t *id) {
  if (!s || !src || !depth_out || !id || n < 1) return fail(IBFS_ERR_ARG, "ibfs_server_submit: bad arguments");
  for (int i = 0; i < n; ++i)
    if (src[i] < 0 || src[i] >= s->g->V)
      return fail(IBFS_ERR_ARG, "source %d = %d out of range", i, src[i]);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->error) return fail(s->error, "server stopped: %s", s->error_msg.c_str());
  Query q;
  q.id = s->next_id++;
  q.n = n;
  q.srcbuf.assign(src, src + n);
  q.out = depth_out;
  q.remaining = n;
  q.t_submit = now_ms();
  auto &ref = s->queries[q.id] = std::move(q);
  ref.src = ref.srcbuf.data();
  for (int i = 0; i < n; ++i) s->incoming.push_back(PendingSrc{ref.id, i});
  *id = ref.id;
  s->cv_work.notify_one();
  return IBFS_OK;
}

int ibfs_server_wait(ibfs_server_t s, int64_t id, double timeout_ms, ibfs_query_info *info) {
  if (!s) return fail(IBFS_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->queries.find(id);
  if (it == s->queries.end()) return fail(IBFS_ERR_ARG, "unknown query %lld", (long long)id);
  // a query released by another thread while this one waits ends the wait
  // (looked up, never inserted)
  auto ready = [&] {
    auto q = s->queries.find(id);
    return s->error || q == s->queries.end() || q->second.done;
  };
  if (timeout_ms < 0) s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
    return fail(IBFS_ERR_STATE, "query %lld not finished within %.1f ms", (long long)id, timeout_ms);
  if (s->error) return fail(s->error, "server stopped: %s", s->error_msg.c_str());
  it = s->queries.find(id);
  if (it == s->queries.end())
    return fail(IBFS_ERR_STATE, "query %lld was released while waiting", (long long)id);
  const Query &q = it->second;
  if (info) {
    info->id = q.id;
    info->n_sources = q.n;
    info->done = q.done;
    info->submit_ms = q.t_submit;
    info->admit_ms = q.t_admit;
    info->done_ms = q.t_done;
  }
  return IBFS_OK;
}

int ibfs_server_release(ibfs_server_t s, int64_t id) {
  if (!s) return fail(IBFS_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->queries.find(id);
  if (it == s->queries.end()) return fail(IBFS_ERR_ARG, "unknown query %lld", (long long)id);
  if (!it->second.done) return fail(IBFS_ERR_STATE, "query %lld still running", (long long)id);
  s->queries.erase(it);
  return IBFS_OK;
}

int ibfs_server_destroy(ibfs_server_t s) {
  if (!s) return IBFS_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->g->device);
  for (auto &c : s->ctx) release_ctx(c);
  delete s;
  return IBFS_OK;
}

}  // extern "C"
