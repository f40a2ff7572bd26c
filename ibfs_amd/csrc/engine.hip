// Engine: device graph, per-group traversal driver (level loop, direction
// switch, snapshot ring, depth emission) and GroupBy, behind the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
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

struct ibfs_engine_s {
  ibfs_graph_s *g = nullptr;
  ibfs_config cfg{};
  // K: lanes per group (lane_bits); W: status words per vertex of the current
  // traversal, the narrowest power of two that holds its sources (at most
  // Wmax = lane_bits / 64, which sizes the buffers)
  int K = 128, W = 2, Wmax = 2, R = 32, chunk_shift = 12;
  int hy_grid = 8192;  // workgroups of a hybrid expansion
  HubCut hc;           // hubs a hybrid level pulls through
  hipStream_t st = nullptr;
  hipEvent_t ev_wait = nullptr;  // recorded after each host copy the loop waits for
  u64 *ring = nullptr;
  LevelCtl *d_ctl = nullptr;
  LevelCtl *h_ctl = nullptr;  // pinned
  CtlPost *post = nullptr;    // coherent pinned: control blocks posted by k_post_ctl
  u64 post_seq = 0;
  Queues q{};
  u64 *d_masks = nullptr;     // [0,W) lanes_all, [W,2W) sel, [2W, 2W+(R+1)W) level masks
  u64 *h_masks = nullptr;     // pinned mirror
  int32_t *d_src = nullptr;
  int32_t *h_src = nullptr;   // pinned, K entries
  LaneOut *d_lanes = nullptr;
  LaneOut *h_lanes = nullptr; // pinned, K entries
  int32_t *d_out = nullptr;   // staging for host-placed output
  size_t d_out_cap = 0;
  u64 *d_count = nullptr;
  u64 *h_count = nullptr;
  u64 *hs = nullptr;          // hub cache, H * W words
  uint8_t *flags = nullptr;   // changed-at-this-level byte per traversal vertex
  int emit_vec4 = 0;          // every lane output of this traversal is 16-byte aligned
  std::vector<ibfs_level_info> log;
  std::vector<hipEvent_t> ev;  // event pool
  size_t ev_used = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> emit_ev;  // emission timings of a traversal
  // a traversal failed with work possibly still enqueued on st (a wait past
  // the limit): the next one drains it first, since its set-up rewrites the
  // pinned blocks that work may still copy into
  bool unsettled = false;
};

namespace {

size_t masks_words(const ibfs_engine_s *e) { return (size_t)(2 + (e->R + 1)) * e->W; }

// Waits for the work enqueued so far by polling an event recorded behind it
// (the end of a group: its edge count and timing events).  A blocking wait's
// wake-up costs tens of microseconds, and an event query reads one completion
// signal where a stream query takes the stream's lock.  The level loop waits
// on posted control blocks instead (wait_post).
// Past the wait limit (ibfs_set_wait_timeout) it fails with IBFS_ERR_STATE.
int poll_stream(hipStream_t st, hipEvent_t ev) {
  const hipError_t r = hipEventRecord(ev, st);
  if (r != hipSuccess) return fail(IBFS_ERR_HIP, "hipEventRecord: %s", hipGetErrorString(r));
  WaitClock clk;
  for (;;) {
    const hipError_t q = hipEventQuery(ev);
    if (q == hipSuccess) return IBFS_OK;
    if (q != hipErrorNotReady) return fail(IBFS_ERR_HIP, "stream error: %s", hipGetErrorString(q));
    if (clk.expired())
      return fail(IBFS_ERR_STATE, "group did not finish within the wait limit (%.1f s)",
                  wait_timeout_s());
  }
}

// Waits for the control block posted with tag seq (a spin on the tag; every
// 1,024 reads the stream is queried, so an error, or an idle stream without
// the tag, ends the wait with an error), then copies it to e->h_ctl.
int wait_post(ibfs_engine_s *e, u64 seq) {
  WaitClock clk;
  for (unsigned i = 1;; ++i) {
    if (__atomic_load_n(&e->post->tag, __ATOMIC_ACQUIRE) == seq) break;
    if ((i & 1023) == 0) {
      const hipError_t q = hipStreamQuery(e->st);
      if (q == hipErrorNotReady) {
        if (clk.expired())
          return fail(IBFS_ERR_STATE, "level did not finish within the wait limit (%.1f s)",
                      wait_timeout_s());
        continue;
      }
      if (q != hipSuccess) return fail(IBFS_ERR_HIP, "stream error: %s", hipGetErrorString(q));
      if (__atomic_load_n(&e->post->tag, __ATOMIC_ACQUIRE) == seq) break;
      return fail(IBFS_ERR_STATE, "control block post did not arrive");
    }
  }
  std::memcpy(e->h_ctl, &e->post->c, sizeof(LevelCtl));
  return IBFS_OK;
}

// Elapsed time between two events the device has passed.  A posted control
// block can reach the host before the runtime has seen the completion of the
// events enqueued ahead of it, so each is queried until it reports completion.
// Past the wait limit it returns hipErrorNotReady (reported by HIP_TRY).
hipError_t elapsed_ms(float *ms, hipEvent_t a, hipEvent_t b) {
  for (hipEvent_t x : {a, b}) {
    hipError_t q;
    WaitClock clk;
    while ((q = hipEventQuery(x)) == hipErrorNotReady) {
      if (clk.expired()) return hipErrorNotReady;
    }
    if (q != hipSuccess) return q;
  }
  return hipEventElapsedTime(ms, a, b);
}

int post_and_wait(ibfs_engine_s *e, const LevelCtl *d) {
  const u64 seq = ++e->post_seq;
  HIP_TRY(launch_post_ctl(d, e->post, seq, e->st));
  return wait_post(e, seq);
}

hipEvent_t next_event(ibfs_engine_s *e) {
  if (e->ev_used == e->ev.size()) {
    hipEvent_t x;
    if (hipEventCreate(&x) != hipSuccess) return nullptr;
    e->ev.push_back(x);
  }
  return e->ev[e->ev_used++];
}

int check_device(int dev) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(IBFS_ERR_NODEVICE, "no HIP device available");
  if (dev < 0 || dev >= n) return fail(IBFS_ERR_ARG, "device %d out of range [0,%d)", dev, n);
  if (hipSetDevice(dev) != hipSuccess) return fail(IBFS_ERR_HIP, "hipSetDevice(%d) failed", dev);
  return IBFS_OK;
}

// One non-blocking stream per device for the library's synchronous copies
// (created on first use, kept until the process ends).
int copy_stream(int dev, hipStream_t *out) {
  static std::mutex mu;
  static std::vector<hipStream_t> streams;
  std::lock_guard<std::mutex> lk(mu);
  if (dev < 0) return fail(IBFS_ERR_ARG, "bad device %d", dev);
  if ((size_t)dev >= streams.size()) streams.resize((size_t)dev + 1, nullptr);
  if (!streams[(size_t)dev]) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    HIP_TRY(hipSetDevice(dev));
    const hipError_t r = hipStreamCreateWithFlags(&streams[(size_t)dev], hipStreamNonBlocking);
    (void)hipSetDevice(cur);
    if (r != hipSuccess) {
      streams[(size_t)dev] = nullptr;
      return fail(IBFS_ERR_HIP, "copy stream: %s", hipGetErrorString(r));
    }
  }
  *out = streams[(size_t)dev];
  return IBFS_OK;
}

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

void engine_release(ibfs_engine_s *e) {
  if (!e) return;
  if (e->g) (void)hipSetDevice(e->g->device);
  if (e->st) (void)hipStreamSynchronize(e->st);
  dfree(e->ring);
  dfree(e->d_ctl);
  hfree(e->h_ctl);
  hfree(e->post);
  for (int i = 0; i < 6; ++i) dfree(e->q.q[i]);
  dfree(e->q.chunks);
  dfree(e->d_masks);
  hfree(e->h_masks);
  dfree(e->d_src);
  hfree(e->h_src);
  dfree(e->d_lanes);
  hfree(e->h_lanes);
  dfree(e->d_out);
  dfree(e->d_count);
  hfree(e->h_count);
  dfree(e->hs);
  dfree(e->flags);
  for (auto x : e->ev) (void)hipEventDestroy(x);
  e->ev.clear();
  if (e->ev_wait) (void)hipEventDestroy(e->ev_wait);
  e->ev_wait = nullptr;
  if (e->st) (void)hipStreamDestroy(e->st);
  e->st = nullptr;
}

u64 *slot(ibfs_engine_s *e, int L) {
  return e->ring + (size_t)(L % e->R) * (size_t)e->g->Vt * (size_t)e->W;
}

EmitSpace emit_space(const ibfs_graph_s *g) { return EmitSpace{g->d_cid, g->V, g->Vt, g->d_deg}; }

// Distinct sources without a status word (isolated vertices): they change
// at their admission level like any source, but outside the status words.
int64_t isolated_sources(const ibfs_graph_s *g, const int32_t *src, int k) {
  std::vector<int32_t> iso;
  for (int i = 0; i < k; ++i)
    if (g->h_cid[src[i]] < 0) iso.push_back(src[i]);
  std::sort(iso.begin(), iso.end());
  return (int64_t)(std::unique(iso.begin(), iso.end()) - iso.begin());
}

// direction of level L (0 top-down, 1 bottom-up, 2 hybrid) from the
// counters of L-1
int decide(const ibfs_engine_s *e, int L, const LevelCtl &c, int dir_prev, u64 nf_prev) {
  const ibfs_config &k = e->cfg;
  return choose_direction(k.policy, k.sw_level, k.alpha, k.beta, k.hybrid_gamma, e->hc.rho_q10,
                          (u64)e->g->V, L, dir_prev, nf_prev, c.n_front, c.m_front, c.m_cand,
                          c.m_front_nh);
}

// Host-driven inspector arguments (the pass also refreshes the hub cache
// from cur for the next bottom-up level).
InspectArgs inspect_args(ibfs_engine_s *e, const u64 *prev, const u64 *cur, u64 *next,
                         const u64 *lam_prev, LevelCtl *ctl, LevelCtl *ctl_zero, int full_diff) {
  InspectArgs a{};
  a.full_diff = full_diff;
  a.prev = prev;
  a.cur = cur;
  a.next = next;
  a.lam_prev = lam_prev;
  a.ctl = ctl;
  a.ctl_zero = ctl_zero;
  a.deg = e->g->d_deg;
  a.V = e->g->Vt;
  a.q = e->q;
  a.sml = e->cfg.sml_shed;
  a.lrg = e->cfg.lrg_shed;
  a.chunk_shift = e->chunk_shift;
  a.hubs = e->g->d_hubs;
  a.H = e->g->H;
  a.hs = e->hs;
  a.flags = e->flags;
  a.vpt = e->cfg.inspect_vpt;
  a.nw = e->cfg.wg_size / 64;
  a.hsplit = HubSplit{e->hc.deg, e->hc.tie, e->hc.cut};
  return a;
}

// Emits the window [b, en] for the lanes in e->h_lanes (already on device).
// edges (final window only, or null): the pass also adds the TEPS numerator.
int emit_window(ibfs_engine_s *e, int b, int en, int k, ibfs_stats *stats, u64 *edges) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e->cfg.kernel_timing) {
    e0 = next_event(e);
    HIP_TRY(hipEventRecord(e0, e->st));
  }
  HIP_TRY(launch_emit(e->W, e->cfg.emit_mode, e->emit_vec4, e->ring, e->R, emit_space(e->g), b,
                      en, e->d_lanes, e->d_masks + 2 * e->W, e->d_masks + e->W, k, edges,
                      e->st));
  if (e->cfg.kernel_timing) {
    e1 = next_event(e);
    HIP_TRY(hipEventRecord(e1, e->st));
    e->emit_ev.emplace_back(e0, e1);  // read after the traversal's last wait
  }
  if (stats) stats->n_emit_launches++;
  return IBFS_OK;
}

// One group of k sources (lanes 0..k-1), depths into dout (device, k*V).
int traverse(ibfs_engine_s *e, const int32_t *src, int k, int32_t *dout, ibfs_stats *stats) {
  ibfs_graph_s *g = e->g;
  const int64_t V = g->V, Vt = g->Vt;
  // a partial group streams only the status words its lanes need
  e->W = 1;
  while (e->W * 64 < k && e->W < e->Wmax) e->W *= 2;
  const int W = e->W;
  e->ev_used = 0;
  e->emit_ev.clear();
  e->log.clear();
  clear_hip_error();

  // lanes, masks, sources (pinned staging -> device)
  std::memset(e->h_masks, 0, sizeof(u64) * masks_words(e));
  for (int i = 0; i < k; ++i) {
    e->h_masks[i >> 6] |= 1ull << (i & 63);
    e->h_src[i] = g->h_cid[src[i]];
    e->h_lanes[i].out = dout + (size_t)i * (size_t)V;
    e->h_lanes[i].admit = 0;
    e->h_lanes[i].src = src[i];
  }
  e->emit_vec4 = (V % 4 == 0) && ((uintptr_t)dout % 16 == 0);
  for (int w = 0; w < W; ++w) e->h_masks[W + w] = e->h_masks[w];
  for (size_t i = 0; i < (size_t)(e->R + 1) * W; ++i) e->h_masks[2 * W + i] = ~0ull;
  HIP_TRY(hipMemcpyAsync(e->d_masks, e->h_masks, sizeof(u64) * masks_words(e),
                         hipMemcpyHostToDevice, e->st));
  HIP_TRY(hipMemcpyAsync(e->d_src, e->h_src, sizeof(int32_t) * k, hipMemcpyHostToDevice, e->st));
  HIP_TRY(hipMemcpyAsync(e->d_lanes, e->h_lanes, sizeof(LaneOut) * k, hipMemcpyHostToDevice,
                         e->st));

  const bool timing = e->cfg.kernel_timing != 0;
  hipEvent_t t_begin = next_event(e);
  HIP_TRY(hipEventRecord(t_begin, e->st));

  // level 0: S_0 = source bits; inspection builds the level-1 queues
  HIP_TRY(launch_zero3_u64(slot(e, 0), Vt * W, reinterpret_cast<u64 *>(e->flags), (Vt + 7) / 8,
                           reinterpret_cast<u64 *>(e->d_ctl), (int64_t)e->R * CTL_WORDS, e->st));
  HIP_TRY(launch_init_sources(W, e->d_src, k, 0, slot(e, 0), e->flags, e->st));
  hipEvent_t ti0 = nullptr, ti1 = nullptr;
  if (timing) { ti0 = next_event(e); HIP_TRY(hipEventRecord(ti0, e->st)); }
  HIP_TRY(launch_inspect(W, inspect_args(e, nullptr, slot(e, 0), slot(e, 1), e->d_masks,
                                         &e->d_ctl[0], &e->d_ctl[1 % e->R], 1),
                         e->st));
  if (timing) { ti1 = next_event(e); HIP_TRY(hipEventRecord(ti1, e->st)); }
  {
    const int rc = post_and_wait(e, &e->d_ctl[0]);
    if (rc) return rc;
  }
  auto record = [&](int L, int dir, const LevelCtl &c, float exp_ms, float insp_ms) {
    ibfs_level_info li{};
    li.level = L;
    li.direction = dir;
    li.n_changed = (int64_t)c.n_changed;
    li.n_frontier = (int64_t)c.n_front;
    li.m_frontier = (int64_t)c.m_front;
    li.n_unvisited = (int64_t)c.n_cand;
    li.m_unvisited = (int64_t)c.m_cand;
    for (int i = 0; i < 3; ++i) {
      li.q_td[i] = c.q[i];
      li.q_bu[i] = c.q[3 + i];
    }
    li.expand_ms = exp_ms;
    li.inspect_ms = insp_ms;
    li.m_frontier_nonhub = (int64_t)c.m_front_nh;
    e->log.push_back(li);
  };
  float insp0 = 0;
  if (timing) {
    HIP_TRY(elapsed_ms(&insp0, ti0, ti1));
    if (stats) stats->inspect_ms += insp0;
  }
  if (stats) stats->n_inspect_launches++;
  {
    LevelCtl c0 = *e->h_ctl;
    c0.n_changed += (u64)isolated_sources(g, src, k);
    record(0, -1, c0, 0.f, insp0);
  }

  // A level's event timings and log entry are settled after the next level's
  // kernels are enqueued, so the host does that bookkeeping while they run
  // instead of between the control-block wait and the next launch.
  struct Pending {
    bool on = false;
    int L = 0, dir = 0;
    hipEvent_t x0 = nullptr, x1 = nullptr, y0 = nullptr, y1 = nullptr;
    LevelCtl c{};
    int64_t bu_work = 0, bu_prefix = 0;
  } pend;
  auto settle = [&]() -> int {
    if (!pend.on) return IBFS_OK;
    pend.on = false;
    float ems = 0, ims = 0;
    if (timing) {
      HIP_TRY(elapsed_ms(&ems, pend.x0, pend.x1));
      HIP_TRY(elapsed_ms(&ims, pend.y0, pend.y1));
      if (stats) {
        (pend.dir == 0 ? stats->expand_td_ms
                       : pend.dir == 2 ? stats->expand_hy_ms : stats->expand_bu_ms) += ems;
        stats->inspect_ms += ims;
      }
    }
    record(pend.L, pend.dir, pend.c, ems, ims);
    if (e->cfg.work_model && pend.dir == 1) {
      e->log.back().bu_work = pend.bu_work;
      e->log.back().bu_prefix = pend.bu_prefix;
    }
    return IBFS_OK;
  };

  int base = 0, last = 0, dir_prev = -1;
  u64 nf_prev = ~0ull;
  for (int L = 1;; ++L) {
    const LevelCtl c = *e->h_ctl;  // counters of level L-1
    if (c.n_changed == 0) {        // only possible at L == 1 with no sources
      last = L - 1;
      break;
    }
    if (L > 1 << 20) return fail(IBFS_ERR_STATE, "level limit exceeded");
    const int dir = decide(e, L, c, dir_prev, nf_prev);
    const bool tdf = dir == 0 && flag_top_down(c.m_front, (u64)g->V);
    // lam of level L-1 is the first member of its control block
    const u64 *lamp = reinterpret_cast<const u64 *>(e->d_ctl + (L - 1) % e->R);
    hipEvent_t x0 = nullptr, x1 = nullptr, y1 = nullptr;
    if (timing) { x0 = next_event(e); HIP_TRY(hipEventRecord(x0, e->st)); }
    if (dir == 0) {
      HIP_TRY(launch_expand_td(W, g->d_beg, g->d_adj, slot(e, L - 1), slot(e, L), lamp, e->q,
                               c.q[Q_TD_S], c.q[Q_TD_M], c.q[Q_TD_L], e->chunk_shift,
                               tdf ? e->flags : nullptr, e->st));
      if (stats) stats->n_td_launches++;
    } else if (dir == 2) {
      HIP_TRY(launch_expand_hy(W, g->d_beg, g->d_adj, g->d_tadj, slot(e, L - 1), e->hs,
                               slot(e, L), &e->d_ctl[(L - 1) % e->R], e->q, e->chunk_shift,
                               HubSplit{e->hc.deg, e->hc.tie, e->hc.cut}, e->hy_grid, e->st));
      if (stats) stats->n_hy_launches++;
    } else {
      if (e->cfg.work_model) {  // instrumentation: exact early-termination prefixes
        HIP_TRY(hipMemsetAsync(e->d_count + 1, 0, 2 * sizeof(u64), e->st));
        HIP_TRY(launch_bu_model(W, g->d_beg, g->d_tadj, slot(e, L - 1), e->hs, lamp, Vt,
                                e->d_count + 1, e->st));
        HIP_TRY(hipMemcpyAsync(e->h_count + 1, e->d_count + 1, 2 * sizeof(u64),
                               hipMemcpyDeviceToHost, e->st));
        if (timing) HIP_TRY(hipEventRecord(x0, e->st));
      }
      HIP_TRY(launch_expand_bu(W, g->d_beg, g->d_tadj, slot(e, L - 1), e->hs, slot(e, L),
                               e->flags, lamp, e->q, c.q[Q_BU_S], c.q[Q_BU_M], c.q[Q_BU_L],
                               e->st));
      if (stats) stats->n_bu_launches++;
    }
    if (timing) { x1 = next_event(e); HIP_TRY(hipEventRecord(x1, e->st)); }
    // the inspection's start event is the expansion's end event unless a
    // window flush runs in between
    hipEvent_t y0 = x1;
    if ((L + 1) - base >= e->R) {  // the next copy would overwrite S_base
      int rc = emit_window(e, base, L, k, stats, nullptr);
      if (rc) return rc;
      if (stats) stats->emit_windows++;
      base = L;
      if (timing) { y0 = next_event(e); HIP_TRY(hipEventRecord(y0, e->st)); }
    }
    HIP_TRY(launch_inspect(W, inspect_args(e, slot(e, L - 1), slot(e, L), slot(e, L + 1), lamp,
                                           &e->d_ctl[L % e->R], &e->d_ctl[(L + 1) % e->R],
                                           (dir == 1 || tdf) ? 0 : 1),
                           e->st));
    if (timing) { y1 = next_event(e); HIP_TRY(hipEventRecord(y1, e->st)); }
    const u64 seq = ++e->post_seq;
    HIP_TRY(launch_post_ctl(&e->d_ctl[L % e->R], e->post, seq, e->st));
    {
      const int rc = settle();  // level L-1, while level L runs
      if (rc) return rc;
    }
    {
      const int rc = wait_post(e, seq);
      if (rc) return rc;
    }
    if (stats) stats->n_inspect_launches++;
    pend.on = true;
    pend.L = L;
    pend.dir = dir;
    pend.x0 = x0;
    pend.x1 = x1;
    pend.y0 = y0;
    pend.y1 = y1;
    pend.c = *e->h_ctl;
    pend.bu_work = (int64_t)e->h_count[1];
    pend.bu_prefix = (int64_t)e->h_count[2];
    dir_prev = dir;
    nf_prev = c.n_front;
    if (pend.c.n_changed == 0) {
      last = L - 1;
      break;
    }
  }
  {
    const int rc = settle();
    if (rc) return rc;
  }
  // the final window also counts the TEPS numerator (over S_last, which its
  // pass reads anyway)
  bool counted = false;
  if (last >= base) {
    if (stats) HIP_TRY(hipMemsetAsync(e->d_count, 0, sizeof(u64), e->st));
    int rc = emit_window(e, base, last, k, stats, stats ? e->d_count : nullptr);
    if (rc) return rc;
    if (stats) stats->emit_windows++;
    counted = stats != nullptr;
  }
  // one wait for the group: its end event, the edge count and its copy
  hipEvent_t t_end = next_event(e);
  HIP_TRY(hipEventRecord(t_end, e->st));
  if (stats) {
    if (!counted) {  // no final window (the last flush ended at last + 1)
      HIP_TRY(hipMemsetAsync(e->d_count, 0, sizeof(u64), e->st));
      HIP_TRY(launch_count_edges(W, slot(e, last), g->d_deg, Vt, e->d_masks, e->d_count, e->st));
    }
    HIP_TRY(hipMemcpyAsync(e->h_count, e->d_count, sizeof(u64), hipMemcpyDeviceToHost, e->st));
  }
  if (int rc = poll_stream(e->st, e->ev_wait)) return rc;
  if (stats) {
    float ms = 0;
    HIP_TRY(elapsed_ms(&ms, t_begin, t_end));
    stats->total_ms += ms;
    for (const auto &x : e->emit_ev) {
      HIP_TRY(elapsed_ms(&ms, x.first, x.second));
      stats->emit_ms += ms;
    }
    stats->levels = std::max(stats->levels, last + 1);
    stats->traversed_edges += (int64_t)*e->h_count;
  }
  return IBFS_OK;
}

int ensure_out(ibfs_engine_s *e, size_t elems) {
  if (e->d_out_cap >= elems) return IBFS_OK;
  dfree(e->d_out);
  e->d_out_cap = 0;
  HIP_TRY(hipMalloc((void **)&e->d_out, elems * sizeof(int32_t)));
  e->d_out_cap = elems;
  return IBFS_OK;
}

}  // namespace

HubCut hub_cut(const ibfs_graph_s *g, int64_t r) {
  HubCut c;
  const int64_t H = g->H;
  if (r <= 0 || r > H) r = H;
  c.cut = (int32_t)r;
  if (r > 0) {
    c.deg = g->h_hub_deg[(size_t)r - 1];
    c.tie = g->h_hub_id[(size_t)r - 1];
  }
  c.rho_q10 = g->E > 0 ? (int32_t)((g->h_hub_entries[(size_t)r] * 1024) / g->E) : 0;
  return c;
}

extern "C" {

int ibfs_config_default(ibfs_config *cfg) {
  if (!cfg) return fail(IBFS_ERR_ARG, "null config");
  std::memset(cfg, 0, sizeof *cfg);
  cfg->lane_bits = 128;
  // wave64: a vertex with fewer than 64 edges would leave lanes of its wave
  // idle, so it takes the one-thread class (the reference's 32 matches a
  // 32-wide warp); measured 3% faster on R-MAT s22
  cfg->sml_shed = 64;
  cfg->lrg_shed = 512;
  cfg->policy = IBFS_POLICY_ALPHA_BETA;
  cfg->sw_level = 2;
  // With 128 concurrent lanes a bottom-up vertex rarely covers every active
  // lane early, so top-down stays cheaper far longer than for one source:
  // switch when the frontier's edges exceed half of the candidates' edges.
  cfg->alpha = 2;
  cfg->beta = 48;
  // A top-down level whose frontier edges mostly leave the top 4,096 hubs
  // runs hybrid: the candidates pull through those hubs (a pulled entry
  // costs about a quarter of a pushed edge), the rest of the frontier pushes.
  // Measured on R-MAT s22: plain 7.5K -> 7.85K GTEPS (128 lanes), batched
  // 9.7K -> 10.4K.
  cfg->hybrid_gamma = 16;
  cfg->hybrid_hubs = 4096;
  // batched runs group sources by two-hop size: lanes that start slowly (a
  // degree-1 source next to low-degree vertices) no longer keep every
  // bottom-up candidate of a fast group scanning its whole list (R-MAT s22,
  // 256 lanes x 2 contexts: 10.3K -> 11.7K GTEPS)
  cfg->pack_sources = 1;
  cfg->split_cov = 0;
  cfg->phase_offsets = 0;
  // batched runs: the last levels (frontier under V/32 vertices and E/32
  // edges) run as sparse tail levels (R-MAT s22, 1,024 sources in 256-lane
  // groups: 10.23 -> 10.00 ms per batch)
  cfg->tail_div = 32;
  cfg->wg_size = 512;
  cfg->td_chunk = 2048;
  cfg->ring_levels = 32;
  cfg->kernel_timing = 0;
  cfg->device = 0;
  cfg->emit_mode = 0;
  return IBFS_OK;
}

int ibfs_device_count(int *n) {
  if (!n) return fail(IBFS_ERR_ARG, "null argument");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return IBFS_OK;
}

int ibfs_device_alloc(int device, size_t bytes, void **dptr) {
  if (!dptr) return fail(IBFS_ERR_ARG, "null argument");
  int rc = check_device(device);
  if (rc) return rc;
  if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess)
    return fail(IBFS_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
  return IBFS_OK;
}

int ibfs_device_free(void *dptr) {
  if (dptr) HIP_TRY(hipFree(dptr));
  return IBFS_OK;
}

int ibfs_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  if (bytes && (!dst || !src)) return fail(IBFS_ERR_ARG, "ibfs_memcpy_d2h: null pointer");
  if (!bytes) return IBFS_OK;
  // on a library stream of the source's device, as ibfs_memcpy_d2d
  hipPointerAttribute_t pa{};
  HIP_TRY(hipPointerGetAttributes(&pa, src));
  int cur = 0;
  HIP_TRY(hipGetDevice(&cur));
  const int dev = pa.device >= 0 ? pa.device : cur;
  hipStream_t st = nullptr;
  if (int rc = copy_stream(dev, &st)) return rc;
  HIP_TRY(hipSetDevice(dev));
  const D2HPiece pc{dst, src, bytes};
  const int rc = copy_to_host(&pc, 1, st);
  (void)hipSetDevice(cur);
  return rc;
}

int ibfs_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return IBFS_OK;
}

int ibfs_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  if (bytes && (!dst || !src)) return fail(IBFS_ERR_ARG, "null device pointer");
  if (!bytes) return IBFS_OK;
  // on a library stream of the destination's device, waiting for that
  // stream only: engines, batches and servers on the device keep running
  hipPointerAttribute_t pa{};
  HIP_TRY(hipPointerGetAttributes(&pa, dst));
  int dev = pa.device;
  if (dev < 0) HIP_TRY(hipGetDevice(&dev));
  hipStream_t st = nullptr;
  if (int rc = copy_stream(dev, &st)) return rc;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  return IBFS_OK;
}

int ibfs_graph_opts_default(ibfs_graph_opts *o) {
  if (!o) return fail(IBFS_ERR_ARG, "null options");
  std::memset(o, 0, sizeof *o);
  o->hub_count = 131072;
  o->nthreads = 0;
  return IBFS_OK;
}

int ibfs_graph_create(const int64_t *beg, int64_t V, const int32_t *adj, int64_t E, int device,
                      ibfs_graph_t *out) {
  return ibfs_graph_create_ex(beg, V, adj, E, device, nullptr, out);
}

int ibfs_graph_create_ex(const int64_t *beg, int64_t V, const int32_t *adj, int64_t E, int device,
                         const ibfs_graph_opts *opts_in, ibfs_graph_t *out) {
  ibfs_graph_opts opts;
  if (opts_in) opts = *opts_in;
  else ibfs_graph_opts_default(&opts);
  if (opts.hub_count < 0) return fail(IBFS_ERR_ARG, "hub_count must be >= 0");
  if (!out || !beg || V < 1 || E < 0 || (E > 0 && !adj))
    return fail(IBFS_ERR_ARG, "ibfs_graph_create: bad arguments (V=%lld, E=%lld)", (long long)V,
                (long long)E);
  if (V > 0x7fffffffLL) return fail(IBFS_ERR_ARG, "vertex count exceeds int32 ids");
  if (beg[0] != 0 || beg[V] != E)
    return fail(IBFS_ERR_FORMAT, "beg_pos[0]=%lld, beg_pos[V]=%lld, expected 0 and %lld",
                (long long)beg[0], (long long)beg[V], (long long)E);
  std::vector<int32_t> deg((size_t)V);
  int64_t maxd = 0;
  for (int64_t v = 0; v < V; ++v) {
    const int64_t d = beg[v + 1] - beg[v];
    if (d < 0) return fail(IBFS_ERR_FORMAT, "beg_pos decreases at vertex %lld", (long long)v);
    if (d > 0x7fffffffLL) return fail(IBFS_ERR_FORMAT, "degree exceeds int32");
    deg[v] = (int32_t)d;
    maxd = std::max(maxd, d);
  }
  for (int64_t i = 0; i < E; ++i)
    if (adj[i] < 0 || adj[i] >= V)
      return fail(IBFS_ERR_FORMAT, "adjacency entry %lld = %d out of range", (long long)i, adj[i]);
  int rc = check_device(device);
  if (rc) return rc;
  int nthreads = opts.nthreads > 0 ? opts.nthreads : (int)std::thread::hardware_concurrency();
  nthreads = std::max(1, std::min(nthreads, 64));
  // Traversal vertex space: vertices with an out- or in-edge, original order.
  std::vector<uint8_t> has_in((size_t)V, 0);
  struct MarkCtx {
    const int32_t *adj;
    uint8_t *has_in;
    int64_t E;
    int nt;
  } mx{adj, has_in.data(), E, nthreads};
  parallel_for_threads(nthreads, [](int t, void *p) {
    MarkCtx &c = *(MarkCtx *)p;
    const int64_t lo = c.E * t / c.nt, hi = c.E * (t + 1) / c.nt;
    for (int64_t j = lo; j < hi; ++j) __atomic_store_n(&c.has_in[c.adj[j]], 1, __ATOMIC_RELAXED);
  }, &mx);
  std::vector<int32_t> cid((size_t)V, -1);
  int64_t Vt = 0;
  for (int64_t v = 0; v < V; ++v)
    if (deg[v] > 0 || has_in[v]) cid[v] = (int32_t)Vt++;
  std::vector<int64_t> tbeg((size_t)Vt + 1);
  std::vector<int32_t> tdeg((size_t)std::max<int64_t>(Vt, 1));
  for (int64_t v = 0; v < V; ++v)
    if (cid[v] >= 0) {
      tbeg[cid[v]] = beg[v];
      tdeg[cid[v]] = deg[v];
    }
  tbeg[Vt] = E;
  // Hubs: the H highest-degree traversal vertices (ties by id).  Traversal
  // adjacency: each list ordered by descending neighbour degree (ties by
  // id), hub neighbours encoded as 0x80000000 | hub slot.
  const int64_t H = std::min<int64_t>(opts.hub_count, Vt);
  std::vector<int32_t> order((size_t)Vt);
  for (int64_t v = 0; v < Vt; ++v) order[v] = (int32_t)v;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return tdeg[a] > tdeg[b]; });
  std::vector<int32_t> slot_of((size_t)std::max<int64_t>(Vt, 1), -1);
  for (int64_t i = 0; i < H; ++i) slot_of[order[i]] = (int32_t)i;
  std::vector<int32_t> xadj((size_t)std::max<int64_t>(E, 1)), tadj((size_t)std::max<int64_t>(E, 1));
  struct Ctx {
    const int64_t *beg;
    const int32_t *adj, *deg, *slot_of, *cid;
    int32_t *xadj, *tadj;
    int64_t Vt;
    int nt;
  } cx{tbeg.data(), adj, tdeg.data(), slot_of.data(), cid.data(), xadj.data(), tadj.data(), Vt,
       nthreads};
  parallel_for_threads(cx.nt, [](int t, void *p) {
    Ctx &c = *(Ctx *)p;
    for (int64_t v = t; v < c.Vt; v += c.nt) {
      const int64_t s = c.beg[v], e = c.beg[v + 1];
      for (int64_t j = s; j < e; ++j) c.xadj[j] = c.cid[c.adj[j]];
      int32_t *a = c.tadj + s, *ae = c.tadj + e;
      std::copy(c.xadj + s, c.xadj + e, a);
      std::sort(a, ae, [&](int32_t x, int32_t y) {
        return c.deg[x] != c.deg[y] ? c.deg[x] > c.deg[y] : x < y;
      });
      for (int32_t *x = a; x < ae; ++x)
        if (c.slot_of[*x] >= 0) *x = (int32_t)(0x80000000u | (uint32_t)c.slot_of[*x]);
    }
  }, &cx);
  auto *g = new ibfs_graph_s;
  g->device = device;
  g->V = V;
  g->E = E;
  g->Vt = Vt;
  g->H = H;
  g->max_deg = maxd;
  g->h_hub_deg.resize((size_t)H);
  g->h_hub_id.resize((size_t)H);
  for (int64_t i = 0; i < H; ++i) {
    g->h_hub_id[i] = order[i];
    g->h_hub_deg[i] = tdeg[order[i]];
  }
  {  // entries per hub slot, then the prefix over slots
    std::vector<int64_t> per((size_t)H + 1, 0);
    for (int64_t j = 0; j < E; ++j)
      if (tadj[j] < 0) ++per[(size_t)(tadj[j] & 0x7fffffff) + 1];
    for (int64_t i = 1; i <= H; ++i) per[i] += per[i - 1];
    g->h_hub_entries = std::move(per);
  }
  const size_t eb = sizeof(int32_t) * (size_t)std::max<int64_t>(E, 1);
  const size_t vtb = sizeof(int32_t) * (size_t)std::max<int64_t>(Vt, 1);
  hipError_t err = hipMalloc((void **)&g->d_beg, sizeof(int64_t) * (size_t)(Vt + 1));
  if (err == hipSuccess) err = hipMalloc((void **)&g->d_adj, eb);
  if (err == hipSuccess) err = hipMalloc((void **)&g->d_tadj, eb);
  if (err == hipSuccess) err = hipMalloc((void **)&g->d_deg, vtb);
  if (err == hipSuccess) err = hipMalloc((void **)&g->d_hubs, sizeof(int32_t) * (size_t)std::max<int64_t>(H, 1));
  if (err == hipSuccess) err = hipMalloc((void **)&g->d_cid, sizeof(int32_t) * (size_t)V);
  if (err == hipSuccess) err = hipMalloc((void **)&g->d_vid, vtb);
  if (err == hipSuccess) err = hipMemcpy(g->d_beg, tbeg.data(), sizeof(int64_t) * (size_t)(Vt + 1), hipMemcpyHostToDevice);
  if (err == hipSuccess && E > 0) err = hipMemcpy(g->d_adj, xadj.data(), sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice);
  if (err == hipSuccess && E > 0) err = hipMemcpy(g->d_tadj, tadj.data(), sizeof(int32_t) * (size_t)E, hipMemcpyHostToDevice);
  if (err == hipSuccess && Vt > 0) err = hipMemcpy(g->d_deg, tdeg.data(), sizeof(int32_t) * (size_t)Vt, hipMemcpyHostToDevice);
  if (err == hipSuccess && H > 0) err = hipMemcpy(g->d_hubs, order.data(), sizeof(int32_t) * (size_t)H, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(g->d_cid, cid.data(), sizeof(int32_t) * (size_t)V, hipMemcpyHostToDevice);
  if (err == hipSuccess && Vt > 0) {
    std::vector<int32_t> vid((size_t)Vt);
    for (int64_t v = 0; v < V; ++v)
      if (cid[v] >= 0) vid[(size_t)cid[v]] = (int32_t)v;
    err = hipMemcpy(g->d_vid, vid.data(), sizeof(int32_t) * (size_t)Vt, hipMemcpyHostToDevice);
  }
  if (err == hipSuccess) g->h_cid = std::move(cid);
  if (err != hipSuccess) {
    dfree(g->d_beg);
    dfree(g->d_adj);
    dfree(g->d_tadj);
    dfree(g->d_deg);
    dfree(g->d_hubs);
    dfree(g->d_cid);
    dfree(g->d_vid);
    delete g;
    return fail(err == hipErrorOutOfMemory ? IBFS_ERR_NOMEM : IBFS_ERR_HIP,
                "graph upload failed: %s", hipGetErrorString(err));
  }
  *out = g;
  return IBFS_OK;
}

int ibfs_graph_from_host(ibfs_hostcsr_t h, int device, ibfs_graph_t *out) {
  if (!h) return fail(IBFS_ERR_ARG, "null host csr");
  return ibfs_graph_create(h->csr.beg.data(), h->csr.V, h->csr.adj.data(), h->csr.E, device, out);
}

int ibfs_graph_load_files(const char *beg_path, const char *csr_path, int device,
                          ibfs_graph_t *out) {
  ibfs_hostcsr_t h = nullptr;
  int rc = ibfs_hostcsr_load(beg_path, csr_path, &h);
  if (rc) return rc;
  rc = ibfs_graph_from_host(h, device, out);
  ibfs_hostcsr_free(h);
  return rc;
}

int ibfs_graph_info(ibfs_graph_t g, int64_t *V, int64_t *E) {
  if (!g) return fail(IBFS_ERR_ARG, "null graph");
  if (V) *V = g->V;
  if (E) *E = g->E;
  return IBFS_OK;
}

int ibfs_graph_layout(ibfs_graph_t g, int64_t *traversal_vertices, int64_t *hubs) {
  if (!g) return fail(IBFS_ERR_ARG, "null graph");
  if (traversal_vertices) *traversal_vertices = g->Vt;
  if (hubs) *hubs = g->H;
  return IBFS_OK;
}

int ibfs_graph_hub_entries(ibfs_graph_t g, int64_t top, int64_t *entries) {
  if (!g || !entries) return fail(IBFS_ERR_ARG, "ibfs_graph_hub_entries: null argument");
  if (top <= 0 || top > g->H) top = g->H;
  *entries = g->h_hub_entries[(size_t)top];
  return IBFS_OK;
}

int ibfs_graph_destroy(ibfs_graph_t g) {
  if (!g) return IBFS_OK;
  (void)hipSetDevice(g->device);
  dfree(g->d_beg);
  dfree(g->d_adj);
  dfree(g->d_tadj);
  dfree(g->d_deg);
  dfree(g->d_hubs);
  dfree(g->d_cid);
  dfree(g->d_vid);
  delete g;
  return IBFS_OK;
}

int ibfs_groupby(ibfs_graph_t g, const int32_t *src, int64_t n, int32_t *out) {
  if (!g || (n > 0 && (!src || !out)) || n < 0) return fail(IBFS_ERR_ARG, "ibfs_groupby: bad arguments");
  if (n == 0) return IBFS_OK;
  for (int64_t i = 0; i < n; ++i)
    if (src[i] < 0 || src[i] >= g->V)
      return fail(IBFS_ERR_ARG, "source %lld = %d out of range", (long long)i, src[i]);
  int rc = check_device(g->device);
  if (rc) return rc;
  int32_t *d = nullptr;
  void *ws = nullptr;
  const size_t wsb = groupby_workspace_bytes(n);
  HIP_TRY(hipMalloc((void **)&d, sizeof(int32_t) * (size_t)n * 3));
  hipError_t err = hipMalloc(&ws, wsb);
  if (err == hipSuccess) err = hipMemcpy(d, src, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice);
  if (err == hipSuccess)
    err = launch_groupby(g->d_beg, g->d_adj, g->d_deg, g->d_cid, d, n, d + n, d + 2 * n, ws, wsb,
                         nullptr);
  if (err == hipSuccess)
    err = hipMemcpy(out, d + 2 * n, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost);
  (void)hipFree(ws);
  (void)hipFree(d);
  if (err != hipSuccess) return fail(IBFS_ERR_HIP, "groupby failed: %s", hipGetErrorString(err));
  return IBFS_OK;
}

int ibfs_engine_create(ibfs_graph_t g, const ibfs_config *cfg_in, ibfs_engine_t *out) {
  if (!g || !out) return fail(IBFS_ERR_ARG, "ibfs_engine_create: null argument");
  ibfs_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else ibfs_config_default(&cfg);
  if (cfg.lane_bits != 64 && cfg.lane_bits != 128 && cfg.lane_bits != 256 && cfg.lane_bits != 512)
    return fail(IBFS_ERR_ARG, "lane_bits must be 64, 128, 256 or 512 (got %d)", cfg.lane_bits);
  if (cfg.sml_shed < 1 || cfg.lrg_shed < cfg.sml_shed)
    return fail(IBFS_ERR_ARG, "need 1 <= sml_shed <= lrg_shed (got %d, %d)", cfg.sml_shed, cfg.lrg_shed);
  if (int rc = check_direction_config(cfg)) return rc;
  if (cfg.td_chunk < 256 || (cfg.td_chunk & (cfg.td_chunk - 1)))
    return fail(IBFS_ERR_ARG, "td_chunk must be a power of two >= 256 (got %d)", cfg.td_chunk);
  if (cfg.emit_mode < 0 || cfg.emit_mode > 3)
    return fail(IBFS_ERR_ARG, "emit_mode must be 0, 1, 2 or 3 (got %d)", cfg.emit_mode);
  if (cfg.ring_levels < 4 || cfg.ring_levels > 255)
    return fail(IBFS_ERR_ARG, "ring_levels must be in [4, 255] (got %d)", cfg.ring_levels);
  if (!valid_inspect_vpt(cfg.inspect_vpt))
    return fail(IBFS_ERR_ARG, "inspect_vpt must be 0, 1, 2, 4, 8 or 16 (got %d)", cfg.inspect_vpt);
  if (!valid_wg_size(cfg.wg_size))
    return fail(IBFS_ERR_ARG, "wg_size must be 256, 512 or 1024 (got %d)", cfg.wg_size);
  if (cfg.device != g->device)
    return fail(IBFS_ERR_ARG, "config device %d differs from the graph's device %d", cfg.device, g->device);
  int rc = check_device(g->device);
  if (rc) return rc;
  auto *e = new ibfs_engine_s;
  e->g = g;
  e->cfg = cfg;
  e->K = cfg.lane_bits;
  e->W = e->Wmax = cfg.lane_bits / 64;
  e->chunk_shift = __builtin_ctz((unsigned)cfg.td_chunk);
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device) !=
            hipSuccess || cus <= 0)
      cus = 256;
    e->hy_grid = cfg.dd_grid > 0 ? cfg.dd_grid : cus * 32;
  }
  e->hc = hub_cut(g, cfg.hybrid_hubs);
  const int64_t Vt = std::max<int64_t>(g->Vt, 1);
  // ring: at most 8 GiB of snapshots, at least 4 levels
  const size_t slot_bytes = (size_t)Vt * (size_t)e->W * sizeof(u64);
  int R = cfg.ring_levels;
  const size_t budget = (size_t)8 << 30;
  while (R > 4 && (size_t)R * slot_bytes > budget) --R;
  e->R = R;
  hipError_t err = hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking);
  if (err == hipSuccess) err = hipEventCreateWithFlags(&e->ev_wait, hipEventDisableTiming);
  auto A = [&](void **p, size_t bytes) {
    if (err == hipSuccess) err = hipMalloc(p, bytes ? bytes : 1);
  };
  auto H = [&](void **p, size_t bytes) {
    if (err == hipSuccess) err = hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault);
  };
  A((void **)&e->ring, (size_t)R * slot_bytes);
  A((void **)&e->d_ctl, sizeof(LevelCtl) * (size_t)R);
  H((void **)&e->h_ctl, sizeof(LevelCtl));
  if (err == hipSuccess) err = hipHostMalloc((void **)&e->post, sizeof(CtlPost), hipHostMallocCoherent);
  if (err == hipSuccess) std::memset(e->post, 0, sizeof(CtlPost));
  for (int i = 0; i < 6; ++i)
    if (i != Q_TD_L) A((void **)&e->q.q[i], sizeof(int32_t) * (size_t)Vt);
  const size_t nchunks = (size_t)Vt + (size_t)(g->E >> e->chunk_shift) + 2;
  A((void **)&e->q.chunks, sizeof(int2) * nchunks);
  A((void **)&e->d_masks, sizeof(u64) * masks_words(e));
  H((void **)&e->h_masks, sizeof(u64) * masks_words(e));
  A((void **)&e->d_src, sizeof(int32_t) * (size_t)e->K);
  H((void **)&e->h_src, sizeof(int32_t) * (size_t)e->K);
  A((void **)&e->d_lanes, sizeof(LaneOut) * (size_t)e->K);
  H((void **)&e->h_lanes, sizeof(LaneOut) * (size_t)e->K);
  A((void **)&e->hs, sizeof(u64) * (size_t)e->W * (size_t)std::max<int64_t>(g->H, 1));
  A((void **)&e->flags, sizeof(u64) * (size_t)((Vt + 7) / 8));
  A((void **)&e->d_count, 4 * sizeof(u64));
  H((void **)&e->h_count, 4 * sizeof(u64));
  if (err != hipSuccess) {
    engine_release(e);
    delete e;
    return fail(err == hipErrorOutOfMemory ? IBFS_ERR_NOMEM : IBFS_ERR_HIP,
                "engine allocation failed: %s", hipGetErrorString(err));
  }
  *out = e;
  return IBFS_OK;
}

int ibfs_engine_destroy(ibfs_engine_t e) {
  if (!e) return IBFS_OK;
  engine_release(e);
  delete e;
  return IBFS_OK;
}

int ibfs_traversal(ibfs_engine_t e, const int32_t *src, int32_t k, int32_t *depth_out,
                   int32_t out_place, ibfs_stats *stats) {
  if (!e) return fail(IBFS_ERR_ARG, "null engine");
  if (k < 0 || k > e->K) return fail(IBFS_ERR_ARG, "k=%d outside [0, %d]", k, e->K);
  if (k > 0 && (!src || !depth_out)) return fail(IBFS_ERR_ARG, "null sources or output");
  if (out_place != IBFS_OUT_HOST && out_place != IBFS_OUT_DEVICE)
    return fail(IBFS_ERR_ARG, "bad out_place %d", out_place);
  const int64_t V = e->g->V;
  for (int i = 0; i < k; ++i)
    if (src[i] < 0 || src[i] >= V)
      return fail(IBFS_ERR_ARG, "source %d = %d out of range [0,%lld)", i, src[i], (long long)V);
  if (stats) std::memset(stats, 0, sizeof *stats);
  if (k == 0) return IBFS_OK;
  int rc = check_device(e->g->device);
  if (rc) return rc;
  if (e->unsettled) {
    rc = poll_stream(e->st, e->ev_wait);
    if (rc) return rc;
    e->unsettled = false;
  }
  int32_t *dout = depth_out;
  if (out_place == IBFS_OUT_HOST) {
    rc = ensure_out(e, (size_t)k * (size_t)V);
    if (rc) return rc;
    dout = e->d_out;
  }
  rc = traverse(e, src, k, dout, stats);
  if (rc == IBFS_OK && out_place == IBFS_OUT_HOST) {
    const D2HPiece pc{depth_out, dout, sizeof(int32_t) * (size_t)k * (size_t)V};
    rc = copy_to_host(&pc, 1, e->st);
  }
  if (rc) e->unsettled = true;
  return rc;
}

int ibfs_run_groups(ibfs_engine_t e, const int32_t *src, int64_t n, int32_t *depth_out,
                    int32_t out_place, ibfs_stats *stats) {
  if (!e || n < 0 || (n > 0 && (!src || !depth_out))) return fail(IBFS_ERR_ARG, "ibfs_run_groups: bad arguments");
  if (stats) std::memset(stats, 0, sizeof *stats);
  const int64_t V = e->g->V;
  for (int64_t g0 = 0; g0 < n; g0 += e->K) {
    const int k = (int)std::min<int64_t>(e->K, n - g0);
    ibfs_stats s{};
    int rc = ibfs_traversal(e, src + g0, k, depth_out + (size_t)g0 * (size_t)V, out_place,
                            stats ? &s : nullptr);
    if (rc) return rc;
    if (stats) {
      stats->levels = std::max(stats->levels, s.levels);
      stats->emit_windows += s.emit_windows;
      stats->traversed_edges += s.traversed_edges;
      stats->total_ms += s.total_ms;
      stats->init_ms += s.init_ms;
      stats->expand_td_ms += s.expand_td_ms;
      stats->expand_bu_ms += s.expand_bu_ms;
      stats->inspect_ms += s.inspect_ms;
      stats->emit_ms += s.emit_ms;
      stats->n_td_launches += s.n_td_launches;
      stats->n_bu_launches += s.n_bu_launches;
      stats->n_inspect_launches += s.n_inspect_launches;
      stats->n_emit_launches += s.n_emit_launches;
      stats->expand_hy_ms += s.expand_hy_ms;
      stats->n_hy_launches += s.n_hy_launches;
    }
  }
  return IBFS_OK;
}

int ibfs_engine_level_log(ibfs_engine_t e, ibfs_level_info *out, int32_t cap, int32_t *n) {
  if (!e || !n || cap < 0 || (cap > 0 && !out)) return fail(IBFS_ERR_ARG, "bad arguments");
  *n = (int32_t)e->log.size();
  for (int32_t i = 0; i < cap && i < *n; ++i) out[i] = e->log[i];
  return IBFS_OK;
}

}  // extern "C"
