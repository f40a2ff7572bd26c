// Allocation-free traversal: one group of up to 128 sources with every
// scratch buffer (status snapshots, queues, control blocks, hub cache,
// changed flags, lane table, masks, plan, edge counter) in one file-scope
// device array, and the device-driven level loop (captured-level kernels
// launched directly on the null stream, direction decided on the device,
// the host reading only the plan's head between bursts).  For callers that
// cannot allocate: code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "kernels.h"

using namespace ibfs;

namespace {

// Limits of the scratch array (each GPU holds one copy until the process
// ends): up to 128 lanes (two status words), 2^18 vertices with an edge
// (R-MAT scale 18), 2^24 adjacency entries, the default hub cache (131,072
// hubs), 8 ring snapshots, top-down chunks of at least 2,048 edges.
constexpr int NA_W = 2;
constexpr int NA_K = 64 * NA_W;
constexpr int64_t NA_VT = int64_t(1) << 18;
constexpr int64_t NA_E = int64_t(1) << 24;
constexpr int64_t NA_H = 131072;
constexpr int NA_R = 8;
constexpr int NA_MIN_CHUNK_SHIFT = 11;

constexpr size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

// byte offsets of each buffer inside the array
constexpr size_t OFF_RING = 0;
constexpr size_t SZ_RING = size_t(NA_R) * NA_VT * NA_W * sizeof(u64);
constexpr size_t OFF_CTL = OFF_RING + up256(SZ_RING);
constexpr size_t SZ_CTL = size_t(NA_R) * sizeof(LevelCtl);
constexpr size_t OFF_Q = OFF_CTL + up256(SZ_CTL);
constexpr size_t SZ_Q1 = up256(size_t(NA_VT) * sizeof(int32_t));
constexpr size_t OFF_CHUNKS = OFF_Q + 5 * SZ_Q1;
constexpr size_t SZ_CHUNKS = (size_t(NA_VT) + (size_t(NA_E) >> NA_MIN_CHUNK_SHIFT) + 2) * sizeof(int2);
constexpr size_t OFF_HS = OFF_CHUNKS + up256(SZ_CHUNKS);
constexpr size_t SZ_HS = size_t(NA_H) * NA_W * sizeof(u64);
constexpr size_t OFF_FLAGS = OFF_HS + up256(SZ_HS);
constexpr size_t SZ_FLAGS = size_t(NA_VT);
constexpr size_t OFF_MASKS = OFF_FLAGS + up256(SZ_FLAGS);
constexpr size_t SZ_MASKS = size_t(2 + NA_R + 1) * NA_W * sizeof(u64);
constexpr size_t OFF_LANES = OFF_MASKS + up256(SZ_MASKS);
constexpr size_t SZ_LANES = size_t(NA_K) * sizeof(LaneOut);
constexpr size_t OFF_PLAN = OFF_LANES + up256(SZ_LANES);
constexpr size_t OFF_COUNT = OFF_PLAN + up256(sizeof(Plan));
constexpr size_t NA_BYTES = OFF_COUNT + 256;
static_assert(NA_BYTES == IBFS_NOALLOC_SCRATCH_BYTES, "scratch size in ibfs.h");

// The scratch array.  In host code its name is a host-side stand-in, not a
// device address: host code uses only the per-device address from
// hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

struct SourceList {  // sources of the group, passed by value to the set-up kernel
  int32_t src[NA_K];
  int32_t k;
};

struct PlanInit {
  int policy, sw_level, alpha, beta, gamma, rho;
  u64 V;
};

// Set-up of one traversal (slot 0, flags and control blocks already zero):
// lane table, masks, source bits of S_0, plan.  One workgroup.
template <int W>
__global__ void k_noalloc_setup(SourceList sl, const int32_t *__restrict__ cid, int32_t *dout,
                                int64_t V, u64 *S0, uint8_t *flags, LaneOut *lanes, u64 *masks,
                                int R, Plan *plan, PlanInit pi) {
  const int t = threadIdx.x;
  for (int i = t; i < (int)(sizeof(Plan) / 4); i += blockDim.x)
    reinterpret_cast<uint32_t *>(plan)[i] = 0;
  for (int i = t; i < (2 + R + 1) * W; i += blockDim.x) {
    u64 m = ~0ull;  // level masks: every lane admitted at level 0
    if (i < 2 * W) {
      const int w = i % W, lo = w * 64;
      const int n = sl.k - lo;
      m = n >= 64 ? ~0ull : (n <= 0 ? 0ull : (1ull << n) - 1);
    }
    masks[i] = m;
  }
  if (t < sl.k) {
    const int32_t s = sl.src[t];
    lanes[t].out = dout + (int64_t)t * V;
    lanes[t].admit = 0;
    lanes[t].src = s;
    const int32_t v = cid[s];
    if (v >= 0) {
      atomicOr(&S0[(int64_t)v * W + (t >> 6)], 1ull << (t & 63));
      flags[v] = 1;
    }
  }
  __syncthreads();
  if (t == 0) {
    plan->policy = pi.policy;
    plan->sw_level = pi.sw_level;
    plan->alpha = pi.alpha;
    plan->beta = pi.beta;
    plan->gamma = pi.gamma;
    plan->rho = pi.rho;
    plan->split_cov = 0;
    plan->nf_prev = ~0ull;
    plan->V = pi.V;
    plan->dir = -1;
  }
}

std::mutex g_mu;  // one traversal at a time per process (the array is shared)
std::vector<unsigned char *> g_ptr;  // per device: the array's device address

// Device address of the scratch array on the current device `dev`, looked
// up once per device.
int scratch_ptr(int dev, unsigned char **out) {
  if ((size_t)dev >= g_ptr.size()) g_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_ptr[(size_t)dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(IBFS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_ptr[(size_t)dev] = static_cast<unsigned char *>(p);
  }
  *out = g_ptr[(size_t)dev];
  return IBFS_OK;
}

#define NA_TRY(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(IBFS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                  __FILE__, __LINE__);                                                 \
  } while (0)

// Waits for the null stream within the library's wait limit.
int wait_null_stream() {
  WaitClock clk;
  for (;;) {
    const hipError_t q = hipStreamQuery(nullptr);
    if (q == hipSuccess) return IBFS_OK;
    if (q != hipErrorNotReady) return fail(IBFS_ERR_HIP, "stream error: %s", hipGetErrorString(q));
    if (clk.expired())
      return fail(IBFS_ERR_STATE, "traversal did not finish within the wait limit (%.1f s)",
                  wait_timeout_s());
  }
}

struct PlanHead {  // the first members of Plan
  int level, dir, done, last;
};

int run(ibfs_graph_s *g, const ibfs_config &cfg, const int32_t *src, int k, int32_t *dout,
        int64_t *edges) {
  int dev = 0;
  NA_TRY(hipGetDevice(&dev));
  unsigned char *base = nullptr;
  if (int rc = scratch_ptr(dev, &base)) return rc;
  const int W = k <= 64 ? 1 : 2;
  const int R = NA_R;
  const int64_t V = g->V, Vt = g->Vt;
  const int cs = __builtin_ctz((unsigned)cfg.td_chunk);
  u64 *ring = reinterpret_cast<u64 *>(base + OFF_RING);
  LevelCtl *ctl = reinterpret_cast<LevelCtl *>(base + OFF_CTL);
  Queues q{};
  for (int i = 0, j = 0; i < 6; ++i)
    if (i != Q_TD_L) q.q[i] = reinterpret_cast<int32_t *>(base + OFF_Q + (size_t)(j++) * SZ_Q1);
  q.chunks = reinterpret_cast<int2 *>(base + OFF_CHUNKS);
  u64 *hs = reinterpret_cast<u64 *>(base + OFF_HS);
  uint8_t *flags = base + OFF_FLAGS;
  u64 *masks = reinterpret_cast<u64 *>(base + OFF_MASKS);
  LaneOut *lanes = reinterpret_cast<LaneOut *>(base + OFF_LANES);
  Plan *plan = reinterpret_cast<Plan *>(base + OFF_PLAN);
  u64 *count = reinterpret_cast<u64 *>(base + OFF_COUNT);
  const HubCut hc = hub_cut(g, cfg.hybrid_hubs);
  const HubSplit hsp{hc.deg, hc.tie, hc.cut};
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  const int grid = cfg.dd_grid > 0 ? cfg.dd_grid : cus * 32;
  hipStream_t st = nullptr;  // the null stream: nothing to create

  // the previous call's contents stay in the array: clear what starts at zero
  NA_TRY(launch_fill_u64(ring, Vt * W, 0ull, st));  // slot 0
  NA_TRY(launch_fill_u64(reinterpret_cast<u64 *>(flags), (Vt + 7) / 8, 0ull, st));
  NA_TRY(launch_fill_u64(reinterpret_cast<u64 *>(ctl), (int64_t)R * CTL_WORDS, 0ull, st));
  NA_TRY(launch_fill_u64(count, 1, 0ull, st));
  SourceList sl{};
  for (int i = 0; i < k; ++i) sl.src[i] = src[i];
  sl.k = k;
  const PlanInit pi{cfg.policy, cfg.sw_level, cfg.alpha, cfg.beta, cfg.hybrid_gamma, hc.rho_q10,
                    (u64)V};
  if (W == 1)
    hipLaunchKernelGGL(k_noalloc_setup<1>, dim3(1), dim3(NA_K), 0, st, sl, g->d_cid, dout, V,
                       ring, flags, lanes, masks, R, plan, pi);
  else
    hipLaunchKernelGGL(k_noalloc_setup<2>, dim3(1), dim3(NA_K), 0, st, sl, g->d_cid, dout, V,
                       ring, flags, lanes, masks, R, plan, pi);
  NA_TRY(hipGetLastError());

  InspectArgs a{};
  a.deg = g->d_deg;
  a.V = Vt;
  a.q = q;
  a.sml = cfg.sml_shed;
  a.lrg = cfg.lrg_shed;
  a.chunk_shift = cs;
  a.hubs = g->d_hubs;
  a.H = g->H;
  a.hs = hs;
  a.plan = plan;
  a.ring = ring;
  a.R = R;
  a.ctl_ring = ctl;
  a.lanes_all = masks;
  a.flags = flags;
  a.vpt = cfg.inspect_vpt;
  a.nw = cfg.wg_size / 64;
  a.hsplit = hsp;
  NA_TRY(launch_inspect(W, a, st));
  NA_TRY(launch_plan_decide(plan, ctl, R, hs, W, 0, st));

  const EmitSpace sp{g->d_cid, V, Vt, g->d_deg};
  const int vec4 = (V % 4 == 0) && ((uintptr_t)dout % 16 == 0);
  int level = 1, lo = 0;
  for (;;) {
    int B = std::min(8, R - 1 - (level - lo));
    if (B < 1) {  // the next levels would overwrite S_lo: emit [lo, level-1]
      NA_TRY(launch_emit(W, cfg.emit_mode, vec4, ring, R, sp, lo, level - 1, lanes, masks + 2 * W,
                         masks + W, k, nullptr, st));
      lo = level - 1;
      B = std::min(8, R - 1 - (level - lo));
    }
    for (int i = 0; i < B; ++i) {
      NA_TRY(launch_expand_dd(W, g->d_beg, g->d_adj, g->d_tadj, ring, flags, R, Vt, ctl, plan, hs,
                              q, cs, hsp, TailLog{}, grid, st));
      NA_TRY(launch_inspect(W, a, st));
      NA_TRY(launch_plan_decide(plan, ctl, R, hs, W, 0, st));
    }
    if (int rc = wait_null_stream()) return rc;
    PlanHead ph{};
    NA_TRY(hipMemcpy(&ph, plan, sizeof ph, hipMemcpyDeviceToHost));
    if (ph.done) {
      if (ph.last >= lo)
        NA_TRY(launch_emit(W, cfg.emit_mode, vec4, ring, R, sp, lo, ph.last, lanes, masks + 2 * W,
                           masks + W, k, edges ? count : nullptr, st));
      else if (edges)  // no final window: count over S_last
        NA_TRY(launch_count_edges(W, ring + (int64_t)(ph.last % R) * Vt * W, g->d_deg, Vt, masks,
                                  count, st));
      if (int rc = wait_null_stream()) return rc;
      if (edges) {
        u64 c = 0;
        NA_TRY(hipMemcpy(&c, count, sizeof c, hipMemcpyDeviceToHost));
        // sources without a status word reach themselves, degree 0: nothing to add
        *edges = (int64_t)c;
      }
      return IBFS_OK;
    }
    if (ph.level > Vt + 2) return fail(IBFS_ERR_STATE, "traversal exceeded V+2 levels");
    level = ph.level;
  }
}

}  // namespace

extern "C" {

int ibfs_traversal_noalloc(ibfs_graph_t g, const ibfs_config *cfg_in, const int32_t *src,
                           int32_t k, int32_t *depth_out, int64_t *traversed_edges) {
  if (!g) return fail(IBFS_ERR_ARG, "ibfs_traversal_noalloc: null graph");
  if (k < 1 || k > NA_K) return fail(IBFS_ERR_ARG, "k=%d outside [1, %d]", k, NA_K);
  if (!src || !depth_out) return fail(IBFS_ERR_ARG, "null sources or output");
  if (g->Vt > NA_VT || g->E > NA_E || g->H > NA_H)
    return fail(IBFS_ERR_ARG,
                "graph beyond the scratch limits: %lld vertices with an edge (max %lld), %lld "
                "adjacency entries (max %lld), %lld hubs (max %lld)",
                (long long)g->Vt, (long long)NA_VT, (long long)g->E, (long long)NA_E,
                (long long)g->H, (long long)NA_H);
  ibfs_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else ibfs_config_default(&cfg);
  if (cfg.sml_shed < 1 || cfg.lrg_shed < cfg.sml_shed)
    return fail(IBFS_ERR_ARG, "need 1 <= sml_shed <= lrg_shed");
  if (int rc = check_direction_config(cfg)) return rc;
  if (cfg.td_chunk < (1 << NA_MIN_CHUNK_SHIFT) || (cfg.td_chunk & (cfg.td_chunk - 1)))
    return fail(IBFS_ERR_ARG, "td_chunk must be a power of two >= %d here (got %d)",
                1 << NA_MIN_CHUNK_SHIFT, cfg.td_chunk);
  if (cfg.emit_mode < 0 || cfg.emit_mode > 3) return fail(IBFS_ERR_ARG, "bad emit_mode");
  if (!valid_inspect_vpt(cfg.inspect_vpt)) return fail(IBFS_ERR_ARG, "bad inspect_vpt");
  if (!valid_wg_size(cfg.wg_size)) return fail(IBFS_ERR_ARG, "wg_size must be 256, 512 or 1024");
  const int64_t V = g->V;
  for (int i = 0; i < k; ++i)
    if (src[i] < 0 || src[i] >= V)
      return fail(IBFS_ERR_ARG, "source %d = %d out of range [0,%lld)", i, src[i], (long long)V);
  std::lock_guard<std::mutex> lk(g_mu);
  int cur = 0;
  NA_TRY(hipGetDevice(&cur));
  NA_TRY(hipSetDevice(g->device));
  const int rc = run(g, cfg, src, k, depth_out, traversed_edges);
  (void)hipSetDevice(cur);
  return rc;
}

}  // extern "C"
