// Traversal kernels (gfx950, wave64): lane init, inspector, top-down and
// bottom-up expanders (host-driven and device-driven), bottom-up work model.
#include <hip/hip_runtime.h>

#include "../../include/ibfs.h"
#include "device_util.h"

namespace ibfs {
namespace {

using namespace dev;

__global__ void k_fill_u64(u64 *p, int64_t n, u64 value) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    p[i] = value;
}

// Three ranges filled with zeros in one launch (a group's first snapshot,
// changed flags and control blocks), over their concatenated index space.
__global__ void k_zero3_u64(u64 *p0, int64_t n0, u64 *p1, int64_t n1, u64 *p2, int64_t n2) {
  const int64_t n01 = n0 + n1, n = n01 + n2;
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS) {
    if (i < n0) p0[i] = 0;
    else if (i < n01) p1[i - n0] = 0;
    else p2[i - n01] = 0;
  }
}

// One wave: thread t < CTL_WORDS copies word t of the block (vector stores),
// makes it visible system-wide, and after the barrier thread 0 releases the tag.
__global__ void k_post_ctl(const LevelCtl *src, CtlPost *dst, u64 tag) {
  const int t = threadIdx.x;
  if (t < (int)CTL_WORDS) {
    reinterpret_cast<u64 *>(&dst->c)[t] = atomic_read(reinterpret_cast<const u64 *>(src) + t);
    __threadfence_system();
  }
  __syncthreads();
  if (t == 0) __hip_atomic_store(&dst->tag, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t value) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    p[i] = value;
}

// Lane init: lane lane0+i gets bit of src[i] (a traversal id).  OR, so
// duplicate sources keep their bit.  src[i] < 0 is an isolated source: it
// has no status word and its depth array is -1 except 0 at the source (the
// reference stores and skips zero-degree sources, expander.cu:63,70-71).
template <int W>
__global__ void k_init_sources(const int32_t *src, int k, int lane0, u64 *S0, uint8_t *flags) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= k || src[i] < 0) return;
  int lane = lane0 + i;
  atomicOr(&S0[(int64_t)src[i] * W + (lane >> 6)], 1ull << (lane & 63));
  flags[src[i]] = 1;
}

// Admission: bit lane[i] set in S[src[i]] (lanes of any position).
template <int W>
__global__ void k_admit(const int32_t *src, const int32_t *lane, int n, u64 *S, uint8_t *flags) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n || src[i] < 0) return;
  const int k = lane[i];
  atomicOr(&S[(int64_t)src[i] * W + (k >> 6)], 1ull << (k & 63));
  flags[src[i]] = 1;
}

// Admission into a device-driven traversal between the expansion and the
// inspection of level L = plan->level: when the posted list is tagged for L,
// lane lane[i] (clean in every snapshot so far) gets its bit at src[i] in
// S_L, the vertex is flagged, and the lanes join lam_{L-1} so that the
// inspection of L counts a vertex that misses one of them as a bottom-up
// candidate of L+1 (lam_{L-1} is not read again after that).  A lane whose
// source is isolated gains nothing in any status word and stays out of
// lam.  One workgroup; nothing to do when the plan is done or the list is
// for another level.
template <int W>
__global__ void k_admit_dd(const AdmitList *al, const Plan *plan, u64 *ring, int R, int64_t V,
                           LevelCtl *ctl_ring, uint8_t *flags) {
  const int L = plan->level;
  if (plan->done || plan->tail_start || L < 1 || al->level != L) return;
  const int n = al->n;
  u64 *S = ring + (int64_t)(L % R) * V * W;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int32_t v = al->src[i];
    const int k = al->lane[i];
    if (v >= 0) {
      atomicOr(&S[(int64_t)v * W + (k >> 6)], 1ull << (k & 63));
      flags[v] = 1;
    }
  }
  if (threadIdx.x < W) {
    u64 m = 0;
    for (int i = 0; i < n; ++i)
      if (al->src[i] >= 0 && (al->lane[i] >> 6) == (int)threadIdx.x)
        m |= 1ull << (al->lane[i] & 63);
    if (m) atomicOr(&ctl_ring[(L - 1) % R].lam[threadIdx.x], m);
  }
}

// Direction of level L+1 from the counters of level L (same rule as the
// host-driven driver): fixed switch level, or alpha/beta; or the entry into
// and the progress of the sparse tail (Plan::tail_div).
__device__ void plan_decide(Plan *p, const LevelCtl *c, int L) {
  if (p->tail_start) {
    // sparse tail level L: the records it appended are the next frontier;
    // none ends the traversal, an overflowing log ends it for a dense rerun
    const u64 cnt = atomic_read(&p->tail_count);
    if (cnt > p->tail_cap || cnt == p->tail_hi) {
      p->done = 1;
      p->last = L - 1;
      return;
    }
    p->tail_lo = p->tail_hi;
    p->tail_hi = cnt;
    p->dirlog[(L + 1) & 255] = 4;
    p->level = L + 1;
    return;
  }
  const u64 nch = atomic_read(&c->n_changed);
  const u64 nf = atomic_read(&c->n_front);
  const u64 mf = atomic_read(&c->m_front);
  const u64 mc = atomic_read(&c->m_cand);
  const u64 nc = atomic_read(&c->n_cand);
  const u64 mfn = atomic_read(&c->m_front_nh);
  if (nch == 0 && L >= p->hold) {
    p->done = 1;
    // level 0 always ends a window: lanes whose source is isolated change
    // nothing in the status words but still have depth 0 at the source
    p->last = L > 0 ? L - 1 : 0;
    return;
  }
  const int Ln = L + 1;
  // the candidates bound what the tail can still append (every record holds
  // a bit one of them gains): on a long-diameter graph (a grid, a path) most
  // vertices still miss some lane when the frontier is small, and their
  // records would overflow the log and cost a dense rerun
  if (p->tail_div > 0 && L >= 2 && L >= p->hold && nf <= p->nf_prev &&
      nf * (u64)p->tail_div <= p->V && mf * (u64)p->tail_div <= p->E &&
      nc * (u64)p->tail_div <= p->V) {
    p->tail_start = Ln;
    p->tail_lo = p->tail_hi = 0;
    p->nf_prev = nf;
    p->dir = 4;
    p->td_flag = 0;
    p->dirlog[Ln & 255] = 4;
    p->level = Ln;
    return;
  }
  const int dir_prev = (L == 0) ? -1 : p->dir;
  const int dir = choose_direction(p->policy, p->sw_level, p->alpha, p->beta, p->gamma, p->rho,
                                   p->V, Ln, dir_prev, p->nf_prev, nf, mf, mc, mfn);
  p->nf_prev = nf;
  p->dir = dir;
  p->td_flag = dir == 0 && flag_top_down(mf, p->V);
  p->dirlog[Ln & 255] = dir;
  p->level = Ln;
}

// ---------------------------------------------------------------------------
// Inspector of level L: one pass over the traversal vertices.  Copies S_L
// (cur) into S_{L+1} (next, the slot the expanders of L+1 update), refreshes
// the hub cache from S_L, diffs S_L against S_{L-1} and accumulates the lanes
// that advanced (lam of L).  After a bottom-up or a sparse top-down level only
// the vertices the expander (or the source admission) flagged are diffed
// (prev is read only for them); after other levels and at level 0 every
// vertex is.  Flags
// are cleared as they are read.  It compacts, per degree class, both the top-down
// frontier of L+1 (changed, degree > 0) and the bottom-up candidates
// (degree > 0 and missing a lane of lam_prev).  Replaces
// td/sw_frontier_count + insp_scan + td/sw_frontier_gather + extract_depth
// of inspector.cu:37-500 and scan.cuh:455-904 with a single kernel.
//
// A workgroup of NW waves takes tiles of NW*64*VPT consecutive vertices;
// lane l of wave w handles vertex (w*VPT + i)*64 + l in round i.  A wave
// issues all loads of its rounds before any store, classifies in registers,
// and the waves' class totals meet in LDS, so each tile reserves its queue
// ranges with three 64-bit atomics (two classes each): same-address device
// atomics serialize, so their number per level is what this layout keeps
// small.  Entries go to ballot prefix positions, ascending vertex order
// within a tile.  Device-driven mode: the level and its slots come from the
// plan; k_plan_decide then picks the next direction.
__device__ __forceinline__ unsigned lanes_below(u64 m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

template <int W, int VPT, int NW>
__global__ __launch_bounds__(NW * 64) void k_inspect(InspectArgs a) {
  constexpr int IBS = NW * 64;
  __shared__ unsigned wsum[2][NW][6];   // per-wave class totals, double-buffered by tile parity
  __shared__ unsigned wbase[2][NW][6];  // per-wave queue bases
  __shared__ u64 red[NW][6 + W];

  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const u64 *prev = a.prev, *cur = a.cur;
  u64 *next = a.next;
  const u64 *lam_prev = a.lam_prev;
  LevelCtl *ctl = a.ctl, *ctl_zero = a.ctl_zero;
  int full = a.full_diff;
  if (a.plan) {
    if (a.plan->done || a.plan->tail_start) return;  // sparse tail levels skip the pass
    const int L = a.plan->level;
    full = (L == 0 || !(a.plan->dir == 1 || (a.plan->dir == 0 && a.plan->td_flag)));
    const int64_t slot = a.V * W;
    prev = L > 0 ? a.ring + (int64_t)((L - 1) % a.R) * slot : nullptr;
    cur = a.ring + (int64_t)(L % a.R) * slot;
    next = a.ring + (int64_t)((L + 1) % a.R) * slot;
    ctl = a.ctl_ring + L % a.R;
    ctl_zero = a.ctl_ring + (L + 1) % a.R;
    lam_prev = L > 0 ? a.ctl_ring[(L - 1) % a.R].lam : a.lanes_all;
  }
  if (blockIdx.x == 0 && ctl_zero && t < (int)(sizeof(LevelCtl) / 8))
    reinterpret_cast<u64 *>(ctl_zero)[t] = 0;

  u64 lp[W], keep[W], nand[W], lamacc[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    lp[w] = lam_prev[w];
    keep[w] = a.prev_keep ? a.prev_keep[w] : ~0ull;
    nand[w] = a.next_and ? a.next_and[w] : ~0ull;
    lamacc[w] = 0;
  }
  for (int64_t i = (int64_t)blockIdx.x * IBS + t; i < a.H; i += (int64_t)gridDim.x * IBS) {
    u64 x[W];
    load_word<W>(cur + (int64_t)a.hubs[i] * W, x);
    store_word<W>(a.hs + i * W, x);
  }
  u64 n_changed = 0, n_front = 0, m_front = 0, n_cand = 0, m_cand = 0, m_front_nh = 0;
  const HubSplit hsp = a.hsplit;
  const int64_t V = a.V;
  const int sml = a.sml, lrg = a.lrg, cs = a.chunk_shift;
  uint8_t *flags = a.flags;
  const u64 *__restrict__ curp = cur;
  const u64 *__restrict__ prevp = prev;
  const int32_t *__restrict__ degp = a.deg;
  // queue lengths as three 64-bit pairs (q[2j] low half, q[2j+1] high half)
  unsigned long long *qpair = reinterpret_cast<unsigned long long *>(ctl->q);

  constexpr int TILE = IBS * VPT;
  const int64_t ntiles = (V + TILE - 1) / TILE;
  int par = 0;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x, par ^= 1) {
    const int64_t c0 = tile * TILE + (int64_t)wv * 64 * VPT;
    // every load of the wave's rounds first, so that none waits behind a store
    uint8_t fl[VPT];
    int dgv[VPT];
    u64 cw[VPT][W], pw[VPT][W];
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int64_t v = c0 + i * 64 + lane;
      fl[i] = 0;
      dgv[i] = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) cw[i][w] = pw[i][w] = 0;
      if (v < V) {
        fl[i] = flags[v];
        dgv[i] = degp[v];
        load_word<W>(curp + v * W, cw[i]);
      }
    }
    if (prevp) {
#pragma unroll
      for (int i = 0; i < VPT; ++i)
        if ((full || fl[i]) && c0 + i * 64 + lane < V)
          load_word<W>(prevp + (c0 + i * 64 + lane) * W, pw[i]);
    }
    uint8_t code[VPT];  // bits 0-1 top-down class, bits 2-3 bottom-up class
    unsigned tot[6] = {0, 0, 0, 0, 0, 0};  // wave totals (uniform)
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int64_t v = c0 + i * 64 + lane;
      int tc = 0, bc = 0, nck = 0;
      bool ch = false;
      if (v < V) {
        bool cover = true;
        u64 n[W];
        if (fl[i]) flags[v] = 0;
        if (full || fl[i]) {
#pragma unroll
          for (int w = 0; w < W; ++w) {
            const u64 d = cw[i][w] & ~(pw[i][w] & keep[w]);
            lamacc[w] |= d;
            ch |= (d != 0);
          }
        }
#pragma unroll
        for (int w = 0; w < W; ++w) {
          n[w] = cw[i][w] & nand[w];
          cover &= ((cw[i][w] & lp[w]) == lp[w]);
        }
        store_word<W>(next + v * W, n);
        const int dg = dgv[i];
        if (ch) {
          ++n_changed;
          if (dg > 0) {
            ++n_front;
            m_front += (u64)dg;
            if (!is_hub(hsp, v, dg)) m_front_nh += (u64)dg;
            tc = degree_class(dg, sml, lrg);
            if (tc == 3) nck = (dg + (1 << cs) - 1) >> cs;
          }
        }
        if (dg > 0 && !cover) {
          ++n_cand;
          m_cand += (u64)dg;
          bc = degree_class(dg, sml, lrg);
        }
      }
      code[i] = (uint8_t)(tc | (bc << 2));
      tot[Q_TD_S] += (unsigned)__popcll(__ballot(tc == 1));
      tot[Q_TD_M] += (unsigned)__popcll(__ballot(tc == 2));
      if (__ballot(tc == 3)) tot[Q_TD_L] += (unsigned)wave_sum((u64)nck);
      tot[Q_BU_S] += (unsigned)__popcll(__ballot(bc == 1));
      tot[Q_BU_M] += (unsigned)__popcll(__ballot(bc == 2));
      tot[Q_BU_L] += (unsigned)__popcll(__ballot(bc == 3));
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) wsum[par][wv][k] = tot[k];
    }
    __syncthreads();
    // reservation for the tile: thread j < 3 adds classes 2j, 2j+1 in one
    // 64-bit atomic and lays out the waves' bases in wave order
    if (t < 3) {
      unsigned lo = 0, hi = 0;
      for (int w = 0; w < NW; ++w) {
        lo += wsum[par][w][2 * t];
        hi += wsum[par][w][2 * t + 1];
      }
      const unsigned long long add = (unsigned long long)lo | ((unsigned long long)hi << 32);
      const unsigned long long got = add ? atomicAdd(&qpair[t], add) : 0ull;
      unsigned blo = (unsigned)got, bhi = (unsigned)(got >> 32);
      for (int w = 0; w < NW; ++w) {
        wbase[par][w][2 * t] = blo;
        wbase[par][w][2 * t + 1] = bhi;
        blo += wsum[par][w][2 * t];
        bhi += wsum[par][w][2 * t + 1];
      }
    }
    __syncthreads();
    unsigned base[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) base[k] = wbase[par][wv][k];
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int tc = code[i] & 3, bc = code[i] >> 2;
      const int32_t v = (int32_t)(c0 + i * 64 + lane);
      u64 m = __ballot(tc == 1);
      if (tc == 1) a.q.q[Q_TD_S][base[Q_TD_S] + lanes_below(m)] = v;
      base[Q_TD_S] += (unsigned)__popcll(m);
      m = __ballot(tc == 2);
      if (tc == 2) a.q.q[Q_TD_M][base[Q_TD_M] + lanes_below(m)] = v;
      base[Q_TD_M] += (unsigned)__popcll(m);
      if (__ballot(tc == 3)) {  // hub frontier vertices: (vertex, chunk) pairs
        const int nck = (tc == 3) ? (dgv[i] + (1 << cs) - 1) >> cs : 0;
        const unsigned inc = wave_incl_scan((unsigned)nck, lane);
        if (tc == 3) {
          const unsigned p0 = base[Q_TD_L] + inc - (unsigned)nck;
          for (int c = 0; c < nck; ++c) a.q.chunks[p0 + c] = make_int2(v, c);
        }
        base[Q_TD_L] += (unsigned)__shfl((int)inc, 63, 64);
      }
      m = __ballot(bc == 1);
      if (bc == 1) a.q.q[Q_BU_S][base[Q_BU_S] + lanes_below(m)] = v;
      base[Q_BU_S] += (unsigned)__popcll(m);
      m = __ballot(bc == 2);
      if (bc == 2) a.q.q[Q_BU_M][base[Q_BU_M] + lanes_below(m)] = v;
      base[Q_BU_M] += (unsigned)__popcll(m);
      m = __ballot(bc == 3);
      if (bc == 3) a.q.q[Q_BU_L][base[Q_BU_L] + lanes_below(m)] = v;
      base[Q_BU_L] += (unsigned)__popcll(m);
    }
  }

  // workgroup reduction of the level counters, one atomic each
  u64 vals[6 + W] = {n_changed, n_front, m_front, n_cand, m_cand, m_front_nh};
#pragma unroll
  for (int w = 0; w < W; ++w) vals[6 + w] = lamacc[w];
#pragma unroll
  for (int i = 0; i < 6; ++i) vals[i] = wave_sum(vals[i]);
#pragma unroll
  for (int w = 0; w < W; ++w) vals[6 + w] = wave_or(vals[6 + w]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 6 + W; ++i) red[wv][i] = vals[i];
  }
  __syncthreads();
  if (t < 6) {  // n_changed .. m_front_nh are consecutive
    u64 s = 0;
    for (int w = 0; w < NW; ++w) s += red[w][t];
    if (s) atomicAdd(&(&ctl->n_changed)[t], s);
  } else if (t < 6 + W) {
    u64 s = 0;
    for (int w = 0; w < NW; ++w) s |= red[w][t];
    if (s) atomicOr(&ctl->lam[t - 6], s);
  }
}

// Device-driven decision after the inspection of plan->level.  A separate
// launch instead of a last-workgroup election inside the inspector: on this
// multi-XCD part each workgroup's agent-scope fence would write back and
// invalidate its L2 in the middle of the pass.
// Thread 0 applies the direction rule.  A bottom-up choice with split_cov > 0
// then samples 64 of the top `hubs` hubs in the hub cache hs (S_L words, W per
// hub; wave w takes word w): an active lane set in at least split_cov % of
// the sample is dense and stays bottom-up, the other active lanes push
// top-down.  The level splits (dir 3) when both kinds are active.  A lane
// that holds few hubs has reached little of the graph: its candidates would
// scan their whole lists without covering it, while its own frontier is small.
__global__ __launch_bounds__(512) void k_plan_decide(Plan *plan, const LevelCtl *ctl_ring, int R,
                                                     const u64 *hs, int W, int hubs) {
  __shared__ int dir_s;
  __shared__ u64 dense[MAX_W];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int L = plan->level, done = plan->done, cov = plan->split_cov;
  __syncthreads();  // every wave has read the plan before thread 0 changes it
  if (done) return;
  if (t == 0) {
    plan_decide(plan, ctl_ring + L % R, L);
    dir_s = plan->done ? -1 : plan->dir;
  }
  __syncthreads();
  if (dir_s != 1 || cov <= 0 || hubs <= 0 || R < 3) return;  // (hubs > 0: 64 * W threads)
  const int ns = hubs < 64 ? hubs : 64, step = hubs / ns;
  if (wv < W) {
    const u64 x = lane < ns ? hs[(int64_t)lane * step * W + wv] : 0;
    u64 d = 0;
    for (int b = 0; b < 64; ++b) {
      const u64 m = __ballot((x >> b) & 1);
      if (__popcll(m) * 100 >= cov * ns) d |= 1ull << b;
    }
    if (lane == 0) dense[wv] = d;
  }
  __syncthreads();
  if (t == 0) {
    const LevelCtl *c = ctl_ring + L % R;
    u64 dn = 0, sp = 0;
    for (int w = 0; w < W; ++w) {
      const u64 lw = atomic_read(&c->lam[w]);
      dn |= lw & dense[w];
      sp |= lw & ~dense[w];
    }
    if (dn && sp) {
      for (int w = 0; w < W; ++w) plan->split[w] = dense[w];
      plan->dir = 3;
      plan->dirlog[(L + 1) & 255] = 3;
    }
  }
}

// ---------------------------------------------------------------------------
// Top-down: a frontier vertex u pushes S_{L-1}[u] & lam_{L-1} into S_L of its
// neighbours with 64-bit atomic ORs, skipping neighbours that already hold
// the bits (expander.cu:100-252).
// A dense top-down level flags nothing: its changes are scattered, and a byte
// store per changed vertex costs a partial-line write once the line is
// evicted; the inspector diffs every vertex after it instead.  A sparse one
// (flag_top_down) flags the vertices it changes.
// NEW (split levels): only the bits u gained at L-1 (S_{L-1} & ~S_{L-2}, pp
// = S_{L-2} or null) are pushed, and u is skipped when none is in lam.
template <int W, bool NEW = false>
__device__ __forceinline__ bool td_bits(const u64 *__restrict__ prev, const u64 *__restrict__ pp,
                                        const u64 (&lam)[W], int32_t u, u64 (&p)[W]) {
  load_word<W>(prev + (int64_t)u * W, p);
  u64 any = 0;
  if (NEW && pp) {
    u64 o[W];
    load_word<W>(pp + (int64_t)u * W, o);
#pragma unroll
    for (int w = 0; w < W; ++w) p[w] &= ~o[w];
  }
#pragma unroll
  for (int w = 0; w < W; ++w) {
    p[w] &= lam[w];
    any |= p[w];
  }
  return !NEW || any != 0;
}

template <int W, bool NEW = false>
__device__ __forceinline__ void td_push(const int64_t *__restrict__ beg,
                                        const int32_t *__restrict__ adj,
                                        const u64 *__restrict__ prev, u64 *cur,
                                        const u64 (&lam)[W], int32_t u, int64_t j0, int64_t j1,
                                        int stride, const u64 *pp = nullptr,
                                        uint8_t *flags = nullptr) {
  u64 p[W];
  if (!td_bits<W, NEW>(prev, pp, lam, u, p)) return;
  for (int64_t j = j0; j < j1; j += stride) {
    const int64_t v = adj[j];
    u64 c[W];
    load_word<W>(cur + v * W, c);
    bool ch = false;
#pragma unroll
    for (int w = 0; w < W; ++w)
      if (p[w] & ~c[w]) {
        atomicOr(&cur[v * W + w], p[w]);
        ch = true;
      }
    if (flags && ch) flags[v] = 1;
  }
}

// One thread per small frontier vertex: four neighbours per step, their
// adjacency entries and words loaded before any atomic, so four dependent
// load chains are in flight instead of one.
template <int W, bool NEW = false>
__device__ __forceinline__ void td_push_small(const int64_t *__restrict__ beg,
                                              const int32_t *__restrict__ adj,
                                              const u64 *__restrict__ prev, u64 *cur,
                                              const u64 (&lam)[W], int32_t u,
                                              const u64 *pp = nullptr, uint8_t *flags = nullptr) {
  u64 p[W];
  if (!td_bits<W, NEW>(prev, pp, lam, u, p)) return;
  int64_t j = beg[u];
  const int64_t e = beg[u + 1];
  for (; j + 4 <= e; j += 4) {
    int64_t v[4];
    u64 c[4][W];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = adj[j + r];
#pragma unroll
    for (int r = 0; r < 4; ++r) load_word<W>(cur + v[r] * W, c[r]);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      bool ch = false;
#pragma unroll
      for (int w = 0; w < W; ++w)
        if (p[w] & ~c[r][w]) {
          atomicOr(&cur[v[r] * W + w], p[w]);
          ch = true;
        }
      if (flags && ch) flags[v[r]] = 1;
    }
  }
  for (; j < e; ++j) {
    const int64_t v = adj[j];
    u64 c[W];
    load_word<W>(cur + v * W, c);
    bool ch = false;
#pragma unroll
    for (int w = 0; w < W; ++w)
      if (p[w] & ~c[w]) {
        atomicOr(&cur[v * W + w], p[w]);
        ch = true;
      }
    if (flags && ch) flags[v] = 1;
  }
}

// chunk i of a large top-down vertex: edges [lo, hi)
__device__ __forceinline__ void td_chunk_range(const int64_t *__restrict__ beg, int2 e, int cs,
                                               int64_t &lo, int64_t &hi) {
  lo = beg[e.x] + ((int64_t)e.y << cs);
  const int64_t end = beg[e.x + 1];
  hi = (lo + (1ll << cs)) < end ? (lo + (1ll << cs)) : end;
}

// neighbours gathered per step by one thread's bottom-up candidate
#ifndef IBFS_BU_STEP
#define IBFS_BU_STEP 4
#endif

// Result of a bottom-up candidate: x (= x0 | new bits) into cur.  A plain
// store and a flag for the inspector's flagged diff; a hybrid level ORs the
// changed words in atomically instead (pushes may target the same vertex).
template <int W, bool HY>
__device__ __forceinline__ void bu_commit(u64 *cur, uint8_t *flags, int64_t v, const u64 (&x)[W],
                                          const u64 (&x0)[W]) {
  if (!differs<W>(x, x0)) return;
  if (HY) {
#pragma unroll
    for (int w = 0; w < W; ++w)
      if (x[w] != x0[w]) atomicOr(&cur[v * W + w], x[w]);
  } else {
    store_word<W>(cur + v * W, x);
    flags[v] = 1;
  }
}

// Bottom-up: a candidate v ORs S_{L-1} of its neighbours (restricted to
// lam_{L-1}, the only lanes that can still be new) in traversal-adjacency
// order (descending degree, hubs first) and stops as soon as its word
// covers lam_{L-1} (bitwise early termination, expander.cu:258-359).
// One candidate per thread, IBFS_BU_STEP neighbours per step.
// MODE 0: bottom-up level.  MODE 1 (hybrid level): only the prefix of hub
// entries with slot < cut is read (from the hub cache).  MODE 2 (split
// level): the whole list, for the dense lanes in lam.  In modes 1 and 2 new
// bits go in with atomic ORs (frontier vertices push into the same words
// concurrently) and nothing is flagged.
__device__ __forceinline__ bool pulled(int32_t a, int cut) {
  return a < 0 && (a & 0x7fffffff) < cut;
}

template <int W, int MODE = 0>
__device__ __forceinline__ void bu_thread(const int64_t *__restrict__ beg,
                                          const int32_t *__restrict__ tadj, const Gather &gs,
                                          u64 *cur, uint8_t *flags,
                                          const u64 (&lam)[W], int64_t v, int cut = 0) {
  u64 x[W], x0[W];
  load_word<W>(cur + v * W, x);
#pragma unroll
  for (int w = 0; w < W; ++w) x0[w] = x[w];
  if (covers<W>(x, lam)) return;
  int64_t j = beg[v];
  const int64_t e = beg[v + 1];
  while (j < e) {
    int32_t an[IBFS_BU_STEP];
    an[0] = tadj[j];
#pragma unroll
    for (int r = 1; r < IBFS_BU_STEP; ++r) an[r] = (j + r < e) ? tadj[j + r] : an[0];
    u64 sn[IBFS_BU_STEP][W];
    bool nonhub = false;
#pragma unroll
    for (int r = 0; r < IBFS_BU_STEP; ++r) {
      if (MODE == 1 && !pulled(an[r], cut)) {
        nonhub = true;
#pragma unroll
        for (int w = 0; w < W; ++w) sn[r][w] = 0;
      } else {
        load_nbr<W>(gs, an[r], sn[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < IBFS_BU_STEP; ++r)
#pragma unroll
      for (int w = 0; w < W; ++w) x[w] |= sn[r][w] & lam[w];
    j += IBFS_BU_STEP;
    if (covers<W>(x, lam) || nonhub) break;  // pulled prefix ended
  }
  bu_commit<W, MODE != 0>(cur, flags, v, x, x0);
}

// One candidate per wave: 64 neighbours per step, wave OR-reduction.
template <int W, int MODE = 0>
__device__ __forceinline__ void bu_wave(const int64_t *__restrict__ beg,
                                        const int32_t *__restrict__ tadj, const Gather &gs,
                                        u64 *cur, uint8_t *flags,
                                        const u64 (&lam)[W], int64_t v, int lane, int cut = 0) {
  u64 x[W], x0[W];
  load_word<W>(cur + v * W, x);
#pragma unroll
  for (int w = 0; w < W; ++w) x0[w] = x[w];
  if (covers<W>(x, lam)) return;
  const int64_t s = beg[v], e = beg[v + 1];
  for (int64_t r = s; r < e; r += 64) {
    const int64_t j = r + lane;
    u64 y[W];
#pragma unroll
    for (int w = 0; w < W; ++w) y[w] = 0;
    bool nonhub = false;
    if (j < e) {
      const int32_t a = tadj[j];
      if (MODE == 1 && !pulled(a, cut)) nonhub = true;
      else load_nbr<W>(gs, a, y);
    }
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] |= wave_or(y[w] & lam[w]);
    if (covers<W>(x, lam)) break;
    if (MODE == 1 && __ballot(nonhub)) break;  // pulled prefix ended
  }
  if (lane == 0) bu_commit<W, MODE != 0>(cur, flags, v, x, x0);
}

// One candidate per workgroup: 256 neighbours per step, the four wave
// partials combined through LDS.  Must be called by the whole workgroup.
// nh: four ints of LDS (pulled prefix ended, per wave; used by MODE 1).
template <int W, int MODE = 0>
__device__ __forceinline__ void bu_block(const int64_t *__restrict__ beg,
                                         const int32_t *__restrict__ tadj, const Gather &gs,
                                         u64 *cur, uint8_t *flags,
                                         const u64 (&lam)[W], int64_t v, u64 (*part)[W],
                                         int *nh, int cut = 0) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  u64 x[W], x0[W];
  if (MODE != 0) {
    // One read of the candidate's word for the whole workgroup: in a hybrid
    // or split level pushes from other workgroups can change it between the waves'
    // own reads, and every wave must take the same covers/stop branches
    // around the barriers below.
    if (t == 0) {
      load_word<W>(cur + v * W, x);
#pragma unroll
      for (int w = 0; w < W; ++w) part[0][w] = x[w];
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] = part[0][w];
    __syncthreads();
  } else {
    load_word<W>(cur + v * W, x);
  }
#pragma unroll
  for (int w = 0; w < W; ++w) x0[w] = x[w];
  if (covers<W>(x, lam)) return;
  const int64_t s = beg[v], e = beg[v + 1];
  for (int64_t r = s; r < e; r += BS) {
    const int64_t j = r + t;
    u64 y[W];
#pragma unroll
    for (int w = 0; w < W; ++w) y[w] = 0;
    bool nonhub = false;
    if (j < e) {
      const int32_t a = tadj[j];
      if (MODE == 1 && !pulled(a, cut)) nonhub = true;
      else load_nbr<W>(gs, a, y);
    }
#pragma unroll
    for (int w = 0; w < W; ++w) y[w] = wave_or(y[w] & lam[w]);
    const bool wnh = MODE == 1 && __ballot(nonhub) != 0;
    if (lane == 0) {
#pragma unroll
      for (int w = 0; w < W; ++w) part[wv][w] = y[w];
      if (MODE == 1) nh[wv] = wnh;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] |= part[0][w] | part[1][w] | part[2][w] | part[3][w];
    const bool stop = MODE == 1 && (nh[0] | nh[1] | nh[2] | nh[3]);
    __syncthreads();
    if (covers<W>(x, lam) || stop) break;
  }
  if (t == 0) bu_commit<W, MODE != 0>(cur, flags, v, x, x0);
}

// Host-driven top-down: workgroups [0, nbs) take small vertices one per
// thread, [nbs, nbs+nbm) mid vertices one per wave, the rest one chunk of
// 2^chunk_shift edges of a large vertex per workgroup.
template <int W>
__global__ __launch_bounds__(BS) void k_expand_td(const int64_t *__restrict__ beg,
                                                  const int32_t *__restrict__ adj,
                                                  const u64 *__restrict__ prev, u64 *cur,
                                                  const u64 *__restrict__ lam_prev, Queues q,
                                                  int64_t ns, int64_t nm, int64_t nl, int64_t nbs,
                                                  int64_t nbm, int cs, uint8_t *flags) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t b = blockIdx.x;
  u64 lam[W];
#pragma unroll
  for (int w = 0; w < W; ++w) lam[w] = lam_prev[w];
  if (b < nbs) {
    const int64_t i = b * BS + t;
    if (i >= ns) return;
    td_push_small<W>(beg, adj, prev, cur, lam, q.q[Q_TD_S][i], nullptr, flags);
  } else if (b < nbs + nbm) {
    const int64_t i = (b - nbs) * 4 + wv;
    if (i >= nm) return;
    const int32_t u = q.q[Q_TD_M][i];
    td_push<W>(beg, adj, prev, cur, lam, u, beg[u] + lane, beg[u + 1], 64, nullptr, flags);
  } else {
    const int64_t i = b - nbs - nbm;
    if (i >= nl) return;
    const int2 e = q.chunks[i];
    int64_t lo, hi;
    td_chunk_range(beg, e, cs, lo, hi);
    td_push<W>(beg, adj, prev, cur, lam, e.x, lo + t, hi, BS, nullptr, flags);
  }
}

// Host-driven bottom-up, same workgroup ranges as k_expand_td.
template <int W>
__global__ __launch_bounds__(BS) void k_expand_bu(const int64_t *__restrict__ beg,
                                                  const int32_t *__restrict__ tadj, Gather gs,
                                                  u64 *__restrict__ cur, uint8_t *flags,
                                                  const u64 *__restrict__ lam_prev, Queues q,
                                                  int64_t ns, int64_t nm, int64_t nl, int64_t nbs,
                                                  int64_t nbm) {
  __shared__ u64 part[4][W];
  __shared__ int nh[4];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t b = blockIdx.x;
  u64 lam[W];
#pragma unroll
  for (int w = 0; w < W; ++w) lam[w] = lam_prev[w];
  if (b < nbs) {
    const int64_t i = b * BS + t;
    if (i < ns) bu_thread<W>(beg, tadj, gs, cur, flags, lam, q.q[Q_BU_S][i]);
  } else if (b < nbs + nbm) {
    const int64_t i = (b - nbs) * 4 + wv;
    if (i < nm) bu_wave<W>(beg, tadj, gs, cur, flags, lam, q.q[Q_BU_M][i], lane);
  } else {
    const int64_t i = b - nbs - nbm;
    if (i < nl) bu_block<W>(beg, tadj, gs, cur, flags, lam, q.q[Q_BU_L][i], part, nh);
  }
}

// Hybrid level from the queues of ctl (the control block of the level
// before), grid-strided: the bottom-up candidates pull through their prefix
// of pulled hubs (the top hsp.cut), then the other top-down frontier
// vertices push.  A pulled-hub frontier vertex u reaches every candidate
// neighbour v through that prefix of v's traversal-adjacency list; a
// neighbour that is no candidate already holds every lane u can bring.
template <int W>
__device__ __forceinline__ void expand_hybrid(const int64_t *__restrict__ beg,
                                              const int32_t *__restrict__ adj,
                                              const int32_t *__restrict__ tadj, const Gather &gs,
                                              u64 *cur, const u64 (&lam)[W],
                                              const LevelCtl *ctl, const Queues &q, int cs,
                                              HubSplit hsp, u64 (*part)[W], int *nh) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t G = gridDim.x;
  const int64_t bs = ctl->q[Q_BU_S], bm = ctl->q[Q_BU_M], bl = ctl->q[Q_BU_L];
  for (int64_t i = (int64_t)blockIdx.x * BS + t; i < bs; i += G * BS)
    bu_thread<W, 1>(beg, tadj, gs, cur, nullptr, lam, q.q[Q_BU_S][i], hsp.cut);
  for (int64_t i = (int64_t)blockIdx.x * 4 + wv; i < bm; i += G * 4)
    bu_wave<W, 1>(beg, tadj, gs, cur, nullptr, lam, q.q[Q_BU_M][i], lane, hsp.cut);
  for (int64_t i = blockIdx.x; i < bl; i += G)
    bu_block<W, 1>(beg, tadj, gs, cur, nullptr, lam, q.q[Q_BU_L][i], part, nh, hsp.cut);
  const int64_t ts = ctl->q[Q_TD_S], tm = ctl->q[Q_TD_M], tl = ctl->q[Q_TD_L];
  for (int64_t i = (int64_t)blockIdx.x * BS + t; i < ts; i += G * BS) {
    const int32_t u = q.q[Q_TD_S][i];
    if (!is_hub(hsp, u, (int)(beg[u + 1] - beg[u])))
      td_push_small<W>(beg, adj, gs.prev, cur, lam, u);
  }
  for (int64_t i = (int64_t)blockIdx.x * 4 + wv; i < tm; i += G * 4) {
    const int32_t u = q.q[Q_TD_M][i];
    const int64_t b0 = beg[u], b1 = beg[u + 1];
    if (!is_hub(hsp, u, (int)(b1 - b0))) td_push<W>(beg, adj, gs.prev, cur, lam, u, b0 + lane, b1, 64);
  }
  for (int64_t i = blockIdx.x; i < tl; i += G) {
    const int2 e = q.chunks[i];
    if (is_hub(hsp, e.x, (int)(beg[e.x + 1] - beg[e.x]))) continue;
    int64_t lo, hi;
    td_chunk_range(beg, e, cs, lo, hi);
    td_push<W>(beg, adj, gs.prev, cur, lam, e.x, lo + t, hi, BS);
  }
}

// Host-driven hybrid level (queue lengths read on the device from ctl).
template <int W>
__global__ __launch_bounds__(BS) void k_expand_hy(const int64_t *__restrict__ beg,
                                                  const int32_t *__restrict__ adj,
                                                  const int32_t *__restrict__ tadj, Gather gs,
                                                  u64 *cur, const LevelCtl *__restrict__ ctl,
                                                  Queues q, int cs, HubSplit hsp) {
  __shared__ u64 part[4][W];
  __shared__ int nh[4];
  u64 lam[W];
#pragma unroll
  for (int w = 0; w < W; ++w) lam[w] = ctl->lam[w];
  expand_hybrid<W>(beg, adj, tadj, gs, cur, lam, ctl, q, cs, hsp, part, nh);
}

// Sparse tail level L >= T = plan->tail_start: the frontier items push their
// bits along their adjacency ranges into the cumulative slot of level T.  At
// L = T the items are the top-down queues of level T-1 (small vertices one
// per thread, mid vertices and large-vertex chunks one per wave) with the
// bits S_{T-1} & ~S_{T-2}; after that they are the records of level L-1
// (vertices of degree < 64 one per thread, the others one per wave).  A
// neighbour word the push changes gains the bits that were absent from the
// atomic OR's previous value; each gain is appended as a record of level L
// (one 64-bit reservation per wave and step; records past the capacity are
// dropped and the plan ends the run for a dense rerun).  A bit is gained
// once, so each lane's arrival at a vertex is in exactly one record.
template <int W>
__device__ __forceinline__ bool tail_visit(const int32_t *__restrict__ adj, u64 *cum,
                                           const u64 (&p)[W], int64_t j, int32_t &v,
                                           u64 (&g)[W]) {
  v = adj[j];
  u64 c[W];
  load_word<W>(cum + (int64_t)v * W, c);
  bool any = false;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    g[w] = 0;
    if (p[w] & ~c[w]) {
      g[w] = p[w] & ~atomicOr(&cum[(int64_t)v * W + w], p[w]);
      any |= g[w] != 0;
    }
  }
  return any;
}

// Appends the gains of the active lanes (any: this lane gained g at v); the
// lowest active lane reserves the wave's records.  Called in divergent code:
// the ballots see the active lanes only.
template <int W>
__device__ __forceinline__ void tail_append(const Plan *plan, const TailLog &tl, u64 cap, int L,
                                            bool any, int32_t v, const u64 (&g)[W]) {
  const u64 m = __ballot(any);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_ctzll(__ballot(1));
  u64 base = 0;
  // the plan's only field written during the level
  if (lane == leader) base = atomicAdd(const_cast<u64 *>(&plan->tail_count), (u64)__popcll(m));
  base = __shfl(base, leader, 64);
  if (any) {
    const u64 pos = base + lanes_below(m);
    if (pos < cap) {
      tl.v[pos] = v;
      tl.lev[pos] = L;
      store_word<W>(tl.bits + pos * W, g);
    }
  }
}

// One item's range [lo, hi) pushed by one wave (64 neighbours per step).
template <int W>
__device__ __forceinline__ void tail_wave(const int32_t *__restrict__ adj, u64 *cum,
                                          const Plan *plan, const TailLog &tl, u64 cap, int L,
                                          const u64 (&p)[W], int64_t lo, int64_t hi) {
  const int lane = threadIdx.x & 63;
  for (int64_t j0 = lo; j0 < hi; j0 += 64) {
    u64 g[W];
    int32_t v = -1;
    bool any = false;
    if (j0 + lane < hi) any = tail_visit<W>(adj, cum, p, j0 + lane, v, g);
    tail_append<W>(plan, tl, cap, L, any, v, g);
  }
}

// One item's range [lo, hi) pushed by one thread.
template <int W>
__device__ __forceinline__ void tail_thread(const int32_t *__restrict__ adj, u64 *cum,
                                            const Plan *plan, const TailLog &tl, u64 cap, int L,
                                            const u64 (&p)[W], int64_t lo, int64_t hi) {
  for (int64_t j = lo; j < hi; ++j) {
    u64 g[W];
    int32_t v;
    const bool any = tail_visit<W>(adj, cum, p, j, v, g);
    tail_append<W>(plan, tl, cap, L, any, v, g);
  }
}

constexpr int TAIL_THREAD_DEG = 64;  // records of smaller degree: one thread each

template <int W>
__device__ void tail_level(const int64_t *__restrict__ beg, const int32_t *__restrict__ adj,
                           u64 *ring, int R, int64_t V, const LevelCtl *ctl_ring,
                           const Plan *__restrict__ plan, const Queues &q, int cs,
                           const TailLog &tl) {
  const int L = plan->level, T = plan->tail_start;
  const int64_t slot = V * W;
  u64 *cum = ring + (int64_t)(T % R) * slot;
  const int64_t nthreads = (int64_t)gridDim.x * BS, nwaves = nthreads / 64;
  const int64_t tid = (int64_t)blockIdx.x * BS + threadIdx.x, wid = tid >> 6;
  const u64 cap = plan->tail_cap;
  if (L == T) {
    const LevelCtl *c = ctl_ring + (T - 1) % R;
    const int64_t ns = c->q[Q_TD_S], nm = c->q[Q_TD_M], nl = c->q[Q_TD_L];
    const u64 *s1 = ring + (int64_t)((T - 1) % R) * slot;
    const u64 *s2 = ring + (int64_t)((T - 2) % R) * slot;
    auto bits = [&](int32_t u, u64(&p)[W]) {
      u64 o[W];
      load_word<W>(s1 + (int64_t)u * W, p);
      load_word<W>(s2 + (int64_t)u * W, o);
#pragma unroll
      for (int w = 0; w < W; ++w) p[w] &= ~o[w];
    };
    for (int64_t i = tid; i < ns; i += nthreads) {
      const int32_t u = q.q[Q_TD_S][i];
      u64 p[W];
      bits(u, p);
      tail_thread<W>(adj, cum, plan, tl, cap, L, p, beg[u], beg[u + 1]);
    }
    for (int64_t i = wid; i < nm + nl; i += nwaves) {
      int32_t u;
      int64_t lo, hi;
      if (i < nm) {
        u = q.q[Q_TD_M][i];
        lo = beg[u];
        hi = beg[u + 1];
      } else {
        const int2 e = q.chunks[i - nm];
        u = e.x;
        td_chunk_range(beg, e, cs, lo, hi);
      }
      u64 p[W];
      bits(u, p);
      tail_wave<W>(adj, cum, plan, tl, cap, L, p, lo, hi);
    }
    return;
  }
  const int64_t r0 = (int64_t)plan->tail_lo, n = (int64_t)plan->tail_hi - r0;
  for (int64_t i = tid; i < n; i += nthreads) {
    const int32_t u = tl.v[r0 + i];
    const int64_t lo = beg[u], hi = beg[u + 1];
    if (hi - lo >= TAIL_THREAD_DEG) continue;
    u64 p[W];
    load_word<W>(tl.bits + (r0 + i) * W, p);
    tail_thread<W>(adj, cum, plan, tl, cap, L, p, lo, hi);
  }
  for (int64_t i = wid; i < n; i += nwaves) {
    const int32_t u = tl.v[r0 + i];
    const int64_t lo = beg[u], hi = beg[u + 1];
    if (hi - lo < TAIL_THREAD_DEG) continue;
    u64 p[W];
    load_word<W>(tl.bits + (r0 + i) * W, p);
    tail_wave<W>(adj, cum, plan, tl, cap, L, p, lo, hi);
  }
}

// Device-driven expansion of level plan->level: a fixed grid strides over
// the class queues of the direction the inspector chose, or over the
// frontier of a sparse tail level.
template <int W>
__global__ __launch_bounds__(BS) void k_expand_dd(const int64_t *__restrict__ beg,
                                                  const int32_t *__restrict__ adj,
                                                  const int32_t *__restrict__ tadj, u64 *ring,
                                                  uint8_t *flags, int R, int64_t V,
                                                  const LevelCtl *__restrict__ ctl_ring,
                                                  const Plan *__restrict__ plan,
                                                  const u64 *__restrict__ hs, Queues q, int cs,
                                                  HubSplit hsp, TailLog tl) {
  __shared__ u64 part[4][W];
  __shared__ int nh[4];
  if (plan->done || plan->level < 1) return;
  if (plan->tail_start) {
    tail_level<W>(beg, adj, ring, R, V, ctl_ring, plan, q, cs, tl);
    return;
  }
  const int L = plan->level, dir = plan->dir;
  const LevelCtl *c = ctl_ring + (L - 1) % R;
  const int64_t slot = V * W;
  const u64 *prev = ring + (int64_t)((L - 1) % R) * slot;
  u64 *cur = ring + (int64_t)(L % R) * slot;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t G = gridDim.x;
  u64 lam[W];
#pragma unroll
  for (int w = 0; w < W; ++w) lam[w] = c->lam[w];
  if (dir == 0) {
    uint8_t *tf = plan->td_flag ? flags : nullptr;
    const int64_t ns = c->q[Q_TD_S], nm = c->q[Q_TD_M], nl = c->q[Q_TD_L];
    for (int64_t i = (int64_t)blockIdx.x * BS + t; i < ns; i += G * BS)
      td_push_small<W>(beg, adj, prev, cur, lam, q.q[Q_TD_S][i], nullptr, tf);
    for (int64_t i = (int64_t)blockIdx.x * 4 + wv; i < nm; i += G * 4) {
      const int32_t u = q.q[Q_TD_M][i];
      td_push<W>(beg, adj, prev, cur, lam, u, beg[u] + lane, beg[u + 1], 64, nullptr, tf);
    }
    for (int64_t i = blockIdx.x; i < nl; i += G) {
      const int2 e = q.chunks[i];
      int64_t lo, hi;
      td_chunk_range(beg, e, cs, lo, hi);
      td_push<W>(beg, adj, prev, cur, lam, e.x, lo + t, hi, BS, nullptr, tf);
    }
  } else if (dir == 2) {
    const Gather gs{prev, hs};
    expand_hybrid<W>(beg, adj, tadj, gs, cur, lam, c, q, cs, hsp, part, nh);
  } else if (dir == 3) {
    // split level: the dense lanes pull over whole lists, the sparse lanes
    // push the bits their frontier vertices gained at L-1
    const Gather gs{prev, hs};
    const u64 *pp = L >= 2 ? ring + (int64_t)((L - 2) % R) * slot : nullptr;
    u64 ld[W], ls[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ld[w] = lam[w] & plan->split[w];
      ls[w] = lam[w] & ~plan->split[w];
    }
    const int64_t bs = c->q[Q_BU_S], bm = c->q[Q_BU_M], bl = c->q[Q_BU_L];
    for (int64_t i = (int64_t)blockIdx.x * BS + t; i < bs; i += G * BS)
      bu_thread<W, 2>(beg, tadj, gs, cur, nullptr, ld, q.q[Q_BU_S][i]);
    for (int64_t i = (int64_t)blockIdx.x * 4 + wv; i < bm; i += G * 4)
      bu_wave<W, 2>(beg, tadj, gs, cur, nullptr, ld, q.q[Q_BU_M][i], lane);
    for (int64_t i = blockIdx.x; i < bl; i += G)
      bu_block<W, 2>(beg, tadj, gs, cur, nullptr, ld, q.q[Q_BU_L][i], part, nh);
    const int64_t ns = c->q[Q_TD_S], nm = c->q[Q_TD_M], nl = c->q[Q_TD_L];
    for (int64_t i = (int64_t)blockIdx.x * BS + t; i < ns; i += G * BS)
      td_push_small<W, true>(beg, adj, prev, cur, ls, q.q[Q_TD_S][i], pp);
    for (int64_t i = (int64_t)blockIdx.x * 4 + wv; i < nm; i += G * 4) {
      const int32_t u = q.q[Q_TD_M][i];
      td_push<W, true>(beg, adj, prev, cur, ls, u, beg[u] + lane, beg[u + 1], 64, pp);
    }
    for (int64_t i = blockIdx.x; i < nl; i += G) {
      const int2 e = q.chunks[i];
      int64_t lo, hi;
      td_chunk_range(beg, e, cs, lo, hi);
      td_push<W, true>(beg, adj, prev, cur, ls, e.x, lo + t, hi, BS, pp);
    }
  } else {
    const int64_t ns = c->q[Q_BU_S], nm = c->q[Q_BU_M], nl = c->q[Q_BU_L];
    const Gather gs{prev, hs};
    for (int64_t i = (int64_t)blockIdx.x * BS + t; i < ns; i += G * BS)
      bu_thread<W>(beg, tadj, gs, cur, flags, lam, q.q[Q_BU_S][i]);
    for (int64_t i = (int64_t)blockIdx.x * 4 + wv; i < nm; i += G * 4)
      bu_wave<W>(beg, tadj, gs, cur, flags, lam, q.q[Q_BU_M][i], lane);
    for (int64_t i = blockIdx.x; i < nl; i += G)
      bu_block<W>(beg, tadj, gs, cur, flags, lam, q.q[Q_BU_L][i], part, nh);
  }
}

// Depth patch after a sparse tail (launch_tail_patch): one thread per
// record; record i of level lev > T writes lev - admit into the row of every
// selected lane it holds, or adds lev - T to that lane's closeness sum.  Each
// (vertex, lane) arrival is in one record, so no two threads write the same
// entry.
template <int W>
__global__ __launch_bounds__(BS) void k_tail_patch(TailLog tl, int64_t n, int T,
                                                   const int32_t *__restrict__ vid,
                                                   const LaneOut *__restrict__ lanes,
                                                   const u64 *__restrict__ sel, int k_lanes,
                                                   u64 *sum_out) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * BS) {
    const int lev = tl.lev[i];
    if (lev <= T) continue;
    const int32_t v = vid[tl.v[i]];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      u64 b = tl.bits[i * W + w] & sel[w];
      while (b) {
        const int k = w * 64 + __builtin_ctzll(b);
        b &= b - 1;
        if (k >= k_lanes) break;
        if (sum_out) {
          atomicAdd(&sum_out[k], (u64)(lev - T));
        } else {
          int32_t *out = lanes[k].out;
          if (out) out[v] = lev - lanes[k].admit;
        }
      }
    }
  }
}

// Work model of a bottom-up level (instrumentation, not on the timed path):
// out[0] = vertices of degree > 0 whose S_{L-1} word misses a lane of
// lam_{L-1}; out[1] = sum over them of the sequential traversal-adjacency
// prefix after which the word covers lam_{L-1} (the whole list if never).
template <int W>
__global__ __launch_bounds__(BS) void k_bu_model(const int64_t *__restrict__ beg,
                                                 const int32_t *__restrict__ tadj,
                                                 const u64 *__restrict__ prev,
                                                 const u64 *__restrict__ hs,
                                                 const u64 *__restrict__ lam_prev, int64_t V,
                                                 u64 *out) {
  __shared__ u64 red[2][4];
  u64 lam[W];
#pragma unroll
  for (int w = 0; w < W; ++w) lam[w] = lam_prev[w];
  const Gather gs{prev, hs};
  u64 nv = 0, pref = 0;
  for (int64_t v = (int64_t)blockIdx.x * BS + threadIdx.x; v < V; v += (int64_t)gridDim.x * BS) {
    const int64_t s = beg[v], e = beg[v + 1];
    if (s == e) continue;
    u64 x[W];
    load_word<W>(prev + v * W, x);
    if (covers<W>(x, lam)) continue;
    ++nv;
    for (int64_t j = s; j < e; ++j) {
      u64 y[W];
      load_nbr<W>(gs, tadj[j], y);
#pragma unroll
      for (int w = 0; w < W; ++w) x[w] |= y[w] & lam[w];
      ++pref;
      if (covers<W>(x, lam)) break;
    }
  }
  nv = wave_sum(nv);
  pref = wave_sum(pref);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = nv;
    red[1][threadIdx.x >> 6] = pref;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const u64 s = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] +
                  red[threadIdx.x][3];
    if (s) atomicAdd(&out[threadIdx.x], s);
  }
}

#define IBFS_DISPATCH_W(W, STMT)             \
  switch (W) {                               \
    case 1: { constexpr int WW = 1; STMT; break; } \
    case 2: { constexpr int WW = 2; STMT; break; } \
    case 4: { constexpr int WW = 4; STMT; break; } \
    case 8: { constexpr int WW = 8; STMT; break; } \
    default: return hipErrorInvalidValue;    \
  }

}  // namespace

// ---------------------------------------------------------------------------
hipError_t launch_fill_u64(u64 *p, int64_t n, u64 value, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, BS * 8, 8192)), dim3(BS), 0, st, p, n, value);
  return hipGetLastError();
}

hipError_t launch_zero3_u64(u64 *p0, int64_t n0, u64 *p1, int64_t n1, u64 *p2, int64_t n2,
                            hipStream_t st) {
  if (n0 < 0 || n1 < 0 || n2 < 0) return hipErrorInvalidValue;
  const int64_t n = n0 + n1 + n2;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero3_u64, dim3(grid_for(n, BS * 8, 8192)), dim3(BS), 0, st, p0, n0, p1, n1,
                     p2, n2);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t value, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, BS * 8, 8192)), dim3(BS), 0, st, p, n, value);
  return hipGetLastError();
}

hipError_t launch_init_sources(int W, const int32_t *src, int k, int lane0, u64 *S0,
                               uint8_t *flags, hipStream_t st) {
  if (k <= 0) return hipSuccess;
  dim3 g(grid_for(k, BS, 1 << 20)), blk(BS);
  IBFS_DISPATCH_W(W, hipLaunchKernelGGL(k_init_sources<WW>, g, blk, 0, st, src, k, lane0, S0,
                                        flags));
  return hipGetLastError();
}

hipError_t launch_admit_dd(int W, const AdmitList *al, const Plan *plan, u64 *ring, int R,
                           int64_t V, LevelCtl *ctl_ring, uint8_t *flags, hipStream_t st) {
  IBFS_DISPATCH_W(W, hipLaunchKernelGGL(k_admit_dd<WW>, dim3(1), dim3(BS), 0, st, al, plan, ring,
                                        R, V, ctl_ring, flags));
  return hipGetLastError();
}

hipError_t launch_admit(int W, const int32_t *src, const int32_t *lane, int n, u64 *S,
                        uint8_t *flags, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  dim3 g(grid_for(n, BS, 1 << 20)), blk(BS);
  IBFS_DISPATCH_W(W, hipLaunchKernelGGL(k_admit<WW>, g, blk, 0, st, src, lane, n, S, flags));
  return hipGetLastError();
}

// Vertices per lane and round set: 8/W, at least 2 (8 for 64-lane words, 4
// for 128, 2 for 256 and 512: chunk registers within 8 words per lane, and
// tiles large enough that the per-tile reservation atomics stay rare), or
// less while the graph has fewer than 512 such tiles.  1,024-thread
// workgroups have half the registers per lane, so their words per lane are
// halved.  a.vpt overrides within the same register caps.
int inspect_vpt(const InspectArgs &a, int W, int NW) {
  const int cap = NW >= 16 ? 4 : 8;  // status words of chunk registers per lane
  int v;
  if (a.vpt == 1 || a.vpt == 2 || a.vpt == 4 || a.vpt == 8 || a.vpt == 16) {
    v = a.vpt > 8 ? 8 : a.vpt;
    while (v > 1 && v * W > 2 * cap) v /= 2;
    return v;
  }
  v = cap / W < 2 ? (NW >= 16 ? 1 : 2) : cap / W;
  while (v > 1 && a.V < (int64_t)64 * NW * v * 512) v /= 2;
  return v;
}

// A grid of 512 workgroups (two per CU) strides over the tiles: with one
// workgroup per tile the last round of tiles runs on a half-empty part.
#ifndef IBFS_INSPECT_GRID_CAP
#define IBFS_INSPECT_GRID_CAP 512
#endif

template <int W, int VPT, int NW>
static void inspect_launch(const InspectArgs &a, hipStream_t st) {
  const int64_t tiles = (a.V + 64 * NW * VPT - 1) / (64 * NW * VPT);
  hipLaunchKernelGGL((k_inspect<W, VPT, NW>), dim3(grid_for(tiles, 1, IBFS_INSPECT_GRID_CAP)),
                     dim3(64 * NW), 0, st, a);
}

template <int W, int NW>
static void inspect_nw(const InspectArgs &a, hipStream_t st) {
  switch (inspect_vpt(a, W, NW)) {
    case 8: inspect_launch<W, 8, NW>(a, st); break;
    case 4: inspect_launch<W, 4, NW>(a, st); break;
    case 2: inspect_launch<W, 2, NW>(a, st); break;
    default: inspect_launch<W, 1, NW>(a, st); break;
  }
}

template <int W>
static void inspect_w(const InspectArgs &a, hipStream_t st) {
  // 1,024-thread workgroups are not used for 512-lane words (register spills)
  const int nw = (a.nw == 16 && W < 8) ? 16 : (a.nw == 4 ? 4 : 8);
  if (nw == 16) inspect_nw<W, 16>(a, st);
  else if (nw == 4) inspect_nw<W, 4>(a, st);
  else inspect_nw<W, 8>(a, st);
}

hipError_t launch_inspect(int W, const InspectArgs &a, hipStream_t st) {
  if (a.V <= 0) return hipSuccess;
  IBFS_DISPATCH_W(W, inspect_w<WW>(a, st));
  return hipGetLastError();
}

hipError_t launch_expand_td(int W, const int64_t *beg, const int32_t *adj, const u64 *prev,
                            u64 *cur, const u64 *lam_prev, Queues q, int64_t ns, int64_t nm,
                            int64_t nl, int chunk_shift, uint8_t *flags, hipStream_t st) {
  const int64_t nbs = (ns + BS - 1) / BS, nbm = (nm + 3) / 4;
  const int64_t nb = nbs + nbm + nl;
  if (nb == 0) return hipSuccess;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  dim3 g((unsigned)nb), blk(BS);
  IBFS_DISPATCH_W(W, hipLaunchKernelGGL(k_expand_td<WW>, g, blk, 0, st, beg, adj, prev, cur,
                                        lam_prev, q, ns, nm, nl, nbs, nbm, chunk_shift, flags));
  return hipGetLastError();
}

hipError_t launch_expand_bu(int W, const int64_t *beg, const int32_t *tadj, const u64 *prev,
                            const u64 *hs, u64 *cur, uint8_t *flags, const u64 *lam_prev,
                            Queues q, int64_t ns, int64_t nm, int64_t nl, hipStream_t st) {
  const Gather gs{prev, hs};
  const int64_t nbs = (ns + BS - 1) / BS, nbm = (nm + 3) / 4;
  const int64_t nb = nbs + nbm + nl;
  if (nb == 0) return hipSuccess;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  dim3 g((unsigned)nb), blk(BS);
  IBFS_DISPATCH_W(W, hipLaunchKernelGGL(k_expand_bu<WW>, g, blk, 0, st, beg, tadj, gs, cur,
                                        flags, lam_prev, q, ns, nm, nl, nbs, nbm));
  return hipGetLastError();
}

hipError_t launch_expand_dd(int W, const int64_t *beg, const int32_t *adj, const int32_t *tadj,
                            u64 *ring, uint8_t *flags, int R, int64_t V, const LevelCtl *ctl_ring,
                            Plan *plan, const u64 *hs, Queues q, int chunk_shift,
                            HubSplit hsplit, TailLog tl, int grid, hipStream_t st) {
  dim3 g(grid > 0 ? grid : 1), blk(BS);
  IBFS_DISPATCH_W(W, hipLaunchKernelGGL(k_expand_dd<WW>, g, blk, 0, st, beg, adj, tadj, ring,
                                        flags, R, V, ctl_ring, plan, hs, q, chunk_shift, hsplit,
                                        tl));
  return hipGetLastError();
}

hipError_t launch_tail_patch(int W, TailLog tl, int64_t n, int T, const int32_t *vid,
                             const LaneOut *lanes, const u64 *sel, int k_lanes, u64 *sum_out,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  dim3 g(grid_for(n, BS, 4096)), blk(BS);
  IBFS_DISPATCH_W(W, hipLaunchKernelGGL(k_tail_patch<WW>, g, blk, 0, st, tl, n, T, vid, lanes,
                                        sel, k_lanes, sum_out));
  return hipGetLastError();
}

hipError_t launch_expand_hy(int W, const int64_t *beg, const int32_t *adj, const int32_t *tadj,
                            const u64 *prev, const u64 *hs, u64 *cur, const LevelCtl *ctl_prev,
                            Queues q, int chunk_shift, HubSplit hsplit, int grid, hipStream_t st) {
  dim3 g(grid > 0 ? grid : 1), blk(BS);
  const Gather gs{prev, hs};
  IBFS_DISPATCH_W(W, hipLaunchKernelGGL(k_expand_hy<WW>, g, blk, 0, st, beg, adj, tadj, gs, cur,
                                        ctl_prev, q, chunk_shift, hsplit));
  return hipGetLastError();
}

hipError_t launch_post_ctl(const LevelCtl *src, CtlPost *dst, u64 tag, hipStream_t st) {
  static_assert(CTL_WORDS <= 64, "one wave posts the control block");
  hipLaunchKernelGGL(k_post_ctl, dim3(1), dim3(64), 0, st, src, dst, tag);
  return hipGetLastError();
}

hipError_t launch_plan_decide(Plan *plan, const LevelCtl *ctl_ring, int R, const u64 *hs, int W,
                              int hubs, hipStream_t st) {
  // one wave without the split rule: a wider workgroup waits longer for a CU
  // while another context's expansion fills the device
  hipLaunchKernelGGL(k_plan_decide, dim3(1), dim3(hubs > 0 ? 64 * W : 64), 0, st, plan, ctl_ring,
                     R, hs, W, hubs);
  return hipGetLastError();
}

hipError_t launch_bu_model(int W, const int64_t *beg, const int32_t *tadj, const u64 *prev,
                           const u64 *hs, const u64 *lam_prev, int64_t V, u64 *out,
                           hipStream_t st) {
  if (V <= 0) return hipSuccess;
  dim3 g(grid_for(V, BS, 8192)), blk(BS);
  IBFS_DISPATCH_W(W, hipLaunchKernelGGL(k_bu_model<WW>, g, blk, 0, st, beg, tadj, prev, hs,
                                        lam_prev, V, out));
  return hipGetLastError();
}

}  // namespace ibfs
