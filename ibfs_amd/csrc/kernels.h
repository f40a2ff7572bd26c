// Device-side data layout and kernel launchers of libibfs (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstring>

#include "../../include/ibfs.h"

namespace ibfs {

typedef unsigned long long u64;

constexpr int MAX_W = 8;               // status words per vertex: up to 512 lanes
constexpr int MAX_LANES = 64 * MAX_W;

// Per-level control block, written by the inspector of level L with device
// atomics and read by the expanders of level L+1, the direction decision and
// the host.
struct alignas(64) LevelCtl {
  u64 lam[MAX_W];      // lanes that gained a vertex at level L (OR of new bits)
  u64 n_changed;       // vertices with new bits at L
  u64 n_front;         // ... of degree > 0 (top-down frontier of L+1)
  u64 m_front;         // sum of their degrees
  u64 n_cand;          // bottom-up candidates for L+1
  u64 m_cand;          // sum of their degrees
  u64 m_front_nh;      // sum of the degrees of frontier vertices that are not hubs
  unsigned int q[6];   // queue lengths: td sml, td mid, td lrg chunks, bu sml, bu mid, bu lrg
  unsigned int pad[14];
};
static_assert(sizeof(LevelCtl) == 192, "LevelCtl layout");
constexpr int64_t CTL_WORDS = sizeof(LevelCtl) / 8;
// the inspector reserves queue ranges two classes at a time (q[2j], q[2j+1]
// as one little-endian 64-bit counter)
static_assert(offsetof(LevelCtl, q) % 8 == 0, "LevelCtl queue pairs");

enum { Q_TD_S = 0, Q_TD_M = 1, Q_TD_L = 2, Q_BU_S = 3, Q_BU_M = 4, Q_BU_L = 5 };

// Device-resident plan of a device-driven traversal: k_plan_decide writes the
// next level's direction here after each inspection, the expander reads it.
struct Plan {
  int level;            // level to expand next (== level being inspected)
  int dir;              // direction of `level`: 0 top-down, 1 bottom-up, 2 hybrid, 3 split,
                        // 4 sparse tail
  int done;             // 1 once a level produced no change
  int last;             // deepest level with a change (valid when done)
  unsigned int blocks_done;  // unused (kept for the layout the host copies)
  int policy, sw_level, alpha, beta;
  int gamma;            // hybrid cost factor (1/64 units; 0: never hybrid)
  int rho;              // hub entries per adjacency entry of the graph (1/1024 units)
  int split_cov;        // split levels: % of the top hubs a lane must hold to stay
                        // bottom-up (0: no split levels)
  int td_flag;          // top-down level `level` flags what it changes (flag_top_down)
  int hold;             // levels below `hold` never end the traversal: lanes that
                        // start at a later level are still to be admitted
  // Sparse tail: once the shrinking top-down frontier of a level L >= 2 (at
  // or after `hold`) has at most V/tail_div vertices and E/tail_div edges,
  // every later level pushes only the bits its frontier gained one level
  // before into one cumulative slot (that of level tail_start = L+1) and
  // appends a record (vertex, level, gained bits) per gain to a TailLog; the
  // inspection and its full-slot copy are skipped.  tail_div 0: never.
  int tail_div;
  int tail_start;       // first sparse level (0: every level so far was dense)
  int tail_pad[2];
  u64 tail_cap;         // records the log holds
  u64 tail_lo, tail_hi; // records of level `level`-1: the frontier of `level`
  u64 tail_count;       // records appended (device atomics); > tail_cap: the
                        // log overflowed and the run must be repeated dense
  u64 E;                // adjacency entries (the tail's edge threshold)
  u64 nf_prev;          // frontier size of the level before `level`
  u64 V;
  int dirlog[256];      // direction per level, indexed level & 255
  u64 split[MAX_W];     // split level (dir 3): lanes that go bottom-up
};

// Sources admitted into a running device-driven traversal at level `level`
// (traversal ids, -1: isolated source without a status word; lanes clean so
// far).  Written by the host with a stream-ordered copy before the level's
// expansion is launched; entries of other levels are ignored.
struct AdmitList {
  int level;
  int n;
  int32_t src[MAX_LANES];
  int32_t lane[MAX_LANES];
};

// Records of a sparse tail: record i says that traversal vertex v[i] gained
// the lanes bits[i*W .. i*W+W) at level lev[i].
struct TailLog {
  int32_t *v;
  int32_t *lev;
  u64 *bits;
};


// Plan of a device-driven traversal before its level 0 (V: vertex count
// of the beta rule; rho: the graph's hub-entry fraction, x1024).  The sparse
// tail stays off until the caller sets tail_div, tail_cap and E.
inline void plan_init(Plan *p, const ibfs_config &c, int rho, u64 V) {
  std::memset(p, 0, sizeof(Plan));
  p->policy = c.policy;
  p->sw_level = c.sw_level;
  p->alpha = c.alpha;
  p->beta = c.beta;
  p->gamma = c.hybrid_gamma;
  p->rho = rho;
  p->split_cov = c.split_cov;
  p->nf_prev = ~0ull;
  p->V = V;
  p->dir = -1;
}

// Pulled hubs of a hybrid level: the top `cut` hubs by degree, ties by id,
// so v is one iff deg(v) > deg or (deg(v) == deg and v <= tie) (deg = degree
// of the last pulled hub, tie = its id; none: deg = INT32_MAX, tie = -1).
// Each traversal-adjacency list holds its hub neighbours as a prefix in slot
// order, so the pulled ones are a prefix too.
struct HubSplit {
  int32_t deg;
  int32_t tie;
  int32_t cut;  // hub slots a hybrid level pulls through: [0, cut)
};

// Direction of level Ln (0 top-down, 1 bottom-up, 2 hybrid; the device plan
// may turn a bottom-up choice into a split level, 3) from the
// counters of level Ln-1, shared by the host-driven drivers and the
// device-driven plan.  alpha/beta: top-down -> bottom-up when m_f * alpha >
// m_u, back when n_f * beta < V on a shrinking frontier.  A top-down choice
// becomes hybrid (candidates pull through their entries to the pulled hubs,
// the other frontier vertices push) when gamma/64 * (rho/1024 * m_u) +
// m_f(rest) < m_f: the pulled entries (rho = the pulled hubs' share of all
// adjacency entries), weighted by their cost relative to a pushed edge, plus
// the edges still pushed, against the edges a top-down level pushes.
__host__ __device__ inline int choose_direction(int policy, int sw_level, int alpha, int beta,
                                                int gamma, int rho, u64 V, int Ln, int dir_prev,
                                                u64 nf_prev, u64 nf, u64 mf, u64 mc, u64 mf_nh) {
  switch (policy) {
    case IBFS_POLICY_TD_ONLY: return 0;
    case IBFS_POLICY_BU_ONLY: return 1;
    case IBFS_POLICY_HYBRID: return 2;
    case IBFS_POLICY_FIXED: return Ln <= sw_level ? 0 : 1;
    default: break;
  }
  int dir;
  if (dir_prev != 1) dir = (mf * (u64)alpha > mc) ? 1 : 0;
  else dir = (nf * (u64)beta < V && nf < nf_prev) ? 0 : 1;
  if (dir == 0 && gamma > 0) {
    const u64 pull = ((mc * (u64)rho) >> 10) * (u64)gamma / 64;
    if (pull + mf_nh < mf) dir = 2;
  }
  return dir;
}

// A sparse top-down level (pushed edges m_f under V/8) flags every vertex
// whose word it changes, one byte store per changed vertex, so that the
// inspection after it diffs only those (a dense one would pay a partial-line
// write per flagged byte, and its inspection diffs every vertex instead).
__host__ __device__ inline bool flag_top_down(u64 mf, u64 V) { return mf * 8 < V; }

// Output slot of one lane for the depth emission.
struct LaneOut {
  int32_t *out;   // depth array of V int32 for this lane (null: lane not emitted)
  int32_t admit;  // level at which the lane's source was set (depth 0)
  int32_t src;    // the lane's source, original id (-1: none)
};

// Vertex spaces of an emission: depth arrays span V original ids; status
// words span Vt traversal ids (slot stride Vt*W); cid maps the first to the
// second (-1 = isolated: no status word, depth 0 only at the lane's source).
struct EmitSpace {
  const int32_t *cid;
  int64_t V;
  int64_t Vt;
  const int32_t *deg;  // traversal-space degrees (the edge count of a final window)
};

struct Queues {
  int32_t *q[6];  // q[Q_TD_L] unused: top-down hubs use `chunks`
  int2 *chunks;   // (vertex, chunk index) of top-down large vertices
};

// Arguments of one inspector pass.  Host-driven: prev/cur/next/lam_prev/ctl
// given explicitly and plan == nullptr.  Device-driven: plan != nullptr and
// the level, snapshot slots and control blocks are derived from plan->level.
// flags: one byte per traversal vertex, set by the bottom-up expander and
// the source admission for the vertices they change, cleared by the pass.
// full_diff (host-driven): diff every vertex (after a top-down level and at
// level 0) instead of only the flagged ones; plan mode derives it.
struct InspectArgs {
  const u64 *prev;
  const u64 *cur;
  u64 *next;
  const u64 *lam_prev;
  LevelCtl *ctl;
  LevelCtl *ctl_zero;
  const u64 *prev_keep;  // lanes whose previous bits count (null: all)
  const u64 *next_and;   // mask applied to the copy into next (null: all)
  const int32_t *deg;
  int64_t V;
  Queues q;
  int sml, lrg, chunk_shift;
  const int32_t *hubs;  // hub cache refresh: hs[i] = cur[hubs[i]], i < H
  int64_t H;
  u64 *hs;
  Plan *plan;           // device-driven mode
  u64 *ring;
  int R;
  LevelCtl *ctl_ring;
  const u64 *lanes_all; // lam_prev of level 0
  uint8_t *flags;
  int vpt;              // vertices per lane and chunk (0: by size)
  int full_diff;
  int nw;               // waves per workgroup: 4, 8 or 16 (0: 8)
  HubSplit hsplit;      // frontier degrees outside the pulled hubs are counted separately
};

// The launchers report hipGetLastError() after their launches; a HIP call
// that failed earlier on the same thread (its result already handled) leaves
// that error set, so each operation clears it before its first launch.
inline void clear_hip_error() { (void)hipGetLastError(); }

// Launchers (all asynchronous on `st`).  W = status words (u64) per vertex.
hipError_t launch_fill_u64(u64 *p, int64_t n, u64 value, hipStream_t st);
// zeros in [p0, p0+n0), [p1, p1+n1) and [p2, p2+n2) (u64 words), one launch
hipError_t launch_zero3_u64(u64 *p0, int64_t n0, u64 *p1, int64_t n1, u64 *p2, int64_t n2,
                            hipStream_t st);
// A control block posted to coherent pinned host memory: the block's words,
// then `tag` with a system-scope release, so a host that reads the tag with
// an acquire load sees the block without a copy command or an event.
struct CtlPost {
  LevelCtl c;
  u64 tag;
};
hipError_t launch_post_ctl(const LevelCtl *src, CtlPost *dst, u64 tag, hipStream_t st);
hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t value, hipStream_t st);
// S0[src[i]] |= bit lane0+i, i < k (src in traversal ids; -1 entries skipped).
hipError_t launch_init_sources(int W, const int32_t *src, int k, int lane0, u64 *S0,
                               uint8_t *flags, hipStream_t st);
// S[src[i]] |= bit lane[i], i < n (traversal ids; -1 entries skipped).
hipError_t launch_admit(int W, const int32_t *src, const int32_t *lane, int n, u64 *S,
                        uint8_t *flags, hipStream_t st);
// Device-driven admission between the expansion and the inspection of
// plan->level (k_admit_dd): no effect unless al->level == plan->level.
hipError_t launch_admit_dd(int W, const AdmitList *al, const Plan *plan, u64 *ring, int R,
                           int64_t V, LevelCtl *ctl_ring, uint8_t *flags, hipStream_t st);
hipError_t launch_inspect(int W, const InspectArgs &a, hipStream_t st);
// flags (or null): set for every vertex whose word changes (flag_top_down)
hipError_t launch_expand_td(int W, const int64_t *beg, const int32_t *adj, const u64 *prev,
                            u64 *cur, const u64 *lam_prev, Queues q, int64_t ns, int64_t nm,
                            int64_t nl, int chunk_shift, uint8_t *flags, hipStream_t st);
// tadj: traversal adjacency (neighbours by descending degree, hub entries
// encoded as 0x80000000 | hub slot); hs: hub cache mirroring prev.
hipError_t launch_expand_bu(int W, const int64_t *beg, const int32_t *tadj, const u64 *prev,
                            const u64 *hs, u64 *cur, uint8_t *flags, const u64 *lam_prev,
                            Queues q, int64_t ns, int64_t nm, int64_t nl, hipStream_t st);
// Hybrid level, fixed grid (queue lengths read from ctl_prev on the
// device): the bottom-up candidates OR the hub-cache words of their entries
// to the pulled hubs, the rest of the top-down frontier pushes; both with
// atomic ORs.
hipError_t launch_expand_hy(int W, const int64_t *beg, const int32_t *adj, const int32_t *tadj,
                            const u64 *prev, const u64 *hs, u64 *cur, const LevelCtl *ctl_prev,
                            Queues q, int chunk_shift, HubSplit hsplit, int grid, hipStream_t st);
// Device-driven expansion of plan->level (any direction, or a sparse tail
// level appending to tl), fixed grid.
hipError_t launch_expand_dd(int W, const int64_t *beg, const int32_t *adj, const int32_t *tadj,
                            u64 *ring, uint8_t *flags, int R, int64_t V, const LevelCtl *ctl_ring,
                            Plan *plan, const u64 *hs, Queues q, int chunk_shift,
                            HubSplit hsplit, TailLog tl, int grid, hipStream_t st);
// After the emission of a window that ends at a sparse tail's first level T
// (whose slot holds the final words, so every tail bit was written as
// T - admit): records [0, n) of level > T write lev - admit into the rows of
// the selected lanes (vid: traversal id -> original id), or, with sum_out,
// add lev - T to the lanes' closeness sums.
hipError_t launch_tail_patch(int W, TailLog tl, int64_t n, int T, const int32_t *vid,
                             const LaneOut *lanes, const u64 *sel, int k_lanes, u64 *sum_out,
                             hipStream_t st);
// Device-driven direction decision for plan->level + 1; a bottom-up choice
// becomes a split level when some active lanes hold fewer than
// plan->split_cov % of the top `hubs` hubs in the hub cache hs (W words per
// hub).
hipError_t launch_plan_decide(Plan *plan, const LevelCtl *ctl_ring, int R, const u64 *hs, int W,
                              int hubs, hipStream_t st);
// Depth emission over the snapshot window [b, e] (ring slots L % R, each
// Vt*W words).  level_masks[(L-b)*W + w] = lanes admitted at or before L.
// mode 0: one vertex per thread, per-lane 4-byte stores; mode 1: LDS-tiled,
// 16-byte stores when vec4 (all lane outputs 16-byte aligned, V % 4 == 0);
// mode 2: as 0 with non-temporal stores; mode 3: a wave per status word of
// 256 vertices, 16-byte stores (1-KB pieces of each lane's row).
// edges_out (final window only, or null): += sum over v of deg(v) * |S_e[v] &
// sel|, the TEPS numerator of the selected lanes, counted during the pass.
hipError_t launch_emit(int W, int mode, int vec4, const u64 *ring, int R, EmitSpace sp, int b,
                       int e, const LaneOut *lanes, const u64 *level_masks, const u64 *sel,
                       int k_lanes, u64 *edges_out, hipStream_t st);
// Closeness reduction of the window [b, e]: sum_out[k] += sum of depths and
// reach_out[k] += vertices reached, over the bits of lane k that arrived in
// the window (lane table and masks as for launch_emit).  Isolated sources
// are not counted here (their only reached vertex is themselves, depth 0).
hipError_t launch_close(int W, const u64 *ring, int R, EmitSpace sp, int b, int e,
                        const LaneOut *lanes, const u64 *level_masks, const u64 *sel, int k_lanes,
                        u64 *sum_out, u64 *reach_out, hipStream_t st);
hipError_t launch_count_edges(int W, const u64 *S, const int32_t *deg, int64_t V,
                              const u64 *lanes, u64 *out, hipStream_t st);
// Bottom-up work model of one level: out[0] += candidates with work,
// out[1] += sum of their sequential early-termination prefixes.
hipError_t launch_bu_model(int W, const int64_t *beg, const int32_t *tadj, const u64 *prev,
                           const u64 *hs, const u64 *lam_prev, int64_t V, u64 *out,
                           hipStream_t st);
// key[i] = sum of the degrees of src[i]'s neighbours (src in original ids).
hipError_t launch_two_hop(const int64_t *beg, const int32_t *adj, const int32_t *deg,
                          const int32_t *cid, const int32_t *src, int64_t n, u64 *key,
                          hipStream_t st);
// GroupBy (groupby.cu:36-149): key[i] = first neighbour of degree > 128 of a
// source of degree <= 128, else -1; then the stable placement, as a stable
// radix sort (O(n)).  src holds original ids (cid maps them to the traversal
// CSR); keys are traversal ids, whose order is the original order.  ws: device
// workspace of groupby_workspace_bytes(n) bytes.
size_t groupby_workspace_bytes(int64_t n);
hipError_t launch_groupby(const int64_t *beg, const int32_t *adj, const int32_t *deg,
                          const int32_t *cid, const int32_t *src, int64_t n, int32_t *key,
                          int32_t *out, void *ws, size_t ws_bytes, hipStream_t st);

}  // namespace ibfs
