// Hand-written gfx950 (wave64) kernels of the concurrent multi-source BFS.
//
// Status words: W = lane_bits/64 u64 words per vertex, vertex-major
// (S[v*W + w]); lane k is bit k%64 of word k/64, the same memory image as
// the reference's uint4 with bit k%32 of component k/32 (util.h:117-130).
// Level L's words live in ring slot L % R of R snapshots (V*W words each).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ibfs {
namespace {

constexpr int BS = 256;  // threads per workgroup (4 waves)

template <int W>
__device__ __forceinline__ void load_word(const u64 *p, u64 (&x)[W]) {
  if constexpr (W == 1) {
    x[0] = p[0];
  } else {
    const ulonglong2 *q = reinterpret_cast<const ulonglong2 *>(p);
#pragma unroll
    for (int i = 0; i < W / 2; ++i) {
      ulonglong2 t = q[i];
      x[2 * i] = t.x;
      x[2 * i + 1] = t.y;
    }
  }
}

template <int W>
__device__ __forceinline__ void store_word(u64 *p, const u64 (&x)[W]) {
  if constexpr (W == 1) {
    p[0] = x[0];
  } else {
    ulonglong2 *q = reinterpret_cast<ulonglong2 *>(p);
#pragma unroll
    for (int i = 0; i < W / 2; ++i) q[i] = make_ulonglong2(x[2 * i], x[2 * i + 1]);
  }
}

template <int W>
__device__ __forceinline__ bool covers(const u64 (&x)[W], const u64 (&m)[W]) {
  bool c = true;
#pragma unroll
  for (int w = 0; w < W; ++w) c &= ((x[w] & m[w]) == m[w]);
  return c;
}

__device__ __forceinline__ u64 wave_or(u64 x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ u64 wave_sum(u64 x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ unsigned wave_incl_scan(unsigned x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    unsigned y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

__device__ __forceinline__ int degree_class(int d, int sml, int lrg) {
  return d < sml ? 1 : (d > lrg ? 3 : 2);
}

// ---------------------------------------------------------------------------
__global__ void k_fill_u64(u64 *p, int64_t n, u64 value) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    p[i] = value;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t value) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    p[i] = value;
}

// Lane init: lane lane0+i gets bit of src[i].  OR, so duplicate sources
// and zero-degree sources keep their bit (the reference stores and skips
// zero-degree sources, expander.cu:63,70-71).
template <int W>
__global__ void k_init_sources(const int32_t *src, int k, int lane0, u64 *S0) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= k) return;
  int lane = lane0 + i;
  atomicOr(&S0[(int64_t)src[i] * W + (lane >> 6)], 1ull << (lane & 63));
}

// ---------------------------------------------------------------------------
// Inspector of level L: one pass over all vertices.  Diffs cur (S_L) against
// prev (S_{L-1}), copies S_L into next (S_{L+1}, the slot the expanders of
// L+1 update), accumulates the lanes that advanced (lam of L), and compacts,
// per degree class, both the top-down frontier (changed, degree > 0) and the
// bottom-up candidates (degree > 0 and missing a lane of lam_prev) of level
// L+1.  Replaces td/sw_frontier_count + insp_scan + td/sw_frontier_gather +
// extract_depth of inspector.cu:37-500 and scan.cuh:455-904 with a single
// kernel: per-tile flags in LDS, a workgroup scan and one atomic per class
// and tile.  Queue order is ascending vertex order within a tile.
template <int W, int VPT>
__global__ __launch_bounds__(BS) void k_inspect(
    const u64 *__restrict__ prev, const u64 *__restrict__ cur, u64 *__restrict__ next,
    const int32_t *__restrict__ deg, int64_t V, const u64 *__restrict__ lam_prev,
    const u64 *__restrict__ prev_keep, const u64 *__restrict__ next_and, LevelCtl *ctl,
    LevelCtl *ctl_zero, Queues q, int sml, int lrg, int chunk_shift) {
  constexpr int TILE = BS * VPT;
  __shared__ uint8_t fl[TILE];
  __shared__ unsigned wsum[4][6];
  __shared__ unsigned qbase[6];
  __shared__ u64 red[4][5 + W];

  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (blockIdx.x == 0 && ctl_zero && t < (int)(sizeof(LevelCtl) / 8))
    reinterpret_cast<u64 *>(ctl_zero)[t] = 0;

  u64 lp[W], keep[W], nand[W], lamacc[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    lp[w] = lam_prev[w];
    keep[w] = prev_keep ? prev_keep[w] : ~0ull;
    nand[w] = next_and ? next_and[w] : ~0ull;
    lamacc[w] = 0;
  }
  u64 n_changed = 0, n_front = 0, m_front = 0, n_cand = 0, m_cand = 0;

  const int64_t ntiles = (V + TILE - 1) / TILE;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t t0 = tile * TILE;
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int64_t v = t0 + i * BS + t;
      uint8_t f = 0;
      if (v < V) {
        u64 c[W], p[W], n[W];
        load_word<W>(cur + v * W, c);
        if (prev) {
          load_word<W>(prev + v * W, p);
        } else {
#pragma unroll
          for (int w = 0; w < W; ++w) p[w] = 0;
        }
        bool ch = false, cover = true;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          u64 d = c[w] & ~(p[w] & keep[w]);
          lamacc[w] |= d;
          ch |= (d != 0);
          n[w] = c[w] & nand[w];
          cover &= ((c[w] & lp[w]) == lp[w]);
        }
        store_word<W>(next + v * W, n);
        const int dg = deg[v];
        if (ch) {
          ++n_changed;
          if (dg > 0) {
            ++n_front;
            m_front += (u64)dg;
            f |= (uint8_t)degree_class(dg, sml, lrg);
          }
        }
        if (dg > 0 && !cover) {
          ++n_cand;
          m_cand += (u64)dg;
          f |= (uint8_t)(degree_class(dg, sml, lrg) << 2);
        }
      }
      fl[i * BS + t] = f;
    }
    __syncthreads();

    unsigned cnt[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
      const uint8_t f = fl[t * VPT + j];
      const int tc = f & 3, bc = (f >> 2) & 3;
      if (tc == 1) {
        ++cnt[Q_TD_S];
      } else if (tc == 2) {
        ++cnt[Q_TD_M];
      } else if (tc == 3) {
        const int dg = deg[t0 + t * VPT + j];
        cnt[Q_TD_L] += (unsigned)((dg + (1 << chunk_shift) - 1) >> chunk_shift);
      }
      if (bc) ++cnt[Q_BU_S + bc - 1];
    }
    unsigned ex[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const unsigned inc = wave_incl_scan(cnt[k], lane);
      ex[k] = inc - cnt[k];
      if (lane == 63) wsum[wv][k] = inc;
    }
    __syncthreads();
    if (t < 6) {
      const unsigned tot = wsum[0][t] + wsum[1][t] + wsum[2][t] + wsum[3][t];
      qbase[t] = tot ? atomicAdd(&ctl->q[t], tot) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      unsigned add = qbase[k];
      for (int w = 0; w < wv; ++w) add += wsum[w][k];
      ex[k] += add;
    }
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
      const uint8_t f = fl[t * VPT + j];
      if (!f) continue;
      const int32_t v = (int32_t)(t0 + t * VPT + j);
      const int tc = f & 3, bc = (f >> 2) & 3;
      if (tc == 1) {
        q.q[Q_TD_S][ex[Q_TD_S]++] = v;
      } else if (tc == 2) {
        q.q[Q_TD_M][ex[Q_TD_M]++] = v;
      } else if (tc == 3) {
        const int nch = (deg[v] + (1 << chunk_shift) - 1) >> chunk_shift;
        for (int c = 0; c < nch; ++c) q.chunks[ex[Q_TD_L]++] = make_int2(v, c);
      }
      if (bc == 1) q.q[Q_BU_S][ex[Q_BU_S]++] = v;
      else if (bc == 2) q.q[Q_BU_M][ex[Q_BU_M]++] = v;
      else if (bc == 3) q.q[Q_BU_L][ex[Q_BU_L]++] = v;
    }
    __syncthreads();
  }

  // workgroup reduction of the level counters, one atomic each
  u64 vals[5 + W] = {n_changed, n_front, m_front, n_cand, m_cand};
#pragma unroll
  for (int w = 0; w < W; ++w) vals[5 + w] = lamacc[w];
#pragma unroll
  for (int i = 0; i < 5; ++i) vals[i] = wave_sum(vals[i]);
#pragma unroll
  for (int w = 0; w < W; ++w) vals[5 + w] = wave_or(vals[5 + w]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 5 + W; ++i) red[wv][i] = vals[i];
  }
  __syncthreads();
  if (t < 5) {
    const u64 s = red[0][t] + red[1][t] + red[2][t] + red[3][t];
    if (s) atomicAdd(&(&ctl->n_changed)[t], s);
  } else if (t < 5 + W) {
    const u64 s = red[0][t] | red[1][t] | red[2][t] | red[3][t];
    if (s) atomicOr(&ctl->lam[t - 5], s);
  }
}

// ---------------------------------------------------------------------------
// Top-down expansion of level L: every frontier vertex u pushes
// S_{L-1}[u] & lam_{L-1} into S_L of its neighbours with 64-bit atomic ORs,
// skipping neighbours that already hold the bits (expander.cu:100-252).
// Workgroups [0, nbs) take small vertices one per thread, [nbs, nbs+nbm) mid
// vertices one per wave, the rest one chunk of 2^chunk_shift edges of a
// large vertex per workgroup.
template <int W>
__device__ __forceinline__ void td_push(const int32_t *__restrict__ adj, u64 *cur, const u64 (&p)[W],
                                        int64_t j, int64_t end, int stride) {
  for (; j < end; j += stride) {
    const int64_t v = adj[j];
    u64 c[W];
    load_word<W>(cur + v * W, c);
#pragma unroll
    for (int w = 0; w < W; ++w)
      if (p[w] & ~c[w]) atomicOr(&cur[v * W + w], p[w]);
  }
}

template <int W>
__global__ __launch_bounds__(BS) void k_expand_td(const int64_t *__restrict__ beg,
                                                  const int32_t *__restrict__ adj,
                                                  const u64 *__restrict__ prev, u64 *cur,
                                                  const u64 *__restrict__ lam_prev, Queues q,
                                                  int64_t ns, int64_t nm, int64_t nl, int64_t nbs,
                                                  int64_t nbm, int chunk_shift) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t b = blockIdx.x;
  u64 lam[W];
#pragma unroll
  for (int w = 0; w < W; ++w) lam[w] = lam_prev[w];
  int32_t u;
  int64_t j0, j1;
  int stride;
  if (b < nbs) {
    const int64_t i = b * BS + t;
    if (i >= ns) return;
    u = q.q[Q_TD_S][i];
    j0 = beg[u];
    j1 = beg[u + 1];
    stride = 1;
  } else if (b < nbs + nbm) {
    const int64_t i = (b - nbs) * 4 + wv;
    if (i >= nm) return;
    u = q.q[Q_TD_M][i];
    j0 = beg[u] + lane;
    j1 = beg[u + 1];
    stride = 64;
  } else {
    const int64_t i = b - nbs - nbm;
    if (i >= nl) return;
    const int2 e = q.chunks[i];
    u = e.x;
    const int64_t lo = beg[u] + ((int64_t)e.y << chunk_shift);
    const int64_t hi = beg[u + 1];
    j0 = lo + t;
    j1 = (lo + (1ll << chunk_shift)) < hi ? (lo + (1ll << chunk_shift)) : hi;
    stride = BS;
  }
  u64 p[W];
  load_word<W>(prev + (int64_t)u * W, p);
#pragma unroll
  for (int w = 0; w < W; ++w) p[w] &= lam[w];
  td_push<W>(adj, cur, p, j0, j1, stride);
}

// ---------------------------------------------------------------------------
// Hub status cache: hs[i] = S[hubs[i]] for the H highest-degree vertices,
// refreshed after every expansion so that bottom-up gathers of hub words hit
// a compact, L2-resident array.
template <int W>
__global__ __launch_bounds__(BS) void k_hub_refresh(const u64 *__restrict__ S,
                                                    const int32_t *__restrict__ hubs, int64_t H,
                                                    u64 *__restrict__ hs) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= H) return;
  u64 x[W];
  load_word<W>(S + (int64_t)hubs[i] * W, x);
  store_word<W>(hs + i * W, x);
}

// Word of traversal-adjacency entry a: a >= 0 is a vertex id (read from the
// snapshot), a < 0 is hub slot a & 0x7fffffff (read from the hub cache).
template <int W>
__device__ __forceinline__ void load_nbr(const u64 *__restrict__ prev, const u64 *__restrict__ hs,
                                         int32_t a, u64 (&y)[W]) {
  const u64 *p = (a < 0) ? hs + (int64_t)(a & 0x7fffffff) * W : prev + (int64_t)a * W;
  load_word<W>(p, y);
}

// ---------------------------------------------------------------------------
// Bottom-up expansion of level L: a candidate v ORs S_{L-1} of its
// neighbours (restricted to lam_{L-1}, the only lanes that can still be new)
// in adjacency order and stops as soon as its word covers lam_{L-1}
// (bitwise early termination, expander.cu:258-359).  The traversal adjacency
// `tadj` lists each vertex's neighbours by descending degree (hubs first, so
// coverage is reached sooner) with hub entries redirected to the hub cache.
// Small candidates: one per thread, four neighbours per step; mid: one per
// wave, 64 neighbours per step with a wave OR-reduction; large: one per
// workgroup, 256 neighbours per step with the four wave partials combined
// through LDS.
template <int W>
__global__ __launch_bounds__(BS) void k_expand_bu(const int64_t *__restrict__ beg,
                                                  const int32_t *__restrict__ adj,
                                                  const u64 *__restrict__ prev,
                                                  const u64 *__restrict__ hs,
                                                  u64 *__restrict__ cur,
                                                  const u64 *__restrict__ lam_prev, Queues q,
                                                  int64_t ns, int64_t nm, int64_t nl, int64_t nbs,
                                                  int64_t nbm) {
  __shared__ u64 part[4][W];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t b = blockIdx.x;
  u64 lam[W];
#pragma unroll
  for (int w = 0; w < W; ++w) lam[w] = lam_prev[w];

  if (b < nbs) {
    const int64_t i = b * BS + t;
    if (i >= ns) return;
    const int64_t v = q.q[Q_BU_S][i];
    u64 x[W], x0[W];
    load_word<W>(cur + v * W, x);
#pragma unroll
    for (int w = 0; w < W; ++w) x0[w] = x[w];
    if (covers<W>(x, lam)) return;
    int64_t j = beg[v];
    const int64_t e = beg[v + 1];
    while (j < e) {
      const int32_t a0 = adj[j];
      const int32_t a1 = (j + 1 < e) ? adj[j + 1] : a0;
      const int32_t a2 = (j + 2 < e) ? adj[j + 2] : a0;
      const int32_t a3 = (j + 3 < e) ? adj[j + 3] : a0;
      u64 s0[W], s1[W], s2[W], s3[W];
      load_nbr<W>(prev, hs, a0, s0);
      load_nbr<W>(prev, hs, a1, s1);
      load_nbr<W>(prev, hs, a2, s2);
      load_nbr<W>(prev, hs, a3, s3);
#pragma unroll
      for (int w = 0; w < W; ++w) x[w] |= (s0[w] | s1[w] | s2[w] | s3[w]) & lam[w];
      j += 4;
      if (covers<W>(x, lam)) break;
    }
    bool ch = false;
#pragma unroll
    for (int w = 0; w < W; ++w) ch |= (x[w] != x0[w]);
    if (ch) store_word<W>(cur + v * W, x);
    return;
  }

  if (b < nbs + nbm) {
    const int64_t i = (b - nbs) * 4 + wv;
    if (i >= nm) return;
    const int64_t v = q.q[Q_BU_M][i];
    u64 x[W], x0[W];
    load_word<W>(cur + v * W, x);
#pragma unroll
    for (int w = 0; w < W; ++w) x0[w] = x[w];
    if (covers<W>(x, lam)) return;
    const int64_t s = beg[v], e = beg[v + 1];
    for (int64_t r = s; r < e; r += 64) {
      const int64_t j = r + lane;
      u64 y[W];
      if (j < e) {
        load_nbr<W>(prev, hs, adj[j], y);
      } else {
#pragma unroll
        for (int w = 0; w < W; ++w) y[w] = 0;
      }
#pragma unroll
      for (int w = 0; w < W; ++w) x[w] |= wave_or(y[w] & lam[w]);
      if (covers<W>(x, lam)) break;
    }
    bool ch = false;
#pragma unroll
    for (int w = 0; w < W; ++w) ch |= (x[w] != x0[w]);
    if (lane == 0 && ch) store_word<W>(cur + v * W, x);
    return;
  }

  const int64_t i = b - nbs - nbm;
  if (i >= nl) return;
  const int64_t v = q.q[Q_BU_L][i];
  u64 x[W], x0[W];
  load_word<W>(cur + v * W, x);
#pragma unroll
  for (int w = 0; w < W; ++w) x0[w] = x[w];
  if (covers<W>(x, lam)) return;
  const int64_t s = beg[v], e = beg[v + 1];
  for (int64_t r = s; r < e; r += BS) {
    const int64_t j = r + t;
    u64 y[W];
    if (j < e) {
      load_nbr<W>(prev, hs, adj[j], y);
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w) y[w] = 0;
    }
#pragma unroll
    for (int w = 0; w < W; ++w) y[w] = wave_or(y[w] & lam[w]);
    if (lane == 0) {
#pragma unroll
      for (int w = 0; w < W; ++w) part[wv][w] = y[w];
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] |= part[0][w] | part[1][w] | part[2][w] | part[3][w];
    __syncthreads();
    if (covers<W>(x, lam)) break;
  }
  bool ch = false;
#pragma unroll
  for (int w = 0; w < W; ++w) ch |= (x[w] != x0[w]);
  if (t == 0 && ch) store_word<W>(cur + v * W, x);
}

// ---------------------------------------------------------------------------
// Depth emission over the snapshot window [b, e]: one thread per vertex
// counts, per lane, the admitted levels of the window at which the lane's
// bit is absent (a bit-sliced counter of NS bits), which is the lane's depth
// relative to its admission level because status words only gain bits.
// Stores to lane k's array are consecutive across the wave (one 256-byte
// segment per store instruction).  A lane's first window (admit >= b) writes
// every vertex (-1 where the bit never arrived); later windows write only the
// vertices whose bit arrived inside them.  The lane table is staged in LDS:
// reading it from global memory inside the lane loop would put a memory
// round trip on every lane of every vertex.
typedef __attribute__((address_space(1))) int32_t gint32;

__device__ __forceinline__ void put_depth(int32_t *out, int64_t v, int32_t val, bool nt) {
  gint32 *o = (gint32 *)out;
  if (nt) __builtin_nontemporal_store(val, o + v);
  else o[v] = val;
}

template <int W, int NS>
__device__ __forceinline__ void emit_count(const u64 *__restrict__ ring, int R, int64_t V, int b,
                                           int e, int64_t v, const u64 *__restrict__ level_masks,
                                           u64 (&c)[NS][W], u64 (&fin)[W], u64 (&bs)[W]) {
  const int64_t slot = V * W;
#pragma unroll
  for (int w = 0; w < W; ++w) {
#pragma unroll
    for (int j = 0; j < NS; ++j) c[j][w] = 0;
  }
  load_word<W>(ring + (int64_t)(e % R) * slot + v * W, fin);
  load_word<W>(ring + (int64_t)(b % R) * slot + v * W, bs);
  for (int L = b; L <= e; ++L) {
    u64 s[W];
    load_word<W>(ring + (int64_t)(L % R) * slot + v * W, s);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      u64 carry = ~s[w] & level_masks[(int64_t)(L - b) * W + w];
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const u64 nc = c[j][w] & carry;
        c[j][w] ^= carry;
        carry = nc;
      }
    }
  }
}

template <int W, int NS>
__global__ __launch_bounds__(BS) void k_emit(const u64 *__restrict__ ring, int R, int64_t V, int b,
                                             int e, const LaneOut *__restrict__ lanes,
                                             const u64 *__restrict__ level_masks,
                                             const u64 *__restrict__ sel, int k_lanes, int nt) {
  __shared__ int32_t *s_out[4 * 64];
  __shared__ int32_t s_off[4 * 64];  // b - admit: <= 0 in the lane's first window
  for (int i = threadIdx.x; i < k_lanes; i += BS) {
    const bool on = (sel[i >> 6] >> (i & 63)) & 1ull;
    s_out[i] = on ? lanes[i].out : nullptr;
    s_off[i] = b - lanes[i].admit;
  }
  __syncthreads();
  for (int64_t v = (int64_t)blockIdx.x * BS + threadIdx.x; v < V; v += (int64_t)gridDim.x * BS) {
    u64 c[NS][W], fin[W], bs[W];
    emit_count<W, NS>(ring, R, V, b, e, v, level_masks, c, fin, bs);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      for (int bit = 0; bit < 64; ++bit) {
        const int k = w * 64 + bit;
        if (k >= k_lanes) break;
        int32_t *out = s_out[k];
        if (!out) continue;
        const int off = s_off[k];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < NS; ++j) cnt |= (int)((c[j][w] >> bit) & 1ull) << j;
        const bool set = (fin[w] >> bit) & 1ull;
        if (off <= 0) put_depth(out, v, set ? cnt : -1, nt);
        else if (set && !((bs[w] >> bit) & 1ull)) put_depth(out, v, off + cnt, nt);
      }
    }
  }
}

// Tiled depth emission: same values as k_emit, but a workgroup owns 256
// consecutive vertices and stages 32 lanes at a time through LDS, so that
// each wave writes one lane's 256 depths as a single 1-KB contiguous segment
// with 16-byte stores.  Entries a continuation window must not touch carry
// the sentinel INT_MIN and are skipped.  vec4 = every lane's output is
// 16-byte aligned and V % 4 == 0.
template <int W, int NS>
__global__ __launch_bounds__(BS) void k_emit_tiled(const u64 *__restrict__ ring, int R, int64_t V,
                                                   int b, int e,
                                                   const LaneOut *__restrict__ lanes,
                                                   const u64 *__restrict__ level_masks,
                                                   const u64 *__restrict__ sel, int k_lanes,
                                                   int vec4) {
  constexpr int32_t SKIP = (int32_t)0x80000000;
  __shared__ int4 tile4[32 * 64];  // [row][64 x int4] = 32 lanes x 256 vertices
  __shared__ int32_t *s_out[4 * 64];
  __shared__ int32_t s_off[4 * 64];
  int32_t *tile = reinterpret_cast<int32_t *>(tile4);
  const int t = threadIdx.x;
  for (int i = t; i < k_lanes; i += BS) {
    const bool on = (sel[i >> 6] >> (i & 63)) & 1ull;
    s_out[i] = on ? lanes[i].out : nullptr;
    s_off[i] = b - lanes[i].admit;
  }
  __syncthreads();
  const int64_t nblk = (V + BS - 1) / BS;
  for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int64_t v = blk * BS + t;
    u64 c[NS][W], fin[W], bs[W];
    if (v < V) {
      emit_count<W, NS>(ring, R, V, b, e, v, level_masks, c, fin, bs);
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        fin[w] = bs[w] = 0;
#pragma unroll
        for (int j = 0; j < NS; ++j) c[j][w] = 0;
      }
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int lane0 = w * 64 + h * 32;
        if (lane0 >= k_lanes) break;  // uniform
        for (int r = 0; r < 32; ++r) {
          const int k = lane0 + r;
          const int bit = h * 32 + r;
          int32_t val = SKIP;
          if (k < k_lanes && s_out[k]) {
            const int off = s_off[k];
            int cnt = 0;
#pragma unroll
            for (int j = 0; j < NS; ++j) cnt |= (int)((c[j][w] >> bit) & 1ull) << j;
            const bool set = (fin[w] >> bit) & 1ull;
            if (off <= 0) val = set ? cnt : -1;
            else if (set && !((bs[w] >> bit) & 1ull)) val = off + cnt;
          }
          tile[r * BS + t] = val;
        }
        __syncthreads();
        for (int m = 0; m < 8; ++m) {
          const int i = t + BS * m;
          const int row = i >> 6, col4 = i & 63;
          const int k = lane0 + row;
          int32_t *out = (k < k_lanes) ? s_out[k] : nullptr;
          if (out) {
            const int4 q = tile4[row * 64 + col4];
            const int64_t v0 = blk * BS + 4 * col4;
            const bool full = vec4 && (v0 + 3 < V) && q.x != SKIP && q.y != SKIP &&
                              q.z != SKIP && q.w != SKIP;
            gint32 *o = (gint32 *)out;
            if (full) {
              typedef int v4i __attribute__((ext_vector_type(4)));
              v4i x = {q.x, q.y, q.z, q.w};
              *reinterpret_cast<__attribute__((address_space(1))) v4i *>(o + v0) = x;
            } else {
              if (v0 < V && q.x != SKIP) o[v0] = q.x;
              if (v0 + 1 < V && q.y != SKIP) o[v0 + 1] = q.y;
              if (v0 + 2 < V && q.z != SKIP) o[v0 + 2] = q.z;
              if (v0 + 3 < V && q.w != SKIP) o[v0 + 3] = q.w;
            }
          }
        }
        __syncthreads();
      }
    }
  }
}

// Sum over v of deg(v) * |S[v] & lanes|: the TEPS numerator of all lanes
// (validate.h:59-76 summed over sources).
template <int W>
__global__ __launch_bounds__(BS) void k_count_edges(const u64 *__restrict__ S,
                                                    const int32_t *__restrict__ deg, int64_t V,
                                                    const u64 *__restrict__ lanes, u64 *out) {
  __shared__ u64 red[4];
  u64 acc = 0;
  for (int64_t v = (int64_t)blockIdx.x * BS + threadIdx.x; v < V; v += (int64_t)gridDim.x * BS) {
    u64 x[W];
    load_word<W>(S + v * W, x);
    int pc = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) pc += __popcll(x[w] & lanes[w]);
    acc += (u64)pc * (u64)deg[v];
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const u64 s = red[0] + red[1] + red[2] + red[3];
    if (s) atomicAdd(out, s);
  }
}

// Work model of a bottom-up level (instrumentation, not on the timed path):
// out[0] = vertices of degree > 0 whose S_{L-1} word misses a lane of
// lam_{L-1}; out[1] = sum over them of the sequential adjacency prefix after
// which the word covers lam_{L-1} (the whole list if never).
template <int W>
__global__ __launch_bounds__(BS) void k_bu_model(const int64_t *__restrict__ beg,
                                                 const int32_t *__restrict__ adj,
                                                 const u64 *__restrict__ prev,
                                                 const u64 *__restrict__ hs,
                                                 const u64 *__restrict__ lam_prev, int64_t V,
                                                 u64 *out) {
  __shared__ u64 red[2][4];
  u64 lam[W];
#pragma unroll
  for (int w = 0; w < W; ++w) lam[w] = lam_prev[w];
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
      load_nbr<W>(prev, hs, adj[j], y);
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

// ---------------------------------------------------------------------------
// GroupBy.  Key of source i: its first neighbour (adjacency order) of degree
// > 128 if deg(src[i]) <= 128, else -1 (groupby.cu:49-66); one wave per
// source scans 64 neighbours per step and takes the lowest set ballot lane.
__global__ __launch_bounds__(BS) void k_groupby_key(const int64_t *__restrict__ beg,
                                                    const int32_t *__restrict__ adj,
                                                    const int32_t *__restrict__ deg,
                                                    const int32_t *__restrict__ src, int64_t n,
                                                    int32_t *__restrict__ key) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int32_t s = src[i];
  int32_t h = -1;
  if (deg[s] <= 128) {
    const int64_t e = beg[s + 1];
    for (int64_t j0 = beg[s]; j0 < e; j0 += 64) {
      const int64_t j = j0 + lane;
      const bool hub = (j < e) && (deg[adj[j]] > 128);
      const u64 m = __ballot(hub);
      if (m) {
        h = adj[j0 + __ffsll((unsigned long long)m) - 1];
        break;
      }
    }
  }
  if (lane == 0) key[i] = h;
}

// Stable placement (groupby.cu:69-115): an eligible source goes after every
// eligible source with a smaller hub id or the same hub and a smaller input
// index; ineligible sources follow all eligible ones in input order.  Each
// thread counts its rank over all keys staged through LDS in tiles.
__global__ __launch_bounds__(BS) void k_groupby_place(const int32_t *__restrict__ key,
                                                      const int32_t *__restrict__ src, int64_t n,
                                                      int32_t *__restrict__ out) {
  __shared__ int32_t tile[1024];
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int32_t ki = (i < n) ? key[i] : 0;
  int64_t less = 0, elig = 0, inel_before = 0;
  for (int64_t base = 0; base < n; base += 1024) {
    for (int r = threadIdx.x; r < 1024; r += BS)
      tile[r] = (base + r < n) ? key[base + r] : 0;
    __syncthreads();
    const int m = (n - base) < 1024 ? (int)(n - base) : 1024;
    for (int jj = 0; jj < m; ++jj) {
      const int32_t kj = tile[jj];
      const int64_t j = base + jj;
      if (kj >= 0) {
        ++elig;
        if (ki >= 0 && (kj < ki || (kj == ki && j < i))) ++less;
      } else if (j < i) {
        ++inel_before;
      }
    }
    __syncthreads();
  }
  if (i < n) out[(ki >= 0) ? less : (elig + inel_before)] = src[i];
}

inline int grid_for(int64_t n, int per_block, int cap) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace

// ---------------------------------------------------------------------------
hipError_t launch_fill_u64(u64 *p, int64_t n, u64 value, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, BS * 8, 8192)), dim3(BS), 0, st, p, n, value);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t value, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, BS * 8, 8192)), dim3(BS), 0, st, p, n, value);
  return hipGetLastError();
}

hipError_t launch_init_sources(int W, const int32_t *src, int k, int lane0, u64 *S0,
                               hipStream_t st) {
  if (k <= 0) return hipSuccess;
  dim3 g(grid_for(k, BS, 1 << 20)), blk(BS);
  switch (W) {
    case 1: hipLaunchKernelGGL(k_init_sources<1>, g, blk, 0, st, src, k, lane0, S0); break;
    case 2: hipLaunchKernelGGL(k_init_sources<2>, g, blk, 0, st, src, k, lane0, S0); break;
    case 4: hipLaunchKernelGGL(k_init_sources<4>, g, blk, 0, st, src, k, lane0, S0); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int W>
static void inspect_w(const u64 *prev, const u64 *cur, u64 *next, const int32_t *deg, int64_t V,
                      const u64 *lam_prev, const u64 *keep, const u64 *nand, LevelCtl *ctl,
                      LevelCtl *ctl_zero, Queues q, int sml, int lrg, int cs, hipStream_t st) {
  // vertices per thread: 16 when that still gives >= 1024 tiles, else 4, else 1
  if (V >= (int64_t)BS * 16 * 1024) {
    hipLaunchKernelGGL((k_inspect<W, 16>), dim3(grid_for(V, BS * 16, 4096)), dim3(BS), 0, st,
                       prev, cur, next, deg, V, lam_prev, keep, nand, ctl, ctl_zero, q, sml, lrg, cs);
  } else if (V >= (int64_t)BS * 4 * 512) {
    hipLaunchKernelGGL((k_inspect<W, 4>), dim3(grid_for(V, BS * 4, 4096)), dim3(BS), 0, st,
                       prev, cur, next, deg, V, lam_prev, keep, nand, ctl, ctl_zero, q, sml, lrg, cs);
  } else {
    hipLaunchKernelGGL((k_inspect<W, 1>), dim3(grid_for(V, BS, 4096)), dim3(BS), 0, st, prev,
                       cur, next, deg, V, lam_prev, keep, nand, ctl, ctl_zero, q, sml, lrg, cs);
  }
}

hipError_t launch_inspect(int W, const u64 *prev, const u64 *cur, u64 *next, const int32_t *deg,
                          int64_t V, const u64 *lam_prev, const u64 *prev_keep,
                          const u64 *next_and, LevelCtl *ctl_out, LevelCtl *ctl_zero, Queues q,
                          int sml, int lrg, int chunk_shift, hipStream_t st) {
  if (V <= 0) return hipSuccess;
  switch (W) {
    case 1: inspect_w<1>(prev, cur, next, deg, V, lam_prev, prev_keep, next_and, ctl_out, ctl_zero, q, sml, lrg, chunk_shift, st); break;
    case 2: inspect_w<2>(prev, cur, next, deg, V, lam_prev, prev_keep, next_and, ctl_out, ctl_zero, q, sml, lrg, chunk_shift, st); break;
    case 4: inspect_w<4>(prev, cur, next, deg, V, lam_prev, prev_keep, next_and, ctl_out, ctl_zero, q, sml, lrg, chunk_shift, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_expand_td(int W, const int64_t *beg, const int32_t *adj, const u64 *prev,
                            u64 *cur, const u64 *lam_prev, Queues q, int64_t ns, int64_t nm,
                            int64_t nl, int chunk_shift, hipStream_t st) {
  const int64_t nbs = (ns + BS - 1) / BS, nbm = (nm + 3) / 4;
  const int64_t nb = nbs + nbm + nl;
  if (nb == 0) return hipSuccess;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  dim3 g((unsigned)nb), blk(BS);
  switch (W) {
    case 1: hipLaunchKernelGGL(k_expand_td<1>, g, blk, 0, st, beg, adj, prev, cur, lam_prev, q, ns, nm, nl, nbs, nbm, chunk_shift); break;
    case 2: hipLaunchKernelGGL(k_expand_td<2>, g, blk, 0, st, beg, adj, prev, cur, lam_prev, q, ns, nm, nl, nbs, nbm, chunk_shift); break;
    case 4: hipLaunchKernelGGL(k_expand_td<4>, g, blk, 0, st, beg, adj, prev, cur, lam_prev, q, ns, nm, nl, nbs, nbm, chunk_shift); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_expand_bu(int W, const int64_t *beg, const int32_t *tadj, const u64 *prev,
                            const u64 *hs, u64 *cur, const u64 *lam_prev, Queues q, int64_t ns,
                            int64_t nm, int64_t nl, hipStream_t st) {
  const int64_t nbs = (ns + BS - 1) / BS, nbm = (nm + 3) / 4;
  const int64_t nb = nbs + nbm + nl;
  if (nb == 0) return hipSuccess;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  dim3 g((unsigned)nb), blk(BS);
  switch (W) {
    case 1: hipLaunchKernelGGL(k_expand_bu<1>, g, blk, 0, st, beg, tadj, prev, hs, cur, lam_prev, q, ns, nm, nl, nbs, nbm); break;
    case 2: hipLaunchKernelGGL(k_expand_bu<2>, g, blk, 0, st, beg, tadj, prev, hs, cur, lam_prev, q, ns, nm, nl, nbs, nbm); break;
    case 4: hipLaunchKernelGGL(k_expand_bu<4>, g, blk, 0, st, beg, tadj, prev, hs, cur, lam_prev, q, ns, nm, nl, nbs, nbm); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_hub_refresh(int W, const u64 *S, const int32_t *hubs, int64_t H, u64 *hs,
                              hipStream_t st) {
  if (H <= 0) return hipSuccess;
  dim3 g(grid_for(H, BS, 1 << 30)), blk(BS);
  switch (W) {
    case 1: hipLaunchKernelGGL(k_hub_refresh<1>, g, blk, 0, st, S, hubs, H, hs); break;
    case 2: hipLaunchKernelGGL(k_hub_refresh<2>, g, blk, 0, st, S, hubs, H, hs); break;
    case 4: hipLaunchKernelGGL(k_hub_refresh<4>, g, blk, 0, st, S, hubs, H, hs); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int W, int NS>
static void emit_ns(int mode, int vec4, const u64 *ring, int R, int64_t V, int b, int e,
                    const LaneOut *lanes, const u64 *masks, const u64 *sel, int k,
                    hipStream_t st) {
  if (mode == 1) {
    dim3 g(grid_for(V, BS, 4096)), blk(BS);
    hipLaunchKernelGGL((k_emit_tiled<W, NS>), g, blk, 0, st, ring, R, V, b, e, lanes, masks, sel,
                       k, vec4);
  } else {
    dim3 g(grid_for(V, BS, 8192)), blk(BS);
    hipLaunchKernelGGL((k_emit<W, NS>), g, blk, 0, st, ring, R, V, b, e, lanes, masks, sel, k,
                       mode == 2 ? 1 : 0);
  }
}

template <int W>
static hipError_t emit_w(int mode, int vec4, const u64 *ring, int R, int64_t V, int b, int e,
                         const LaneOut *lanes, const u64 *masks, const u64 *sel, int k,
                         hipStream_t st) {
  const int span = e - b + 1;
  if (span < 16) emit_ns<W, 4>(mode, vec4, ring, R, V, b, e, lanes, masks, sel, k, st);
  else if (span < 64) emit_ns<W, 6>(mode, vec4, ring, R, V, b, e, lanes, masks, sel, k, st);
  else if (span < 256) emit_ns<W, 8>(mode, vec4, ring, R, V, b, e, lanes, masks, sel, k, st);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_emit(int W, int mode, int vec4, const u64 *ring, int R, int64_t V, int b, int e,
                       const LaneOut *lanes, const u64 *level_masks, const u64 *sel, int k_lanes,
                       hipStream_t st) {
  if (V <= 0 || e < b) return hipSuccess;
  switch (W) {
    case 1: return emit_w<1>(mode, vec4, ring, R, V, b, e, lanes, level_masks, sel, k_lanes, st);
    case 2: return emit_w<2>(mode, vec4, ring, R, V, b, e, lanes, level_masks, sel, k_lanes, st);
    case 4: return emit_w<4>(mode, vec4, ring, R, V, b, e, lanes, level_masks, sel, k_lanes, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_count_edges(int W, const u64 *S, const int32_t *deg, int64_t V,
                              const u64 *lanes, u64 *out, hipStream_t st) {
  if (V <= 0) return hipSuccess;
  dim3 g(grid_for(V, BS * 4, 4096)), blk(BS);
  switch (W) {
    case 1: hipLaunchKernelGGL(k_count_edges<1>, g, blk, 0, st, S, deg, V, lanes, out); break;
    case 2: hipLaunchKernelGGL(k_count_edges<2>, g, blk, 0, st, S, deg, V, lanes, out); break;
    case 4: hipLaunchKernelGGL(k_count_edges<4>, g, blk, 0, st, S, deg, V, lanes, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_bu_model(int W, const int64_t *beg, const int32_t *tadj, const u64 *prev,
                           const u64 *hs, const u64 *lam_prev, int64_t V, u64 *out,
                           hipStream_t st) {
  if (V <= 0) return hipSuccess;
  dim3 g(grid_for(V, BS, 8192)), blk(BS);
  switch (W) {
    case 1: hipLaunchKernelGGL(k_bu_model<1>, g, blk, 0, st, beg, tadj, prev, hs, lam_prev, V, out); break;
    case 2: hipLaunchKernelGGL(k_bu_model<2>, g, blk, 0, st, beg, tadj, prev, hs, lam_prev, V, out); break;
    case 4: hipLaunchKernelGGL(k_bu_model<4>, g, blk, 0, st, beg, tadj, prev, hs, lam_prev, V, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_groupby(const int64_t *beg, const int32_t *adj, const int32_t *deg,
                          const int32_t *src, int64_t n, int32_t *key, int32_t *out,
                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_groupby_key, dim3(grid_for(n, 4, 1 << 30)), dim3(BS), 0, st, beg, adj, deg,
                     src, n, key);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(k_groupby_place, dim3(grid_for(n, BS, 1 << 30)), dim3(BS), 0, st, key, src, n,
                     out);
  return hipGetLastError();
}

}  // namespace ibfs
