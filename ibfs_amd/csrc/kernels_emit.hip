// Depth emission, TEPS numerator and GroupBy kernels (gfx950, wave64).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "device_util.h"

namespace ibfs {
namespace {

using namespace dev;

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

__device__ __forceinline__ void put_depth(int32_t *out, int64_t v, int32_t val, bool nt) {
  gint32 *o = (gint32 *)out;
  if (nt) __builtin_nontemporal_store(val, o + v);
  else o[v] = val;
}

template <int W, int NS>
__device__ __forceinline__ void emit_count(const u64 *__restrict__ ring, int R, int64_t Vt, int b,
                                           int e, int64_t i, const u64 *__restrict__ level_masks,
                                           u64 (&c)[NS][W], u64 (&fin)[W], u64 (&bs)[W]) {
  const int64_t slot = Vt * W;
#pragma unroll
  for (int w = 0; w < W; ++w) {
#pragma unroll
    for (int j = 0; j < NS; ++j) c[j][w] = 0;
  }
  // each snapshot of the window is read once: the first is the base word,
  // the last the final word.  NB levels' loads are issued before any of them
  // is counted, so a vertex waits for ceil(levels / NB) round trips rather
  // than one per level (one at a time, the loop waited for each level's load
  // before issuing the next).  R-MAT s22: emission 4.20 -> 3.87 ms per plain
  // step (128 lanes, NB 8), batched step 9.87 -> 9.5 ms (256 lanes, NB 4);
  // the batches cost registers (256 lanes: 6 -> 4 waves per SIMD), which the
  // shorter load chains repay.
  constexpr int NB = W <= 2 ? 8 : (W == 4 ? 4 : 2);
#pragma unroll
  for (int w = 0; w < W; ++w) bs[w] = fin[w] = 0;
  for (int L0 = b; L0 <= e; L0 += NB) {
    u64 s[NB][W];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      if (L0 + k <= e) {
        load_word<W>(ring + (int64_t)((L0 + k) % R) * slot + i * W, s[k]);
      } else {
#pragma unroll
        for (int w = 0; w < W; ++w) s[k][w] = ~0ull;  // nothing absent
      }
    }
    if (L0 == b) {
#pragma unroll
      for (int w = 0; w < W; ++w) bs[w] = s[0][w];
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      if (L0 + k > e) break;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        u64 carry = ~s[k][w] & level_masks[(int64_t)(L0 + k - b) * W + w];
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          const u64 nc = c[j][w] & carry;
          c[j][w] ^= carry;
          carry = nc;
        }
        fin[w] = s[k][w];
      }
    }
  }
}

// Counters of a vertex without a status word (isolated, or past the end).
template <int W, int NS>
__device__ __forceinline__ void emit_zero(u64 (&c)[NS][W], u64 (&fin)[W], u64 (&bs)[W]) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    fin[w] = bs[w] = 0;
#pragma unroll
    for (int j = 0; j < NS; ++j) c[j][w] = 0;
  }
}

// Depth of lane k at vertex v, or SKIP when this window must not write it.
// ci < 0: v is isolated; its only depth is 0 at the lane's own source.
constexpr int32_t SKIP = (int32_t)0x80000000;

template <int W, int NS>
__device__ __forceinline__ int32_t lane_depth(const u64 (&c)[NS][W], const u64 (&fin)[W],
                                              const u64 (&bs)[W], int w, int bit, int off,
                                              bool iso, bool is_src) {
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < NS; ++j) cnt |= (int)((c[j][w] >> bit) & 1ull) << j;
  const bool set = iso ? is_src : (bool)((fin[w] >> bit) & 1ull);
  const bool at_base = iso ? is_src : (bool)((bs[w] >> bit) & 1ull);
  if (off <= 0) return set ? cnt : -1;
  return (set && !at_base) ? off + cnt : SKIP;
}

// Stores of one vertex's depths when every selected lane is in its first
// window and the window has fewer than 16 levels (4 counter bits): the lane
// loop is unrolled over 32-bit halves of the words, so each bit position is
// a constant; eight lane pointers are read from LDS at a time (null for a
// lane that is not emitted), eight depths computed as cnt | (set - 1) (-1
// where the bit never arrived), then stored.  An isolated vertex has no
// word (every depth -1 here); the lane whose source it is gets its 0 from
// k_emit_iso_sources after the pass.
template <int W>
__device__ __forceinline__ void emit_first4(const u64 (&c)[4][W], const u64 (&fin)[W],
                                            int32_t *const *s_out, int64_t v, bool nt) {
  // The counter words' loads complete here, before the lane loop: otherwise
  // lane blocks that branch on a null pointer wait for the vector memory
  // counter at their start, and that count includes earlier lanes' stores,
  // which would serialise the stores.
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt unchanged
#pragma unroll
  for (int w = 0; w < W; ++w) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t c0 = (uint32_t)(c[0][w] >> (32 * h)), c1 = (uint32_t)(c[1][w] >> (32 * h));
      const uint32_t c2 = (uint32_t)(c[2][w] >> (32 * h)), c3 = (uint32_t)(c[3][w] >> (32 * h));
      const uint32_t f = (uint32_t)(fin[w] >> (32 * h));
#pragma unroll
      for (int g = 0; g < 32; g += 8) {
        int32_t *o[8];
        int32_t d[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int bt = g + i;
          o[i] = s_out[w * 64 + h * 32 + bt];
          const uint32_t cnt = ((c0 >> bt) & 1u) | (((c1 >> bt) & 1u) << 1) |
                               (((c2 >> bt) & 1u) << 2) | (((c3 >> bt) & 1u) << 3);
          d[i] = (int32_t)(cnt | (((f >> bt) & 1u) - 1u));
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
          if (o[i]) put_depth(o[i], v, d[i], nt);
      }
    }
  }
}

template <int W, int NS>
__global__ __launch_bounds__(BS) void k_emit(const u64 *__restrict__ ring, int R, EmitSpace sp,
                                             int b, int e, const LaneOut *__restrict__ lanes,
                                             const u64 *__restrict__ level_masks,
                                             const u64 *__restrict__ sel, int k_lanes, int nt,
                                             u64 *edges_out) {
  __shared__ int32_t *s_out[MAX_LANES];
  __shared__ int32_t s_off[MAX_LANES];  // b - admit: <= 0 in the lane's first window
  __shared__ int32_t s_src[MAX_LANES];
  __shared__ u64 red[BS / 64];
  u64 acc = 0;  // edges of the selected lanes' reached vertices (final window)
  int cont = 0;  // some selected lane continues an earlier window
  for (int i = threadIdx.x; i < W * 64; i += BS) {  // null past the group
    const bool on = i < k_lanes && ((sel[i >> 6] >> (i & 63)) & 1ull);
    s_out[i] = on ? lanes[i].out : nullptr;
    s_off[i] = on ? b - lanes[i].admit : 0;
    s_src[i] = on ? lanes[i].src : -1;
    cont |= on && s_off[i] > 0;
  }
  const bool fast = NS == 4 && !__syncthreads_or(cont);
  for (int64_t v = (int64_t)blockIdx.x * BS + threadIdx.x; v < sp.V;
       v += (int64_t)gridDim.x * BS) {
    const int32_t ci = sp.cid[v];
    const bool iso = ci < 0;
    u64 c[NS][W], fin[W], bs[W];
    if (!iso) emit_count<W, NS>(ring, R, sp.Vt, b, e, ci, level_masks, c, fin, bs);
    else emit_zero<W, NS>(c, fin, bs);
    if (edges_out && !iso) {
      int pc = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) pc += __popcll(fin[w] & sel[w]);
      acc += (u64)pc * (u64)sp.deg[ci];
    }
    if constexpr (NS == 4) {
      if (fast) {
        emit_first4<W>(c, fin, s_out, v, nt != 0);
        continue;
      }
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      for (int bit = 0; bit < 64; ++bit) {
        const int k = w * 64 + bit;
        if (k >= k_lanes) break;
        int32_t *out = s_out[k];
        if (!out) continue;
        const int32_t val =
            lane_depth<W, NS>(c, fin, bs, w, bit, s_off[k], iso, s_src[k] == (int32_t)v);
        if (val != SKIP) put_depth(out, v, val, nt);
      }
    }
  }
  if (edges_out) {  // uniform: one atomic per workgroup
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      u64 s = 0;
      for (int i = 0; i < BS / 64; ++i) s += red[i];
      if (s) atomicAdd(edges_out, s);
    }
  }
}

// Depth 0 of a lane whose source is isolated (no status word), in the
// lane's first window: the fast path of k_emit writes -1 at every isolated
// vertex.  One thread per lane, after that pass on the same stream.
__global__ void k_emit_iso_sources(const LaneOut *__restrict__ lanes, const u64 *__restrict__ sel,
                                   const int32_t *__restrict__ cid, int b, int k_lanes) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= k_lanes || !((sel[i >> 6] >> (i & 63)) & 1ull)) return;
  const LaneOut l = lanes[i];
  if (l.out && l.src >= 0 && b - l.admit <= 0 && cid[l.src] < 0) l.out[l.src] = 0;
}

// Tiled depth emission: same values as k_emit, but a workgroup owns 256
// consecutive vertices and stages 32 lanes at a time through LDS, so that
// each wave writes one lane's 256 depths as a single 1-KB contiguous segment
// with 16-byte stores.  Entries a continuation window must not touch carry
// the sentinel INT_MIN and are skipped.  vec4 = every lane's output is
// 16-byte aligned and V % 4 == 0.
template <int W, int NS>
__global__ __launch_bounds__(BS) void k_emit_tiled(const u64 *__restrict__ ring, int R,
                                                   EmitSpace sp, int b, int e,
                                                   const LaneOut *__restrict__ lanes,
                                                   const u64 *__restrict__ level_masks,
                                                   const u64 *__restrict__ sel, int k_lanes,
                                                   int vec4) {
  __shared__ int4 tile4[32 * 64];  // [row][64 x int4] = 32 lanes x 256 vertices
  __shared__ int32_t *s_out[MAX_LANES];
  __shared__ int32_t s_off[MAX_LANES];
  __shared__ int32_t s_src[MAX_LANES];
  int32_t *tile = reinterpret_cast<int32_t *>(tile4);
  const int t = threadIdx.x;
  const int64_t V = sp.V;
  for (int i = t; i < k_lanes; i += BS) {
    const bool on = (sel[i >> 6] >> (i & 63)) & 1ull;
    s_out[i] = on ? lanes[i].out : nullptr;
    s_off[i] = b - lanes[i].admit;
    s_src[i] = lanes[i].src;
  }
  __syncthreads();
  const int64_t nblk = (V + BS - 1) / BS;
  for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int64_t v = blk * BS + t;
    const int32_t ci = (v < V) ? sp.cid[v] : -1;
    const bool iso = ci < 0;
    u64 c[NS][W], fin[W], bs[W];
    if (!iso) emit_count<W, NS>(ring, R, sp.Vt, b, e, ci, level_masks, c, fin, bs);
    else emit_zero<W, NS>(c, fin, bs);
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
          if (k < k_lanes && s_out[k] && v < V)
            val = lane_depth<W, NS>(c, fin, bs, w, bit, s_off[k], iso, s_src[k] == (int32_t)v);
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

// Wide-store depth emission: same values as k_emit.  A wave owns one status
// word w (64 lanes) of 256 consecutive vertices, four per thread, so its
// per-vertex state is NS+2 words of those four vertices and every store
// instruction writes 16 bytes per thread: one lane's 256 depths as a single
// 1-KB contiguous piece of its row (k_emit writes 256-byte pieces, one
// vertex per thread).  The waves of a workgroup take consecutive (chunk,
// word) units, so the words of one vertex are read by neighbouring waves
// from the same lines.  A lane row that is not 16-byte aligned, a chunk
// that ends past V, and a piece holding an entry a continuation window must
// not touch fall back to 4-byte stores.
template <int NS>
__device__ __forceinline__ int32_t lane_depth4(const u64 (&c)[NS][4], const u64 (&fin)[4],
                                               const u64 (&bs)[4], int i, int bit, int off,
                                               bool iso, bool is_src) {
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < NS; ++j) cnt |= (int)((c[j][i] >> bit) & 1ull) << j;
  const bool set = iso ? is_src : (bool)((fin[i] >> bit) & 1ull);
  const bool at_base = iso ? is_src : (bool)((bs[i] >> bit) & 1ull);
  if (off <= 0) return set ? cnt : -1;
  return (set && !at_base) ? off + cnt : SKIP;
}

template <int W, int NS>
__global__ __launch_bounds__(BS) void k_emit_wide(const u64 *__restrict__ ring, int R,
                                                  EmitSpace sp, int b, int e,
                                                  const LaneOut *__restrict__ lanes,
                                                  const u64 *__restrict__ level_masks,
                                                  const u64 *__restrict__ sel, int k_lanes,
                                                  u64 *edges_out) {
  __shared__ int32_t *s_out[MAX_LANES];
  __shared__ int32_t s_off[MAX_LANES];
  __shared__ int32_t s_src[MAX_LANES];
  __shared__ u64 red[BS / 64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int i = t; i < k_lanes; i += BS) {
    const bool on = (sel[i >> 6] >> (i & 63)) & 1ull;
    s_out[i] = on ? lanes[i].out : nullptr;
    s_off[i] = b - lanes[i].admit;
    s_src[i] = lanes[i].src;
  }
  __syncthreads();
  const int64_t V = sp.V, slot = sp.Vt * W;
  const bool v4 = (V & 3) == 0;
  const int64_t units = ((V + 255) >> 8) * W;
  u64 acc = 0;
  for (int64_t u = (int64_t)blockIdx.x * (BS / 64) + wv; u < units;
       u += (int64_t)gridDim.x * (BS / 64)) {
    const int64_t chunk = u / W;
    const int w = (int)(u % W);
    const int64_t v0 = chunk * 256 + 4 * lane;
    int32_t ci[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) ci[i] = (v0 + i < V) ? sp.cid[v0 + i] : -1;
    u64 c[NS][4], fin[4], bs[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      fin[i] = bs[i] = 0;
#pragma unroll
      for (int j = 0; j < NS; ++j) c[j][i] = 0;
    }
    const u64 *Sw = ring + w;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (ci[i] >= 0) {
        fin[i] = Sw[(int64_t)(e % R) * slot + (int64_t)ci[i] * W];
        bs[i] = Sw[(int64_t)(b % R) * slot + (int64_t)ci[i] * W];
      }
    }
    for (int L = b; L <= e; ++L) {
      const u64 lm = level_masks[(int64_t)(L - b) * W + w];
      const u64 *S = Sw + (int64_t)(L % R) * slot;
      u64 x[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = ci[i] >= 0 ? S[(int64_t)ci[i] * W] : ~0ull;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        u64 carry = ~x[i] & lm;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          const u64 nc = c[j][i] & carry;
          c[j][i] ^= carry;
          carry = nc;
        }
      }
    }
    if (edges_out) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (ci[i] >= 0) acc += (u64)__popcll(fin[i] & sel[w]) * (u64)sp.deg[ci[i]];
    }
    for (int bit = 0; bit < 64; ++bit) {
      const int k = w * 64 + bit;
      if (k >= k_lanes) break;
      int32_t *out = s_out[k];
      if (!out) continue;
      const int off = s_off[k];
      const int32_t src = s_src[k];
      int32_t d[4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        d[i] = lane_depth4<NS>(c, fin, bs, i, bit, off, ci[i] < 0, src == (int32_t)(v0 + i));
      gint32 *o = (gint32 *)out;
      const bool full = v4 && ((reinterpret_cast<uintptr_t>(out) & 15) == 0) && v0 + 3 < V &&
                        d[0] != SKIP && d[1] != SKIP && d[2] != SKIP && d[3] != SKIP;
      if (full) {
        typedef int v4i __attribute__((ext_vector_type(4)));
        v4i q = {d[0], d[1], d[2], d[3]};
        *reinterpret_cast<__attribute__((address_space(1))) v4i *>(o + v0) = q;
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (v0 + i < V && d[i] != SKIP) o[v0 + i] = d[i];
      }
    }
  }
  if (edges_out) {  // uniform: one atomic per workgroup
    acc = wave_sum(acc);
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (t == 0) {
      u64 s = 0;
      for (int i = 0; i < BS / 64; ++i) s += red[i];
      if (s) atomicAdd(edges_out, s);
    }
  }
}

// Closeness reduction over the snapshot window [b, e]: per lane k, the sum
// over vertices of depth_k(v) and the number of vertices reached, for the
// vertices whose bit arrived in the window, without writing depth arrays.
// With the per-vertex bit-sliced counters c_j (depth = sum_j 2^j c_j), the
// sum over a wave's 64 vertices is sum_j 2^j popcount(ballot(c_j bit k)),
// so each lane costs NS+1 ballots per 64 vertices.  Lane k of the window is
// a first window (off_k <= 0: every set bit counts, depth = count) or a
// continuation (off_k = b - admit > 0: bits set at e but not at b count,
// depth = off_k + count).  Wave partials meet in LDS, one atomic per lane
// and workgroup.
template <int W, int NS>
__global__ __launch_bounds__(BS) void k_close(const u64 *__restrict__ ring, int R, EmitSpace sp,
                                              int b, int e, const LaneOut *__restrict__ lanes,
                                              const u64 *__restrict__ level_masks,
                                              const u64 *__restrict__ sel, int k_lanes,
                                              unsigned long long *__restrict__ sum_out,
                                              unsigned long long *__restrict__ reach_out) {
  __shared__ int32_t s_off[MAX_LANES];
  __shared__ u64 red[2][4][MAX_LANES];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int i = t; i < MAX_LANES; i += BS) s_off[i] = (i < k_lanes) ? b - lanes[i].admit : 0;
  __syncthreads();
  u64 selw[W], cont[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    selw[w] = sel[w];
    cont[w] = 0;
    for (int bit = 0; bit < 64 && w * 64 + bit < k_lanes; ++bit)
      if (s_off[w * 64 + bit] > 0) cont[w] |= 1ull << bit;
  }
  u64 acc_sum[W], acc_reach[W];
#pragma unroll
  for (int w = 0; w < W; ++w) acc_sum[w] = acc_reach[w] = 0;
  // reduced over traversal ids: an isolated vertex contributes nothing
  const int64_t nblk = (sp.Vt + BS - 1) / BS;
  for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int64_t v = blk * BS + t;
    u64 c[NS][W], fin[W], bs[W], M[W];
    if (v < sp.Vt) emit_count<W, NS>(ring, R, sp.Vt, b, e, v, level_masks, c, fin, bs);
    else emit_zero<W, NS>(c, fin, bs);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      M[w] = fin[w] & ~(bs[w] & cont[w]) & selw[w];
#pragma unroll
      for (int j = 0; j < NS; ++j) c[j][w] &= M[w];
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (!wave_or(M[w])) continue;  // uniform: no counted bit in this word
      for (int bit = 0; bit < 64; ++bit) {
        const int k = w * 64 + bit;
        if (k >= k_lanes) break;
        const u64 pm = __popcll(__ballot((M[w] >> bit) & 1ull));
        if (!pm) continue;
        u64 s = 0;
#pragma unroll
        for (int j = 0; j < NS; ++j) s += (u64)__popcll(__ballot((c[j][w] >> bit) & 1ull)) << j;
        if (lane == bit) {
          acc_reach[w] += pm;
          acc_sum[w] += s + (u64)(s_off[k] > 0 ? s_off[k] : 0) * pm;
        }
      }
    }
  }
#pragma unroll
  for (int w = 0; w < W; ++w) {
    red[0][wv][w * 64 + lane] = acc_sum[w];
    red[1][wv][w * 64 + lane] = acc_reach[w];
  }
  __syncthreads();
  for (int i = t; i < k_lanes; i += BS) {
    const u64 s = red[0][0][i] + red[0][1][i] + red[0][2][i] + red[0][3][i];
    const u64 r = red[1][0][i] + red[1][1][i] + red[1][2][i] + red[1][3][i];
    if (r) {
      atomicAdd(&sum_out[i], s);
      atomicAdd(&reach_out[i], r);
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

// ---------------------------------------------------------------------------
// GroupBy.  Key of source i: its first neighbour (adjacency order) of degree
// > 128 if deg(src[i]) <= 128, else -1 (groupby.cu:49-66); one wave per
// source scans 64 neighbours per step and takes the lowest set ballot lane.
// Two-hop size of source i: the sum of its neighbours' degrees (0 for an
// isolated source), one wave per source.
__global__ __launch_bounds__(BS) void k_two_hop(const int64_t *__restrict__ beg,
                                                const int32_t *__restrict__ adj,
                                                const int32_t *__restrict__ deg,
                                                const int32_t *__restrict__ cid,
                                                const int32_t *__restrict__ src, int64_t n,
                                                u64 *__restrict__ key) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int32_t s = cid[src[i]];
  u64 sum = 0;
  if (s >= 0)
    for (int64_t j = beg[s] + lane; j < beg[s + 1]; j += 64) sum += (u64)deg[adj[j]];
  sum = wave_sum(sum);
  if (lane == 0) key[i] = sum;
}

__global__ __launch_bounds__(BS) void k_groupby_key(const int64_t *__restrict__ beg,
                                                    const int32_t *__restrict__ adj,
                                                    const int32_t *__restrict__ deg,
                                                    const int32_t *__restrict__ cid,
                                                    const int32_t *__restrict__ src, int64_t n,
                                                    int32_t *__restrict__ key) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int32_t s = cid[src[i]];  // < 0: isolated, no neighbour to key on
  int32_t h = -1;
  if (s >= 0 && deg[s] <= 128) {
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

// Sort key of the placement (groupby.cu:69-115): an eligible source sorts by
// its hub id, an ineligible one after every eligible one; a stable sort then
// keeps input order among equal keys, which is the reference's placement
// (eligible sources grouped by hub in input order, then the others in input
// order).
__global__ void k_groupby_sortkey(const int32_t *__restrict__ key, int64_t n,
                                  uint32_t *__restrict__ ukey) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n; i += (int64_t)gridDim.x * BS)
    ukey[i] = key[i] >= 0 ? (uint32_t)key[i] : 0xFFFFFFFFu;
}

}  // namespace

// One vertex per thread up to 16M vertices: fewer, longer-lived threads
// leave the write stream short of stores in flight.
#ifndef IBFS_EMIT_GRID_CAP
#define IBFS_EMIT_GRID_CAP 65536
#endif

template <int W, int NS>
static void emit_ns(int mode, int vec4, const u64 *ring, int R, EmitSpace sp, int b, int e,
                    const LaneOut *lanes, const u64 *masks, const u64 *sel, int k, u64 *edges_out,
                    hipStream_t st) {
  if (mode == 1) {
    dim3 g(grid_for(sp.V, BS, 4096)), blk(BS);
    hipLaunchKernelGGL((k_emit_tiled<W, NS>), g, blk, 0, st, ring, R, sp, b, e, lanes, masks, sel,
                       k, vec4);
    if (edges_out) {
      dim3 gc(grid_for(sp.Vt, BS * 4, 4096));
      hipLaunchKernelGGL(k_count_edges<W>, gc, blk, 0, st, ring + (int64_t)(e % R) * sp.Vt * W,
                         sp.deg, sp.Vt, sel, edges_out);
    }
  } else if (mode == 3) {
    const int64_t units = ((sp.V + 255) >> 8) * W;
    dim3 g(grid_for(units, BS / 64, IBFS_EMIT_GRID_CAP)), blk(BS);
    hipLaunchKernelGGL((k_emit_wide<W, NS>), g, blk, 0, st, ring, R, sp, b, e, lanes, masks, sel,
                       k, edges_out);
  } else {
    dim3 g(grid_for(sp.V, BS, IBFS_EMIT_GRID_CAP)), blk(BS);
    hipLaunchKernelGGL((k_emit<W, NS>), g, blk, 0, st, ring, R, sp, b, e, lanes, masks, sel, k,
                       mode == 2 ? 1 : 0, edges_out);
    if (NS == 4)  // depth 0 at isolated sources after the fast path's -1
      hipLaunchKernelGGL(k_emit_iso_sources, dim3((k + BS - 1) / BS), dim3(BS), 0, st, lanes, sel,
                         sp.cid, b, k);
  }
}

template <int W>
static hipError_t emit_w(int mode, int vec4, const u64 *ring, int R, EmitSpace sp, int b, int e,
                         const LaneOut *lanes, const u64 *masks, const u64 *sel, int k,
                         u64 *edges_out, hipStream_t st) {
  const int span = e - b + 1;
  if (span < 16) emit_ns<W, 4>(mode, vec4, ring, R, sp, b, e, lanes, masks, sel, k, edges_out, st);
  else if (span < 64) emit_ns<W, 6>(mode, vec4, ring, R, sp, b, e, lanes, masks, sel, k, edges_out, st);
  else if (span < 256) emit_ns<W, 8>(mode, vec4, ring, R, sp, b, e, lanes, masks, sel, k, edges_out, st);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_emit(int W, int mode, int vec4, const u64 *ring, int R, EmitSpace sp, int b,
                       int e, const LaneOut *lanes, const u64 *level_masks, const u64 *sel,
                       int k_lanes, u64 *edges_out, hipStream_t st) {
  if (sp.V <= 0 || e < b) return hipSuccess;
  if (edges_out && !sp.deg) return hipErrorInvalidValue;
  switch (W) {
    case 1: return emit_w<1>(mode, vec4, ring, R, sp, b, e, lanes, level_masks, sel, k_lanes, edges_out, st);
    case 2: return emit_w<2>(mode, vec4, ring, R, sp, b, e, lanes, level_masks, sel, k_lanes, edges_out, st);
    case 4: return emit_w<4>(mode, vec4, ring, R, sp, b, e, lanes, level_masks, sel, k_lanes, edges_out, st);
    case 8: return emit_w<8>(mode, vec4, ring, R, sp, b, e, lanes, level_masks, sel, k_lanes, edges_out, st);
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
    case 8: hipLaunchKernelGGL(k_count_edges<8>, g, blk, 0, st, S, deg, V, lanes, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int W>
static hipError_t close_w(const u64 *ring, int R, EmitSpace sp, int b, int e, const LaneOut *lanes,
                          const u64 *masks, const u64 *sel, int k, u64 *sum, u64 *reach,
                          hipStream_t st) {
  const int span = e - b + 1;
  dim3 g(grid_for(sp.Vt, BS * 8, 256)), blk(BS);
  if (span < 16)
    hipLaunchKernelGGL((k_close<W, 4>), g, blk, 0, st, ring, R, sp, b, e, lanes, masks, sel, k, sum, reach);
  else if (span < 64)
    hipLaunchKernelGGL((k_close<W, 6>), g, blk, 0, st, ring, R, sp, b, e, lanes, masks, sel, k, sum, reach);
  else if (span < 256)
    hipLaunchKernelGGL((k_close<W, 8>), g, blk, 0, st, ring, R, sp, b, e, lanes, masks, sel, k, sum, reach);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_close(int W, const u64 *ring, int R, EmitSpace sp, int b, int e,
                        const LaneOut *lanes, const u64 *level_masks, const u64 *sel, int k_lanes,
                        u64 *sum_out, u64 *reach_out, hipStream_t st) {
  if (sp.Vt <= 0 || e < b) return hipSuccess;
  switch (W) {
    case 1: return close_w<1>(ring, R, sp, b, e, lanes, level_masks, sel, k_lanes, sum_out, reach_out, st);
    case 2: return close_w<2>(ring, R, sp, b, e, lanes, level_masks, sel, k_lanes, sum_out, reach_out, st);
    case 4: return close_w<4>(ring, R, sp, b, e, lanes, level_masks, sel, k_lanes, sum_out, reach_out, st);
    case 8: return close_w<8>(ring, R, sp, b, e, lanes, level_masks, sel, k_lanes, sum_out, reach_out, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_two_hop(const int64_t *beg, const int32_t *adj, const int32_t *deg,
                          const int32_t *cid, const int32_t *src, int64_t n, u64 *key,
                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_two_hop, dim3(grid_for(n, 4, 1 << 30)), dim3(BS), 0, st, beg, adj, deg, cid,
                     src, n, key);
  return hipGetLastError();
}

// Workspace of launch_groupby: the sort keys in and out, then the radix
// sort's temporary storage (256-byte aligned).
static size_t groupby_sort_bytes(int64_t n) {
  size_t tb = 0;
  (void)rocprim::radix_sort_pairs(nullptr, tb, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                  (const int32_t *)nullptr, (int32_t *)nullptr, (size_t)n, 0, 32);
  return tb;
}

size_t groupby_workspace_bytes(int64_t n) {
  if (n <= 0) return 0;
  const size_t keys = ((size_t)n * 2 * sizeof(uint32_t) + 255) & ~(size_t)255;
  return keys + groupby_sort_bytes(n);
}

hipError_t launch_groupby(const int64_t *beg, const int32_t *adj, const int32_t *deg,
                          const int32_t *cid, const int32_t *src, int64_t n, int32_t *key,
                          int32_t *out, void *ws, size_t ws_bytes, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (!ws || ws_bytes < groupby_workspace_bytes(n)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_groupby_key, dim3(grid_for(n, 4, 1 << 30)), dim3(BS), 0, st, beg, adj, deg,
                     cid, src, n, key);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  uint32_t *ukey = static_cast<uint32_t *>(ws);
  const size_t keys = ((size_t)n * 2 * sizeof(uint32_t) + 255) & ~(size_t)255;
  hipLaunchKernelGGL(k_groupby_sortkey, dim3(grid_for(n, BS * 4, 4096)), dim3(BS), 0, st, key, n,
                     ukey);
  err = hipGetLastError();
  if (err != hipSuccess) return err;
  // same iterator types as the size query in groupby_sort_bytes
  size_t tb = groupby_sort_bytes(n);
  return rocprim::radix_sort_pairs(static_cast<char *>(ws) + keys, tb,
                                   static_cast<const uint32_t *>(ukey), ukey + n, src, out,
                                   (size_t)n, 0, 32, st);
}

}  // namespace ibfs
