// Device helpers shared by the gfx950 kernel translation units (wave64).
//
// Status words: W = lane_bits/64 u64 words per vertex, vertex-major
// (S[v*W + w]); lane k is bit k%64 of word k/64, the same memory image as the
// reference's uint4 with bit k%32 of component k/32 (util.h:117-130).
// Level L's words live in ring slot L % R of R snapshots (V*W words each).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ibfs {
namespace dev {

constexpr int BS = 256;  // threads per workgroup (4 waves)

typedef __attribute__((address_space(1))) int32_t gint32;

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

template <int W>
__device__ __forceinline__ bool differs(const u64 (&x)[W], const u64 (&y)[W]) {
  bool d = false;
#pragma unroll
  for (int w = 0; w < W; ++w) d |= (x[w] != y[w]);
  return d;
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

// Sources of a bottom-up gather at level L+1: S_L of non-hub neighbours and
// the hub cache (the hubs' words of S_L).
struct Gather {
  const u64 *__restrict__ prev;
  const u64 *__restrict__ hs;
};

// Word of traversal-adjacency entry a: a >= 0 is a vertex id (read from the
// snapshot), a < 0 is hub slot a & 0x7fffffff (read from the hub cache).
template <int W>
__device__ __forceinline__ void load_nbr(const Gather &gs, int32_t a, u64 (&y)[W]) {
  const u64 *p = (a < 0) ? gs.hs + (int64_t)(a & 0x7fffffff) * W : gs.prev + (int64_t)a * W;
  load_word<W>(p, y);
}

__device__ __forceinline__ bool is_hub(const HubSplit &h, int64_t v, int d) {
  return d > h.deg || (d == h.deg && v <= (int64_t)h.tie);
}

// Coherent read of a counter another workgroup updated with device atomics.
__device__ __forceinline__ u64 atomic_read(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline int grid_for(int64_t n, int per_block, int cap) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace dev
}  // namespace ibfs
