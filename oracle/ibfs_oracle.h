/*
 * ibfs_oracle.h - CPU oracle for the concurrent multi-source BFS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the reported CPU baseline.  The product (ibfs_amd) never calls
 * it and fails loudly when its HIP library is missing.
 *
 * Every function restates one piece of the iBFS reference (file:line cited at
 * the definition).  Each function is pinned by tests/test_oracle.py against
 * golden vectors written by a library routine (scipy shortest paths), brute
 * force (Floyd-Warshall), closed forms, hand-derived known answers and the
 * reference's own converter compiled from its source (oracle/Makefile).
 */
#ifndef IBFS_ORACLE_H
#define IBFS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Textbook queue BFS from one source: depth[v] = hop distance, -1 if
 * unreached.  The reference has no CPU BFS; depth semantics are those of its
 * status-word recurrence (first level at which a source's bit reaches v,
 * expander.cu:100-359 + inspector.cu:59-85). */
int oracle_bfs(int64_t V, const int64_t *beg, const int32_t *adj, int32_t src,
               int32_t *depth);

/* oracle_bfs for n sources, depth[i*V + v]; nthreads pthreads over sources. */
int oracle_bfs_many(int64_t V, const int64_t *beg, const int32_t *adj,
                    const int32_t *src, int64_t n, int32_t *depth,
                    int nthreads);

/* Level-synchronous bitwise multi-source BFS following the reference's
 * algorithm: one status word of k bits per vertex (util.h:80-83), lane i =
 * bit i (expander.cu:69), top-down push for level <= sw_level
 * (expander.cu:100-148), bottom-up pull with early termination afterwards
 * (expander.cu:320-359), termination when no status word changes
 * (inspector.cu:59-85).  k <= 256.  depth[i*V+v] = first level at which bit i
 * is present in v's word, -1 if never.  Returns the number of levels. */
int oracle_ibfs_bitwise(int64_t V, const int64_t *beg, const int32_t *adj,
                        const int32_t *src, int k, int sw_level,
                        int32_t *depth);

/* GroupBy (groupby.cu:36-149): out = permutation of src. */
int oracle_groupby(int64_t V, const int64_t *beg, const int32_t *adj,
                   const int32_t *src, int64_t n, int32_t *out);

/* Degree class (inspector.cu:130-134): 0 none (deg 0), 1 sml (< sml),
 * 3 lrg (> lrg), 2 mid otherwise. */
int oracle_degree_class(int64_t deg, int64_t sml, int64_t lrg);

/* TEPS numerator (validate.h:59-76): sum of deg(v) over v with depth >= 0. */
int64_t oracle_traversed_edges(int64_t V, const int64_t *beg,
                               const int32_t *depth);

/* Depth-array consistency (the intent of validate.h:5-56, without its
 * dangling-else defect): returns 0 if depth[src]==0, every edge joins depths
 * that differ by at most one (or two unreached ends), every reached v != src
 * has a neighbour at depth-1, and no reached vertex neighbours an unreached
 * one.  Nonzero codes name the violated rule. */
int oracle_check_depth(int64_t V, const int64_t *beg, const int32_t *adj,
                       int32_t src, const int32_t *depth);

/* Reference CSR reader (graph.cu:63-119): V = size(beg)/8 - 1,
 * E = size(csr)/8, narrowed to 64-bit offsets and 32-bit ids.  Caller frees
 * *beg and *adj with oracle_free.  Returns 0, or -2 (io), -3 (format). */
int oracle_read_csr(const char *beg_path, const char *csr_path, int64_t *V,
                    int64_t *E, int64_t **beg, int32_t **adj);

/* Text edge list -> CSR with the converter's rules
 * (tuple_text_to_bin.cpp:56-272).  Caller frees with oracle_free. */
int oracle_text_to_csr(const char *path, int reverse, int64_t skip_lines,
                       int64_t *V, int64_t *E, int64_t **beg, int32_t **adj);

void oracle_free(void *p);

/* Per-level work model of the engine's algorithm (bottom-up candidates and
 * early-termination prefixes with respect to the level's active lanes).
 * dirs[L] gives the direction of level L (1..levels-1): 0 TD, 1 BU.
 * Outputs per level L (arrays of length `levels`):
 *   n_changed[L], n_front[L], m_front[L]   vertices/edges of the frontier
 *   n_cand[L], m_cand[L]                   bottom-up candidates of level L+1
 *   bu_prefix[L]  sum over candidates of the sequential adjacency prefix
 *                 length read by level L's bottom-up pass (k_v)
 *   lam_pop[L]    lanes active at level L.
 * Returns the number of levels or a negative code. */
int oracle_level_model(int64_t V, const int64_t *beg, const int32_t *adj,
                       const int32_t *src, int k, const int8_t *dirs,
                       int max_levels, int64_t *n_changed, int64_t *n_front,
                       int64_t *m_front, int64_t *n_cand, int64_t *m_cand,
                       int64_t *bu_prefix, int32_t *lam_pop);

#ifdef __cplusplus
}
#endif
#endif
