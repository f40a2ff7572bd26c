/*
 * ibfs.h - C ABI of the MI355X concurrent multi-source BFS engine.
 *
 * Every entry point returns an int status (IBFS_OK == 0) and never exits the
 * process; the message of the last failure on the calling thread is returned
 * by ibfs_last_error().  Handles are opaque.  A graph handle is read-only once
 * built and may be shared by several engines; an engine belongs to one host
 * thread at a time and owns its HIP stream and device buffers.
 *
 * Reference interfaces each entry point replaces are cited as file:line of the
 * iBFS sources (Gen B unless noted).
 */
#ifndef IBFS_H
#define IBFS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IBFS_ABI_VERSION 1

/* status codes */
#define IBFS_OK 0
#define IBFS_ERR_ARG -1      /* invalid argument (null, range, size)            */
#define IBFS_ERR_IO -2       /* file missing, unreadable, truncated              */
#define IBFS_ERR_FORMAT -3   /* file contents inconsistent (sizes, id range)     */
#define IBFS_ERR_HIP -4      /* HIP runtime error                                */
#define IBFS_ERR_NOMEM -5    /* device or host allocation failed / over budget   */
#define IBFS_ERR_STATE -6    /* call not valid in the object's current state     */
#define IBFS_ERR_NODEVICE -7 /* no HIP device available                          */

/* direction policies (reference: fixed switch level, expander.cu:503-515) */
#define IBFS_POLICY_FIXED 0      /* top-down while level <= sw_level, then bottom-up */
#define IBFS_POLICY_ALPHA_BETA 1 /* Beamer-style switch on frontier/unvisited edges  */
#define IBFS_POLICY_TD_ONLY 2
#define IBFS_POLICY_BU_ONLY 3
#define IBFS_POLICY_HYBRID 4     /* every level hybrid: candidates pull through their
                                    pulled-hub neighbours, the rest of the frontier
                                    pushes */

/* output placement for depth arrays */
#define IBFS_OUT_HOST 0
#define IBFS_OUT_DEVICE 1

typedef struct ibfs_graph_s *ibfs_graph_t;
typedef struct ibfs_engine_s *ibfs_engine_t;
typedef struct ibfs_hostcsr_s *ibfs_hostcsr_t;
typedef struct ibfs_server_s *ibfs_server_t;

/* Engine configuration.  Defaults (ibfs_config_default): 128 lanes
 * (concurrent-count, readme.md:27), lrg_shed=512 as the reference
 * (main.cu:56-57), sml_shed=64 where the reference has 32 (main.cu:56): the
 * thread/wave class boundary follows the 64-wide wavefront. */
typedef struct ibfs_config {
  int32_t lane_bits;   /* 64, 128, 256 or 512 source lanes per status word     */
  int32_t sml_shed;    /* degree < sml_shed        -> thread class             */
  int32_t lrg_shed;    /* degree > lrg_shed        -> workgroup class          */
  int32_t policy;      /* IBFS_POLICY_*                                        */
  int32_t sw_level;    /* last top-down level for IBFS_POLICY_FIXED (argv[4])   */
  int32_t alpha;       /* TD->BU when m_f * alpha > m_u   (ALPHA_BETA)          */
  int32_t beta;        /* BU->TD when n_f * beta < V      (ALPHA_BETA)          */
  int32_t wg_size;     /* threads per workgroup of the inspector: 256/512/1024 */
  int32_t td_chunk;    /* edges per workgroup chunk of a top-down hub          */
  int32_t ring_levels; /* status snapshots kept on the device (>= 4)           */
  int32_t kernel_timing; /* 1: time every kernel with HIP events             */
  int32_t device;      /* HIP device ordinal                                   */
  int32_t work_model;  /* 1: measure each bottom-up level's early-termination
                          prefix into the level log (instrumentation)        */
  int32_t emit_mode;   /* depth emission: 0 one vertex per thread, 4-byte
                          stores; 1 LDS-tiled; 2 as 0 with non-temporal
                          stores; 3 one status word of 256 vertices per wave,
                          1-KB pieces of each lane's row per store           */
  int32_t dd_grid;     /* workgroups of the device-driven expander (0: auto)  */
  int32_t inspect_vpt; /* inspector vertices per lane and tile: 1, 2, 4, 8 or
                          16, capped at 16 words per lane (0: by size)       */
  int32_t hybrid_gamma; /* ALPHA_BETA: a top-down level runs hybrid when
                          gamma/64 * (candidates' entries to the pulled hubs) +
                          (edges of the other frontier vertices) < (frontier
                          edges); 0: never                                    */
  int32_t hybrid_hubs; /* hubs a hybrid level pulls through: the top
                          hybrid_hubs by degree (0: every cached hub)         */
  int32_t pack_sources; /* batched runs: 1 = groups of sources with similar
                          two-hop sizes (their lanes progress alike), 0 =
                          groups in the caller's order; outputs stay in the
                          caller's order either way                          */
  int32_t split_cov;    /* batched runs and serving (device-driven levels): a
                          bottom-up level lets the active lanes that hold
                          fewer than split_cov % of the top hybrid_hubs hubs
                          push top-down instead (their sparse frontiers would
                          keep every candidate scanning); 0: never           */
  int32_t phase_offsets; /* batched runs: p in 1..3 = a lane whose source's
                          two-hop size is below its group's median by a
                          factor of d^(p/2) per level (d = average degree;
                          rounded, at most 2 levels) starts that many levels
                          earlier than the median lanes, so that the group's
                          lanes reach their dense levels together (depths
                          stay relative to each lane's own start); 0 = every
                          lane starts at level 0                              */
  int32_t tail_div;     /* batched runs (device-driven levels): once the
                          shrinking top-down frontier of a level >= 2 has at
                          most V/tail_div vertices and E/tail_div edges and at
                          most V/tail_div vertices still miss an active lane,
                          the remaining levels push only newly gained bits into
                          one cumulative status slot and log each gain, with
                          no inspection pass; the depths of those levels are
                          written from the log after the emission.  A log
                          that overflows (more records than traversal
                          vertices + 65,536) reruns the group without it.
                          0: never; default 32                                */
} ibfs_config;

/* Per-level record of the last traversal (the build's monitoring log, in place
 * of ENABLE_MONITORING prints at traversal.cu:127-128). */
typedef struct ibfs_level_info {
  int32_t level;       /* level L: the pass that produced status_L             */
  int32_t direction;   /* 0 top-down, 1 bottom-up, 2 hybrid, 3 split, -1 init */
  int64_t n_changed;   /* vertices whose status gained bits at L               */
  int64_t n_frontier;  /* changed vertices with degree > 0                     */
  int64_t m_frontier;  /* sum of degrees of the frontier                       */
  int64_t n_unvisited; /* bottom-up candidates after L                         */
  int64_t m_unvisited; /* sum of degrees of the candidates                     */
  int64_t q_td[3];     /* top-down queue sizes (sml, mid, lrg chunks)          */
  int64_t q_bu[3];     /* bottom-up queue sizes (sml, mid, lrg)                */
  int64_t bu_work;     /* bottom-up level: candidates missing an active lane   */
  int64_t bu_prefix;   /* ... sum of their early-termination prefixes (edges)  */
  float expand_ms;     /* kernel time of the expansion producing status_L      */
  float inspect_ms;    /* kernel time of the inspection of status_L            */
  int64_t m_frontier_nonhub; /* ... of the frontier vertices that are not hubs */
} ibfs_level_info;

typedef struct ibfs_stats {
  int32_t levels;          /* depth of the deepest source + 1                 */
  int32_t emit_windows;    /* depth-emission passes (1 unless the ring wrapped) */
  int64_t traversed_edges; /* sum over lanes of deg(v) for reached v            */
  double total_ms;         /* device time from first to last kernel (events)   */
  double init_ms, expand_td_ms, expand_bu_ms, inspect_ms, emit_ms; /* timing=1 */
  int32_t n_td_launches, n_bu_launches, n_inspect_launches, n_emit_launches;
  double expand_hy_ms;     /* hybrid expansions (timing=1)                     */
  int32_t n_hy_launches;
  int32_t n_split_levels;  /* batched runs: split levels (dense lanes pull,
                              sparse lanes push; cfg.split_cov)               */
  int32_t n_tail_levels;   /* batched runs: sparse tail expansions, the last
                              one (which found nothing) included (tail_div) */
  int32_t n_tail_reruns;   /* batched runs: groups rerun without the sparse
                              tail after its log overflowed                   */
} ibfs_stats;

/* ---- errors / version ---------------------------------------------------- */
int ibfs_abi_version(void);
/* Copies the last error message of the calling thread into buf.
 * Replaces the print-and-exit of H_ERR (util.h:43-54, comm.h:5-16). */
int ibfs_last_error(char *buf, size_t n);
int ibfs_config_default(ibfs_config *cfg);
int ibfs_device_count(int *n);
/* Limit, in seconds, of every host wait of the library on device work
 * (process-wide; default 120).  A wait that exceeds it fails with
 * IBFS_ERR_STATE and a message naming the wait, instead of spinning on a
 * stream that never drains.  An engine or batch whose call failed this way
 * may be called again: that call first waits, within the limit, for the
 * failed call's work to end (IBFS_ERR_STATE again if it has not).  Servers
 * should be destroyed.  The reference's host loop spins on a mapped flag
 * with no limit (traversal.cu:101-130). */
int ibfs_set_wait_timeout(double seconds);
int ibfs_get_wait_timeout(double *seconds);

/* ---- device memory helpers (callers without a GPU framework) ------------- */
int ibfs_device_alloc(int device, size_t bytes, void **dptr);
int ibfs_device_free(void *dptr);
/* Device-to-host copy into pageable memory (synchronous, on a library stream
 * of the source's device).  Copies of 4 MiB or more stream through two
 * 64-MiB pinned chunks while eight host threads copy into dst (the library
 * keeps them for the process; such copies are serialised across threads).
 * Host-placed depth arrays (IBFS_OUT_HOST) take the same path. */
int ibfs_memcpy_d2h(void *dst, const void *src, size_t bytes);
int ibfs_memcpy_h2d(void *dst, const void *src, size_t bytes);
/* Device-to-device copy (synchronous): moving depth arrays between device
 * buffers, and the plain copy of known size that scales HBM counters. */
int ibfs_memcpy_d2d(void *dst, const void *src, size_t bytes);

/* ---- host CSR: reference file format, generators, converter -------------- */
/* Reads <beg_pos.bin> (V+1 little-endian 8-byte integers) and <csr.bin>
 * (E 8-byte integers), V and E derived from the file sizes and narrowed to
 * int32 ids with a range check.  Replaces graph::graph (graph.cu:63-119). */
int ibfs_hostcsr_load(const char *beg_path, const char *csr_path,
                      ibfs_hostcsr_t *out);
/* Writes the same two-file format (tuple_text_to_bin.cpp:149-178). */
int ibfs_hostcsr_save(ibfs_hostcsr_t h, const char *beg_path,
                      const char *csr_path);
int ibfs_hostcsr_create(const int64_t *beg_pos, int64_t vert_count,
                        const int32_t *adj, int64_t edge_count,
                        ibfs_hostcsr_t *out);
int ibfs_hostcsr_info(ibfs_hostcsr_t h, int64_t *vert_count,
                      int64_t *edge_count);
/* Borrowed views, valid until ibfs_hostcsr_free. */
int ibfs_hostcsr_data(ibfs_hostcsr_t h, const int64_t **beg_pos,
                      const int32_t **adj);
int ibfs_hostcsr_free(ibfs_hostcsr_t h);
/* Seeded undirected R-MAT (Graph500 a,b,c; d = 1-a-b-c), scale/edge factor,
 * self-loops and duplicates removed, symmetrised, optional seeded vertex
 * permutation, adjacency sorted ascending. No reference counterpart (the
 * reference reads SNAP files, readme.md:36-42). */
int ibfs_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                  double c, uint64_t seed, int32_t permute, int32_t nthreads,
                  ibfs_hostcsr_t *out);
/* Seeded Chung-Lu graph: expected degrees ~ power law with exponent gamma,
 * rescaled to mean_degree and capped at max_degree; endpoints drawn with
 * probability proportional to weight, cleaned like ibfs_gen_rmat. */
int ibfs_gen_chung_lu(int64_t vert_count, double gamma, double mean_degree,
                      double max_degree, uint64_t seed, int32_t nthreads,
                      ibfs_hostcsr_t *out);
/* Seeded distinct sources drawn uniformly from vertices with degree > 0
 * (replaces the unseeded rand() rejection loop of graph::gen_src,
 * graph.cu:121-148). */
int ibfs_gen_sources(ibfs_hostcsr_t h, int64_t count, uint64_t seed,
                     int32_t *out);
/* Text edge list -> CSR with the converter's rules (tuple_text_to_bin.cpp:
 * 56-272): skip `skip_lines` header lines, ids shifted by the minimum id,
 * reverse=1 adds the reversed edge, no dedup. */
int ibfs_convert_text(const char *text_path, int32_t reverse,
                      int64_t skip_lines, ibfs_hostcsr_t *out);

/* The converter's files from a text edge list (tuple_text_to_bin.cpp:56-310):
 * <out_prefix>_beg_pos.bin, _csr.bin, _head.bin, _weight.bin and _deg.bin,
 * little-endian int64, with the converter's rules (ids shifted by the
 * minimum, reverse=1 adds the reversed edge, no dedup; head = source of each
 * forward edge, 0 at reversed slots; weight = rand() % 63 + 1 per line from
 * the C library generator seeded with 1, as the reference's unseeded calls).
 * out_prefix null: the text path, as the reference names them.  V and E
 * (optional) are returned. */
int ibfs_convert_text_files(const char *text_path, int32_t reverse,
                            int64_t skip_lines, const char *out_prefix,
                            int64_t *vert_count, int64_t *edge_count);

/* ---- device graph -------------------------------------------------------- */
/* Device-side layout options of a graph (no effect on results). */
typedef struct ibfs_graph_opts {
  int32_t hub_count; /* highest-degree vertices whose status words are mirrored
                        in a compact per-level cache for bottom-up gathers   */
  int32_t nthreads;  /* host threads for the upload preprocessing (0: all)   */
  int32_t reserved[6];
} ibfs_graph_opts;
int ibfs_graph_opts_default(ibfs_graph_opts *opts);
/* Uploads a CSR to the device (int64 beg_pos, int32 adjacency, int32 degree,
 * plus a traversal copy of the adjacency ordered by descending neighbour
 * degree).  The device copy keeps only the vertices with an edge (in or
 * out), renumbered in their original order; isolated vertices get no status
 * words and their depths come from the source list alone.  Replaces the
 * graph upload in the gpu_ibfs constructor (gpu_ibfs.cu:54-116). */
int ibfs_graph_create(const int64_t *beg_pos, int64_t vert_count,
                      const int32_t *adj, int64_t edge_count, int device,
                      ibfs_graph_t *out);
int ibfs_graph_create_ex(const int64_t *beg_pos, int64_t vert_count,
                         const int32_t *adj, int64_t edge_count, int device,
                         const ibfs_graph_opts *opts, ibfs_graph_t *out);
int ibfs_graph_from_host(ibfs_hostcsr_t h, int device, ibfs_graph_t *out);
int ibfs_graph_load_files(const char *beg_path, const char *csr_path,
                          int device, ibfs_graph_t *out);
int ibfs_graph_info(ibfs_graph_t g, int64_t *vert_count, int64_t *edge_count);
/* Device layout sizes: vertices with status words (an edge in or out) and
 * hub-cache entries.  No reference counterpart (layout introspection). */
int ibfs_graph_layout(ibfs_graph_t g, int64_t *traversal_vertices, int64_t *hubs);
/* Adjacency entries that point to the top `top` hubs by degree (top <= 0 or
 * above the hub count: all hubs), the entries a hybrid level pulling through
 * them can read.  No reference counterpart (layout introspection). */
int ibfs_graph_hub_entries(ibfs_graph_t g, int64_t top, int64_t *entries);
int ibfs_graph_destroy(ibfs_graph_t g);

/* GroupBy on the device: a permutation of src_list in which sources of degree
 * <= 128 are ordered by their first neighbour of degree > 128 (stable), the
 * rest follow in input order.  Bit-identical to graph::groupby
 * (groupby.cu:36-149). */
int ibfs_groupby(ibfs_graph_t g, const int32_t *src_list, int64_t src_count,
                 int32_t *out_list);

/* ---- engine: one 128-lane (or 64/256) group per call --------------------- */
/* Replaces gpu_ibfs::gpu_ibfs (gpu_ibfs.cu:36-117). */
int ibfs_engine_create(ibfs_graph_t g, const ibfs_config *cfg,
                       ibfs_engine_t *out);
int ibfs_engine_destroy(ibfs_engine_t e);
/* Runs one group of k <= lane_bits sources (lane i = src[i]; duplicates and
 * zero-degree sources allowed) to completion and writes depth arrays
 * depth_out[i*V + v] (int32, -1 = unreached).  out_place selects a host or a
 * device pointer.  Replaces the per-group status reset (main.cu:79-84) plus
 * gpu_ibfs::traversal (traversal.cu:78-136). */
int ibfs_traversal(ibfs_engine_t e, const int32_t *src, int32_t k,
                   int32_t *depth_out, int32_t out_place, ibfs_stats *stats);
/* Level log of the last traversal. */
int ibfs_engine_level_log(ibfs_engine_t e, ibfs_level_info *out, int32_t cap,
                          int32_t *n_levels);

/* Allocation-free variant of ibfs_traversal.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Runs one
 * group of k <= 128 sources on graph g (cfg: null for the defaults;
 * ring_levels and split_cov are not used) and writes depth_out, a device
 * buffer of k*V int32, as ibfs_traversal does; *traversed_edges (optional)
 * receives the TEPS numerator.  It calls no allocating routine: every
 * scratch buffer lives in one file-scope device array of
 * IBFS_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends, and kernels run on the null stream.  Limits, rejected
 * with IBFS_ERR_ARG: at most 262,144 vertices with an edge, 2^24 adjacency
 * entries, 131,072 hubs, td_chunk >= 2048.  Calls on one device share the
 * array and its previous contents (what must start at zero is cleared by
 * each call); the library serialises calls within a process, and two
 * processes have separate copies of the array. */
#define IBFS_NOALLOC_SCRATCH_BYTES 43324928
int ibfs_traversal_noalloc(ibfs_graph_t g, const ibfs_config *cfg,
                           const int32_t *src, int32_t k, int32_t *depth_out,
                           int64_t *traversed_edges);

/* Runs src_count sources as consecutive groups of lane_bits (the reference's
 * main loop, main.cu:76-90); depth_out is [src_count][V]. */
int ibfs_run_groups(ibfs_engine_t e, const int32_t *src, int64_t src_count,
                    int32_t *depth_out, int32_t out_place, ibfs_stats *stats);

/* ---- batched launcher ---------------------------------------------------- */
/* Runs an arbitrary source set packed into lane groups of cfg->lane_bits with
 * `contexts` traversals in flight on separate HIP streams.  Each level is a
 * captured hipGraph of three kernels (expand, inspect, direction decision):
 * direction and queue sizes are decided on the device; the host polls once
 * per `burst` levels.  depth_out is [n][V] as for ibfs_run_groups.  No reference
 * counterpart: the reference runs one 128-source group at a time with ~13
 * host synchronisations per level (main.cu:76-90, traversal.cu:101-130). */
typedef struct ibfs_batch_s *ibfs_batch_t;
int ibfs_batch_create(ibfs_graph_t g, const ibfs_config *cfg, int32_t contexts,
                      int32_t burst, ibfs_batch_t *out);
int ibfs_batch_run(ibfs_batch_t b, const int32_t *src, int64_t n,
                   int32_t *depth_out, int32_t out_place, ibfs_stats *stats);
int ibfs_batch_destroy(ibfs_batch_t b);
/* Closeness inputs for every source without materialising depth arrays:
 * depth_sum[i] = sum of depth_i(v) over reached v, reach[i] = number of
 * vertices reached (source included).  Host arrays of n int64.  The
 * reference lists closeness as an intended use (readme.md:46-51). */
int ibfs_batch_closeness(ibfs_batch_t b, const int32_t *src, int64_t n,
                         int64_t *depth_sum, int64_t *reach, ibfs_stats *stats);

/* ---- serving: continuous batching ---------------------------------------- */
/* A worker thread serves BFS queries of 1..any sources against one device
 * graph.  Each source takes one status lane of one of the server's traversal
 * contexts from its admission level until its BFS ends; lanes free up as
 * sources finish and are reused two levels later, so queries arriving while
 * others run are admitted as lanes free and share levels with them.  The
 * number of contexts follows from status_bytes (ring_levels * V *
 * lane_bits/8 bytes each, at least one).  depth_out of a query is a device
 * buffer of n*V int32 (row i = source i, -1 unreached), valid once the query
 * is done.  No reference counterpart (the reference is a batch CLI,
 * main.cu:76-90). */
typedef struct ibfs_query_info {
  int64_t id;
  int32_t n_sources;
  int32_t done;
  double submit_ms, admit_ms, done_ms; /* steady-clock milliseconds */
} ibfs_query_info;
int ibfs_server_create(ibfs_graph_t g, const ibfs_config *cfg, int64_t status_bytes,
                       ibfs_server_t *out);
/* Admission policy: a source joins a running context only within
 * window_levels levels of the start of that context's current cohort (a
 * context whose lanes have all finished starts a new cohort with whatever is
 * waiting).  Lanes of one context then stay close in BFS phase, so a level's
 * direction and cost suit all of them.  0 admits only at cohort starts; a
 * large value admits at every level boundary.  Default 0.  No reference
 * counterpart (serving). */
int ibfs_server_set_admission(ibfs_server_t s, int32_t window_levels);
/* Level driver of the server's contexts.  host_levels = 0 (default): a
 * cohort's levels run on the device as captured level graphs (direction
 * decided on the device, one host poll per burst of levels); inside the
 * admission window the host takes every level boundary and waiting sources
 * join the running cohort there (a device-side admission list, applied
 * between the level's expansion and inspection); all lanes of a cohort are
 * emitted when it ends.  host_levels = 1: every level boundary on the host;
 * lanes whose BFS ended are emitted and freed one by one and reused two
 * levels later.  No reference counterpart (serving). */
int ibfs_server_set_levels(ibfs_server_t s, int32_t host_levels);
/* Counters since creation: level bursts/levels issued, sources admitted, and
 * sources that joined a cohort already running (after its first level). */
int ibfs_server_counters(ibfs_server_t s, int64_t *levels, int64_t *admitted,
                         int64_t *admitted_in_flight);
int ibfs_server_contexts(ibfs_server_t s, int32_t *contexts, int32_t *lanes);
int ibfs_server_submit(ibfs_server_t s, const int32_t *src, int32_t n,
                       int32_t *depth_out, int64_t *id);
/* Blocks until query id is done (timeout_ms < 0: no limit). */
int ibfs_server_wait(ibfs_server_t s, int64_t id, double timeout_ms,
                     ibfs_query_info *info);
/* Forgets a finished query. */
int ibfs_server_release(ibfs_server_t s, int64_t id);
int ibfs_server_destroy(ibfs_server_t s);

/* ---- autotuner ------------------------------------------------------------ */
/* Coordinate descent over lane width, alpha/beta, degree-class thresholds,
 * top-down chunk size and emission mode, scored by aggregate GTEPS of the
 * sample sources (best of `reps` timed runs per candidate, `passes` sweeps).
 * *best receives the chosen configuration; report (optional) a JSON record
 * of every trial.  No reference counterpart (thresholds are constants in
 * the reference, main.cu:56-57). */
int ibfs_autotune(ibfs_graph_t g, const int32_t *src, int64_t n,
                  const ibfs_config *start, int32_t passes, int32_t reps,
                  ibfs_config *best, char *report, size_t report_len);
/* The same search through the batched launcher (ibfs_batch_*), which adds
 * the device-driven expander's workgroup count (dd_grid), the number of
 * contexts and the burst length; contexts and burst are in/out. */
int ibfs_autotune_batch(ibfs_graph_t g, const int32_t *src, int64_t n,
                        const ibfs_config *start, int32_t contexts, int32_t burst,
                        int32_t passes, int32_t reps, ibfs_config *best,
                        int32_t *best_contexts, int32_t *best_burst, char *report,
                        size_t report_len);

/* TEPS numerator of one depth array: sum of deg(v) over reached v
 * (validate.h:59-76 report()). depth is a host array. */
int ibfs_traversed_edges(ibfs_hostcsr_t h, const int32_t *depth,
                         int64_t *edges);

#ifdef __cplusplus
}
#endif

#endif /* IBFS_H */
