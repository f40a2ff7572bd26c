/*
 * ibfs_oracle.c - CPU oracle for the concurrent multi-source BFS hot path.
 *
 * TEST INFRASTRUCTURE ONLY: see ibfs_oracle.h.  Plain C99 + pthreads, built
 * by oracle/Makefile into oracle/liboracle.so.
 */
#define _GNU_SOURCE
#include "ibfs_oracle.h"

#include <ctype.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

void oracle_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* Textbook BFS.  Depth = unique hop distance; the reference computes the same
 * quantity bitwise (expander.cu:100-359, inspector.cu:59-85). */
int oracle_bfs(int64_t V, const int64_t *beg, const int32_t *adj, int32_t src,
               int32_t *depth) {
  if (V < 0 || src < 0 || src >= V) return -1;
  int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)(V > 0 ? V : 1));
  if (!queue) return -5;
  for (int64_t v = 0; v < V; ++v) depth[v] = -1;
  int64_t head = 0, tail = 0;
  depth[src] = 0;
  queue[tail++] = src;
  while (head < tail) {
    int32_t u = queue[head++];
    int32_t du = depth[u] + 1;
    for (int64_t j = beg[u]; j < beg[u + 1]; ++j) {
      int32_t w = adj[j];
      if (depth[w] < 0) {
        depth[w] = du;
        queue[tail++] = w;
      }
    }
  }
  free(queue);
  return 0;
}

typedef struct {
  int64_t V;
  const int64_t *beg;
  const int32_t *adj;
  const int32_t *src;
  int64_t n;
  int32_t *depth;
  int64_t next; /* shared work counter */
  pthread_mutex_t mu;
  int err;
} many_ctx;

static void *many_worker(void *arg) {
  many_ctx *c = (many_ctx *)arg;
  for (;;) {
    pthread_mutex_lock(&c->mu);
    int64_t i = c->next++;
    pthread_mutex_unlock(&c->mu);
    if (i >= c->n) break;
    int rc = oracle_bfs(c->V, c->beg, c->adj, c->src[i], c->depth + i * c->V);
    if (rc) c->err = rc;
  }
  return NULL;
}

int oracle_bfs_many(int64_t V, const int64_t *beg, const int32_t *adj,
                    const int32_t *src, int64_t n, int32_t *depth,
                    int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  many_ctx c = {V, beg, adj, src, n, depth, 0, PTHREAD_MUTEX_INITIALIZER, 0};
  pthread_t th[256];
  for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, many_worker, &c);
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  return c.err;
}

/* ------------------------------------------------------------------------ */
/* Bitwise level-synchronous multi-source BFS (the reference's algorithm).
 * Status words of up to 256 bits as 4 x uint64; lane i is bit i%64 of word
 * i/64, the same memory image as the reference's uint4 with bit i%32 of
 * component i/32 (util.h:117-130). */
#define OW 4
typedef struct { uint64_t w[OW]; } oword;

static int ow_eq(const oword *a, const oword *b) {
  for (int i = 0; i < OW; ++i) if (a->w[i] != b->w[i]) return 0;
  return 1;
}

int oracle_ibfs_bitwise(int64_t V, const int64_t *beg, const int32_t *adj,
                        const int32_t *src, int k, int sw_level,
                        int32_t *depth) {
  if (k < 0 || k > 256 || V < 0) return -1;
  oword *last = (oword *)calloc((size_t)(V > 0 ? V : 1), sizeof(oword));
  oword *curr = (oword *)calloc((size_t)(V > 0 ? V : 1), sizeof(oword));
  if (!last || !curr) { free(last); free(curr); return -5; }
  for (int64_t i = 0; i < (int64_t)k * V; ++i) depth[i] = -1;
  oword all;
  memset(&all, 0, sizeof all);
  /* Lane init (expander.cu:42-76): bit i set in both words of src[i].  The
   * reference stores (losing a duplicate) and drops zero-degree sources; this
   * restatement ORs and keeps them, which is the BFS definition. */
  for (int i = 0; i < k; ++i) {
    if (src[i] < 0 || src[i] >= V) { free(last); free(curr); return -1; }
    last[src[i]].w[i >> 6] |= 1ull << (i & 63);
    curr[src[i]].w[i >> 6] |= 1ull << (i & 63);
    all.w[i >> 6] |= 1ull << (i & 63);
    depth[(int64_t)i * V + src[i]] = 0;
  }
  int level;
  for (level = 1;; ++level) {
    if (level <= sw_level) {
      /* top-down: every vertex whose word changed at level-1 pushes its
       * last word into its neighbours' curr words (expander.cu:128-145). */
      for (int64_t u = 0; u < V; ++u) {
        /* frontier = vertices that changed in the previous level; at level 1
         * those are the sources.  Here: last differs from the word two
         * levels back, tracked by `depth` below, so use the depth table. */
        int is_front = 0;
        for (int i = 0; i < k && !is_front; ++i)
          if (depth[(int64_t)i * V + u] == level - 1) is_front = 1;
        if (!is_front) continue;
        for (int64_t j = beg[u]; j < beg[u + 1]; ++j) {
          int32_t v = adj[j];
          for (int w = 0; w < OW; ++w) curr[v].w[w] |= last[u].w[w];
        }
      }
    } else {
      /* bottom-up: a vertex not yet holding every lane ORs its neighbours'
       * last words, stopping once it holds them all (expander.cu:340-356). */
      for (int64_t v = 0; v < V; ++v) {
        oword x = curr[v];
        for (int64_t j = beg[v]; j < beg[v + 1]; ++j) {
          if (ow_eq(&x, &all)) break;
          for (int w = 0; w < OW; ++w) x.w[w] |= last[adj[j]].w[w];
        }
        curr[v] = x;
      }
    }
    /* inspection (inspector.cu:59-85): new bits get depth=level, last=curr,
     * stop when nothing changed. */
    int changed = 0;
    for (int64_t v = 0; v < V; ++v) {
      if (ow_eq(&last[v], &curr[v])) continue;
      changed = 1;
      for (int w = 0; w < OW; ++w) {
        uint64_t nb = curr[v].w[w] & ~last[v].w[w];
        while (nb) {
          int b = __builtin_ctzll(nb);
          nb &= nb - 1;
          int lane = w * 64 + b;
          if (lane < k) depth[(int64_t)lane * V + v] = level;
        }
      }
      last[v] = curr[v];
    }
    if (!changed) break;
  }
  free(last);
  free(curr);
  return level;
}

/* ------------------------------------------------------------------------ */
/* GroupBy, restating groupby.cu:36-149: (1) a source of degree <= 128 whose
 * adjacency (in CSR order) contains a vertex of degree > 128 is "eligible"
 * and counted under the FIRST such hub (groupby.cu:49-66); (2) exclusive
 * prefix sum of the hub counts over all vertices (groupby.cu:69-73);
 * (3) eligible sources scatter to huboff[hub]++ in input order, all others
 * append after the eligible block in input order (groupby.cu:77-115). */
static int32_t first_hub(const int64_t *beg, const int32_t *adj, int32_t s) {
  int64_t b = beg[s], e = beg[s + 1];
  if (e - b > 128) return -1;
  for (int64_t j = b; j < e; ++j) {
    int32_t u = adj[j];
    if (beg[u + 1] - beg[u] > 128) return u;
  }
  return -1;
}

int oracle_groupby(int64_t V, const int64_t *beg, const int32_t *adj,
                   const int32_t *src, int64_t n, int32_t *out) {
  if (V <= 0 || n < 0) return -1;
  int64_t *hubmap = (int64_t *)calloc((size_t)V, sizeof(int64_t));
  int64_t *huboff = (int64_t *)calloc((size_t)V, sizeof(int64_t));
  if (!hubmap || !huboff) { free(hubmap); free(huboff); return -5; }
  for (int64_t i = 0; i < n; ++i) {
    int32_t h = first_hub(beg, adj, src[i]);
    if (h >= 0) hubmap[h]++;
  }
  huboff[0] = 0;
  for (int64_t i = 1; i < V; ++i) huboff[i] = huboff[i - 1] + hubmap[i - 1];
  int64_t remainoff = huboff[V - 1] + hubmap[V - 1];
  for (int64_t i = 0; i < n; ++i) {
    int32_t h = first_hub(beg, adj, src[i]);
    if (h >= 0) out[huboff[h]++] = src[i];
    else out[remainoff++] = src[i];
  }
  free(hubmap);
  free(huboff);
  return 0;
}

int oracle_degree_class(int64_t deg, int64_t sml, int64_t lrg) {
  if (deg <= 0) return 0;
  if (deg < sml) return 1;
  if (deg > lrg) return 3;
  return 2;
}

int64_t oracle_traversed_edges(int64_t V, const int64_t *beg,
                               const int32_t *depth) {
  int64_t e = 0;
  for (int64_t v = 0; v < V; ++v)
    if (depth[v] >= 0) e += beg[v + 1] - beg[v];
  return e;
}

int oracle_check_depth(int64_t V, const int64_t *beg, const int32_t *adj,
                       int32_t src, const int32_t *depth) {
  if (src < 0 || src >= V) return 1;
  if (depth[src] != 0) return 2;
  for (int64_t v = 0; v < V; ++v) {
    int32_t dv = depth[v];
    int has_parent = (v == src);
    for (int64_t j = beg[v]; j < beg[v + 1]; ++j) {
      int32_t du = depth[adj[j]];
      if ((dv < 0) != (du < 0)) return 3;       /* reached next to unreached */
      if (dv < 0) continue;
      if (du - dv > 1 || dv - du > 1) return 4; /* edge spans > 1 level */
      if (du == dv - 1) has_parent = 1;
    }
    if (dv > 0 && !has_parent) return 5;        /* no parent at depth-1 */
    if (dv == 0 && v != src) return 6;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Reference CSR reader, restating graph.cu:71-112. */
static long long file_size(const char *p) {
  struct stat st;
  if (stat(p, &st) != 0) return -1;
  return (long long)st.st_size;
}

int oracle_read_csr(const char *beg_path, const char *csr_path, int64_t *V,
                    int64_t *E, int64_t **beg, int32_t **adj) {
  long long bs = file_size(beg_path), cs = file_size(csr_path);
  if (bs < 0 || cs < 0) return -2;
  if (bs < 8 || bs % 8 || cs % 8) return -3;
  int64_t vc = bs / 8 - 1, ec = cs / 8;
  uint64_t *tb = (uint64_t *)malloc((size_t)bs);
  uint64_t *tc = (uint64_t *)malloc((size_t)(cs > 0 ? cs : 8));
  FILE *f = fopen(beg_path, "rb");
  if (!f || !tb || !tc) { if (f) fclose(f); free(tb); free(tc); return -2; }
  size_t r = fread(tb, 8, (size_t)(vc + 1), f);
  fclose(f);
  if ((int64_t)r != vc + 1) { free(tb); free(tc); return -2; }
  f = fopen(csr_path, "rb");
  if (!f) { free(tb); free(tc); return -2; }
  r = fread(tc, 8, (size_t)ec, f);
  fclose(f);
  if ((int64_t)r != ec) { free(tb); free(tc); return -2; }
  int64_t *b = (int64_t *)malloc(sizeof(int64_t) * (size_t)(vc + 1));
  int32_t *a = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ec > 0 ? ec : 1));
  int bad = (tb[0] != 0 || (int64_t)tb[vc] != ec);
  for (int64_t i = 0; i <= vc && !bad; ++i) {
    b[i] = (int64_t)tb[i];
    if (i && b[i] < b[i - 1]) bad = 1;
  }
  for (int64_t i = 0; i < ec && !bad; ++i) {
    if (tc[i] >= (uint64_t)vc || tc[i] > 0x7fffffffull) bad = 1;
    a[i] = (int32_t)tc[i];
  }
  free(tb);
  free(tc);
  if (bad) { free(b); free(a); return -3; }
  *V = vc; *E = ec; *beg = b; *adj = a;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Converter restatement (tuple_text_to_bin.cpp:56-272): tokens are
 * whitespace-separated integers; the first `skip_lines` lines are skipped;
 * V = max id - min id + 1 over all tokens; consecutive token pairs are edges
 * (src, dst) in file order; reverse=1 also inserts (dst, src).  Adjacency
 * order = file order per source (degree pass, exclusive prefix, scatter). */
int oracle_text_to_csr(const char *path, int reverse, int64_t skip_lines,
                       int64_t *V, int64_t *E, int64_t **beg, int32_t **adj) {
  FILE *f = fopen(path, "rb");
  if (!f) return -2;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *buf = (char *)malloc((size_t)sz + 1);
  if (!buf) { fclose(f); return -5; }
  if ((long)fread(buf, 1, (size_t)sz, f) != sz) { fclose(f); free(buf); return -2; }
  fclose(f);
  buf[sz] = 0;
  long off = 0;
  for (int64_t s = 0; s < skip_lines; ++s) {
    while (off < sz && buf[off] != '\n') ++off;
    if (off >= sz) { free(buf); return -3; }
    ++off;
  }
  /* tokenise */
  int64_t cap = 1024, nt = 0;
  long long *tok = (long long *)malloc(sizeof(long long) * (size_t)cap);
  long p = off;
  while (p < sz) {
    while (p < sz && isspace((unsigned char)buf[p])) ++p;
    if (p >= sz) break;
    char *end;
    long long x = strtoll(buf + p, &end, 10);
    if (end == buf + p) { free(tok); free(buf); return -3; }
    if (nt == cap) { cap *= 2; tok = (long long *)realloc(tok, sizeof(long long) * (size_t)cap); }
    tok[nt++] = x;
    p = (long)(end - buf);
  }
  free(buf);
  if (nt < 2) { free(tok); return -3; }
  long long vmin = tok[0], vmax = tok[0];
  for (int64_t i = 0; i < nt; ++i) {
    if (tok[i] < vmin) vmin = tok[i];
    if (tok[i] > vmax) vmax = tok[i];
  }
  int64_t lines = nt / 2;
  int64_t vc = (int64_t)(vmax - vmin + 1);
  int64_t ec = reverse ? 2 * lines : lines;
  int64_t *b = (int64_t *)calloc((size_t)vc + 1, sizeof(int64_t));
  int64_t *deg = (int64_t *)calloc((size_t)vc, sizeof(int64_t));
  int32_t *a = (int32_t *)calloc((size_t)(ec > 0 ? ec : 1), sizeof(int32_t));
  for (int64_t l = 0; l < lines; ++l) {
    int64_t s = tok[2 * l] - vmin, d = tok[2 * l + 1] - vmin;
    deg[s]++;
    if (reverse) deg[d]++;
  }
  for (int64_t i = 0; i < vc; ++i) b[i + 1] = b[i] + deg[i];
  memset(deg, 0, sizeof(int64_t) * (size_t)vc);
  for (int64_t l = 0; l < lines; ++l) {
    int64_t s = tok[2 * l] - vmin, d = tok[2 * l + 1] - vmin;
    /* both slots are chosen before either degree moves, so a reversed
     * self-loop writes one slot twice and leaves the next at 0
     * (tuple_text_to_bin.cpp:245-258) */
    a[b[s] + deg[s]] = (int32_t)d;
    if (reverse) a[b[d] + deg[d]] = (int32_t)s;
    deg[s]++;
    if (reverse) deg[d]++;
  }
  free(deg);
  free(tok);
  *V = vc; *E = ec; *beg = b; *adj = a;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Work model of the engine's level loop.  Status words as in
 * oracle_ibfs_bitwise; lam[L] = lanes with a vertex at depth L.  Candidates
 * for the bottom-up pass of level L+1 are the vertices of degree > 0 whose
 * word at L misses a lane of lam[L-1] (lam[-1] = all lanes with a source);
 * the pass of level L+1 reads, per candidate, the shortest adjacency prefix
 * after which word | OR(neighbour words & lam[L]) covers lam[L] (the whole
 * list if never), or nothing if the word already covers lam[L]. */
int oracle_level_model(int64_t V, const int64_t *beg, const int32_t *adj,
                       const int32_t *src, int k, const int8_t *dirs,
                       int max_levels, int64_t *n_changed, int64_t *n_front,
                       int64_t *m_front, int64_t *n_cand, int64_t *m_cand,
                       int64_t *bu_prefix, int32_t *lam_pop) {
  if (k < 1 || k > 256 || max_levels < 2) return -1;
  int32_t *depth = (int32_t *)malloc(sizeof(int32_t) * (size_t)k * (size_t)V);
  if (!depth) return -5;
  int levels = oracle_ibfs_bitwise(V, beg, adj, src, k, 1 << 30, depth);
  /* levels = index of the first level without change; depths 0..levels-1 */
  if (levels > max_levels) { free(depth); return -4; }
  oword *S = (oword *)calloc((size_t)V, sizeof(oword));     /* S_L */
  oword *P = (oword *)calloc((size_t)V, sizeof(oword));     /* S_{L-1} */
  oword lamprev, lam;
  memset(&lamprev, 0, sizeof lamprev);
  for (int i = 0; i < k; ++i) lamprev.w[i >> 6] |= 1ull << (i & 63);
  for (int L = 0; L < levels + 1 && L < max_levels; ++L) {
    /* build S_L */
    memcpy(P, S, sizeof(oword) * (size_t)V);
    memset(&lam, 0, sizeof lam);
    for (int64_t v = 0; v < V; ++v)
      for (int i = 0; i < k; ++i) {
        int32_t d = depth[(int64_t)i * V + v];
        if (d >= 0 && d <= L) S[v].w[i >> 6] |= 1ull << (i & 63);
        if (d == L) lam.w[i >> 6] |= 1ull << (i & 63);
      }
    /* bottom-up prefix of level L (reads S_{L-1} = P, target lam[L-1]) */
    int64_t pref = 0;
    if (L >= 1 && dirs && dirs[L] == 1) {
      for (int64_t v = 0; v < V; ++v) {
        int64_t d = beg[v + 1] - beg[v];
        if (d == 0) continue;
        /* candidate test of inspect(L-1): misses a lane of lam[L-2]; the
         * pass itself targets lam[L-1] (== lamprev here) */
        oword x = P[v];
        int cover = 1;
        for (int w = 0; w < OW; ++w) if ((x.w[w] & lamprev.w[w]) != lamprev.w[w]) cover = 0;
        if (cover) continue;
        for (int64_t j = beg[v]; j < beg[v + 1]; ++j) {
          for (int w = 0; w < OW; ++w) x.w[w] |= P[adj[j]].w[w] & lamprev.w[w];
          ++pref;
          cover = 1;
          for (int w = 0; w < OW; ++w) if ((x.w[w] & lamprev.w[w]) != lamprev.w[w]) cover = 0;
          if (cover) break;
        }
      }
    }
    if (bu_prefix) bu_prefix[L] = pref;
    /* inspection of S_L */
    int64_t nc = 0, nf = 0, mf = 0, nu = 0, mu = 0;
    for (int64_t v = 0; v < V; ++v) {
      int64_t d = beg[v + 1] - beg[v];
      int ch = 0;
      for (int w = 0; w < OW; ++w) if (S[v].w[w] != P[v].w[w]) ch = 1;
      if (ch) { nc++; if (d > 0) { nf++; mf += d; } }
      int cover = 1;
      for (int w = 0; w < OW; ++w) if ((S[v].w[w] & lamprev.w[w]) != lamprev.w[w]) cover = 0;
      if (d > 0 && !cover) { nu++; mu += d; }
    }
    if (n_changed) n_changed[L] = nc;
    if (n_front) n_front[L] = nf;
    if (m_front) m_front[L] = mf;
    if (n_cand) n_cand[L] = nu;
    if (m_cand) m_cand[L] = mu;
    int pc = 0;
    for (int w = 0; w < OW; ++w) pc += __builtin_popcountll(lam.w[w]);
    if (lam_pop) lam_pop[L] = pc;
    lamprev = lam;
  }
  free(S);
  free(P);
  free(depth);
  return levels;
}
