// Host CSR: the reference's two-file binary format, seeded graph generators,
// the text-edge-list converter and seeded source selection.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sys/stat.h>
#include <thread>

#include "common.h"

namespace ibfs {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

const std::string &last_error() { return g_last_error; }

static std::atomic<double> g_wait_timeout_s{120.0};

double wait_timeout_s() { return g_wait_timeout_s.load(std::memory_order_relaxed); }

static double steady_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

WaitClock::WaitClock() : t0_(steady_s()) {}

bool WaitClock::expired() {
  if ((++n_ & 15u) != 0) return false;
  return steady_s() - t0_ > wait_timeout_s();
}

void WaitClock::restart() {
  t0_ = steady_s();
  n_ = 0;
}

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void parallel_for_threads(int nthreads, void (*fn)(int, void *), void *ctx) {
  if (nthreads <= 1) {
    fn(0, ctx);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nthreads);
  for (int t = 0; t < nthreads; ++t) th.emplace_back(fn, t, ctx);
  for (auto &x : th) x.join();
}

// SplitMix64: the generators' only random source.  Edge e of a generator
// seeded with s draws from the stream started at mix(s, e), so the output
// does not depend on the thread count.
static inline uint64_t splitmix_next(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t stream_seed(uint64_t seed, uint64_t idx) {
  uint64_t x = seed * 0xD1B54A32D192ED03ull + idx;
  uint64_t a = splitmix_next(x);
  return a ^ (idx * 0x9E3779B97F4A7C15ull);
}
static inline double u01(uint64_t &x) {
  return (double)(splitmix_next(x) >> 11) * (1.0 / 9007199254740992.0);
}

static int default_threads(int n) {
  if (n > 0) return n;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 64u) : 1;
}

// Builds a symmetric, loop-free, duplicate-free CSR with ascending adjacency
// from an undirected edge list (entries with src<0 are skipped).
static int build_undirected(int64_t V, const std::vector<int32_t> &src,
                            const std::vector<int32_t> &dst, int nthreads,
                            HostCsr &out) {
  const int64_t M = (int64_t)src.size();
  std::vector<std::atomic<int64_t>> cnt(V);
  for (int64_t v = 0; v < V; ++v) cnt[v].store(0, std::memory_order_relaxed);
  struct Ctx {
    int64_t V, M;
    const int32_t *s, *d;
    std::atomic<int64_t> *cnt;
    int64_t *beg;
    int32_t *tmp;
    int nt;
  } c{V, M, src.data(), dst.data(), cnt.data(), nullptr, nullptr, nthreads};
  parallel_for_threads(nthreads, [](int t, void *p) {
    Ctx &c = *(Ctx *)p;
    int64_t lo = c.M * t / c.nt, hi = c.M * (t + 1) / c.nt;
    for (int64_t i = lo; i < hi; ++i) {
      int32_t a = c.s[i], b = c.d[i];
      if (a < 0 || a == b) continue;
      c.cnt[a].fetch_add(1, std::memory_order_relaxed);
      c.cnt[b].fetch_add(1, std::memory_order_relaxed);
    }
  }, &c);
  std::vector<int64_t> beg(V + 1, 0);
  for (int64_t v = 0; v < V; ++v) beg[v + 1] = beg[v] + cnt[v].load();
  std::vector<int32_t> tmp((size_t)std::max<int64_t>(beg[V], 1));
  for (int64_t v = 0; v < V; ++v) cnt[v].store(beg[v], std::memory_order_relaxed);
  c.tmp = tmp.data();
  parallel_for_threads(nthreads, [](int t, void *p) {
    Ctx &c = *(Ctx *)p;
    int64_t lo = c.M * t / c.nt, hi = c.M * (t + 1) / c.nt;
    for (int64_t i = lo; i < hi; ++i) {
      int32_t a = c.s[i], b = c.d[i];
      if (a < 0 || a == b) continue;
      c.tmp[c.cnt[a].fetch_add(1, std::memory_order_relaxed)] = b;
      c.tmp[c.cnt[b].fetch_add(1, std::memory_order_relaxed)] = a;
    }
  }, &c);
  // sort + dedup each list in place; new degree into deg
  std::vector<int64_t> deg(V, 0);
  struct Ctx2 {
    int64_t V;
    const int64_t *beg;
    int32_t *tmp;
    int64_t *deg;
    int nt;
  } c2{V, beg.data(), tmp.data(), deg.data(), nthreads};
  parallel_for_threads(nthreads, [](int t, void *p) {
    Ctx2 &c = *(Ctx2 *)p;
    for (int64_t v = t; v < c.V; v += c.nt) {
      int32_t *a = c.tmp + c.beg[v], *e = c.tmp + c.beg[v + 1];
      std::sort(a, e);
      c.deg[v] = std::unique(a, e) - a;
    }
  }, &c2);
  out.V = V;
  out.beg.assign(V + 1, 0);
  for (int64_t v = 0; v < V; ++v) out.beg[v + 1] = out.beg[v] + deg[v];
  out.E = out.beg[V];
  out.adj.resize((size_t)out.E);
  for (int64_t v = 0; v < V; ++v)
    std::memcpy(out.adj.data() + out.beg[v], tmp.data() + beg[v],
                sizeof(int32_t) * (size_t)deg[v]);
  return IBFS_OK;
}

}  // namespace ibfs

using namespace ibfs;

extern "C" {

int ibfs_abi_version(void) { return IBFS_ABI_VERSION; }

int ibfs_set_wait_timeout(double seconds) {
  if (!(seconds > 0) || seconds > 1e7)
    return fail(IBFS_ERR_ARG, "wait timeout must be in (0, 1e7] seconds (got %g)", seconds);
  ibfs::g_wait_timeout_s.store(seconds, std::memory_order_relaxed);
  return IBFS_OK;
}

int ibfs_get_wait_timeout(double *seconds) {
  if (!seconds) return fail(IBFS_ERR_ARG, "null argument");
  *seconds = wait_timeout_s();
  return IBFS_OK;
}

int ibfs_last_error(char *buf, size_t n) {
  if (!buf || n == 0) return IBFS_ERR_ARG;
  std::snprintf(buf, n, "%s", last_error().c_str());
  return IBFS_OK;
}

int ibfs_hostcsr_create(const int64_t *beg_pos, int64_t V, const int32_t *adj,
                        int64_t E, ibfs_hostcsr_t *out) {
  if (!out || V < 0 || E < 0 || !beg_pos || (E > 0 && !adj))
    return fail(IBFS_ERR_ARG, "ibfs_hostcsr_create: bad arguments");
  if (V > 0x7fffffffLL) return fail(IBFS_ERR_ARG, "vertex count %lld exceeds int32 ids", (long long)V);
  if (beg_pos[0] != 0 || beg_pos[V] != E)
    return fail(IBFS_ERR_FORMAT, "beg_pos[0]=%lld beg_pos[V]=%lld, expected 0 and E=%lld",
                (long long)beg_pos[0], (long long)beg_pos[V], (long long)E);
  for (int64_t v = 0; v < V; ++v)
    if (beg_pos[v + 1] < beg_pos[v])
      return fail(IBFS_ERR_FORMAT, "beg_pos decreases at vertex %lld", (long long)v);
  for (int64_t i = 0; i < E; ++i)
    if (adj[i] < 0 || adj[i] >= V)
      return fail(IBFS_ERR_FORMAT, "adjacency entry %lld = %d out of range [0,%lld)",
                  (long long)i, adj[i], (long long)V);
  auto *h = new ibfs_hostcsr_s;
  h->csr.V = V;
  h->csr.E = E;
  h->csr.beg.assign(beg_pos, beg_pos + V + 1);
  h->csr.adj.assign(adj, adj + E);
  *out = h;
  return IBFS_OK;
}

int ibfs_hostcsr_info(ibfs_hostcsr_t h, int64_t *V, int64_t *E) {
  if (!h) return fail(IBFS_ERR_ARG, "null host csr");
  if (V) *V = h->csr.V;
  if (E) *E = h->csr.E;
  return IBFS_OK;
}

int ibfs_hostcsr_data(ibfs_hostcsr_t h, const int64_t **beg, const int32_t **adj) {
  if (!h) return fail(IBFS_ERR_ARG, "null host csr");
  if (beg) *beg = h->csr.beg.data();
  if (adj) *adj = h->csr.adj.data();
  return IBFS_OK;
}

int ibfs_hostcsr_free(ibfs_hostcsr_t h) {
  delete h;
  return IBFS_OK;
}

static long long fsize_bytes(const char *p) {
  struct stat st;
  if (stat(p, &st) != 0) return -1;
  return (long long)st.st_size;
}

// Reader of the reference format: graph.cu:71-112 (uint64 entries, V and E
// from the file sizes, narrowed to 32-bit ids) with the range checks the
// reference leaves to assert().
int ibfs_hostcsr_load(const char *beg_path, const char *csr_path, ibfs_hostcsr_t *out) {
  if (!beg_path || !csr_path || !out) return fail(IBFS_ERR_ARG, "ibfs_hostcsr_load: null argument");
  long long bs = fsize_bytes(beg_path), cs = fsize_bytes(csr_path);
  if (bs < 0) return fail(IBFS_ERR_IO, "%s cannot open", beg_path);
  if (cs < 0) return fail(IBFS_ERR_IO, "%s cannot open", csr_path);
  if (bs < 8 || bs % 8) return fail(IBFS_ERR_FORMAT, "%s: size %lld is not a positive multiple of 8", beg_path, bs);
  if (cs % 8) return fail(IBFS_ERR_FORMAT, "%s: size %lld is not a multiple of 8", csr_path, cs);
  int64_t V = bs / 8 - 1, E = cs / 8;
  if (V > 0x7fffffffLL) return fail(IBFS_ERR_FORMAT, "vertex count %lld exceeds int32 ids", (long long)V);
  std::vector<uint64_t> tb((size_t)V + 1), tc((size_t)E);
  FILE *f = fopen(beg_path, "rb");
  if (!f) return fail(IBFS_ERR_IO, "%s cannot open", beg_path);
  size_t r = fread(tb.data(), 8, (size_t)V + 1, f);
  fclose(f);
  if ((int64_t)r != V + 1) return fail(IBFS_ERR_IO, "%s: short read", beg_path);
  f = fopen(csr_path, "rb");
  if (!f) return fail(IBFS_ERR_IO, "%s cannot open", csr_path);
  r = E ? fread(tc.data(), 8, (size_t)E, f) : 0;
  fclose(f);
  if ((int64_t)r != E) return fail(IBFS_ERR_IO, "%s: short read", csr_path);
  auto *h = new ibfs_hostcsr_s;
  h->csr.V = V;
  h->csr.E = E;
  h->csr.beg.resize((size_t)V + 1);
  h->csr.adj.resize((size_t)E);
  if (tb[0] != 0 || tb[V] != (uint64_t)E) {
    delete h;
    return fail(IBFS_ERR_FORMAT, "beg_pos[0]=%llu beg_pos[V]=%llu, expected 0 and %lld",
                (unsigned long long)tb[0], (unsigned long long)tb[V], (long long)E);
  }
  for (int64_t i = 0; i <= V; ++i) {
    if (i && tb[i] < tb[i - 1]) {
      delete h;
      return fail(IBFS_ERR_FORMAT, "beg_pos decreases at %lld", (long long)i);
    }
    h->csr.beg[i] = (int64_t)tb[i];
  }
  for (int64_t i = 0; i < E; ++i) {
    if (tc[i] >= (uint64_t)V) {
      delete h;
      return fail(IBFS_ERR_FORMAT, "csr entry %lld = %llu out of range", (long long)i,
                  (unsigned long long)tc[i]);
    }
    h->csr.adj[i] = (int32_t)tc[i];
  }
  *out = h;
  return IBFS_OK;
}

int ibfs_hostcsr_save(ibfs_hostcsr_t h, const char *beg_path, const char *csr_path) {
  if (!h || !beg_path || !csr_path) return fail(IBFS_ERR_ARG, "ibfs_hostcsr_save: null argument");
  FILE *f = fopen(beg_path, "wb");
  if (!f) return fail(IBFS_ERR_IO, "%s cannot open for writing", beg_path);
  std::vector<uint64_t> buf(1 << 16);
  auto dump = [&](FILE *f, auto get, int64_t n) -> bool {
    for (int64_t i = 0; i < n; i += (int64_t)buf.size()) {
      int64_t m = std::min<int64_t>((int64_t)buf.size(), n - i);
      for (int64_t j = 0; j < m; ++j) buf[j] = (uint64_t)get(i + j);
      if ((int64_t)fwrite(buf.data(), 8, (size_t)m, f) != m) return false;
    }
    return true;
  };
  bool ok = dump(f, [&](int64_t i) { return h->csr.beg[i]; }, h->csr.V + 1);
  fclose(f);
  if (!ok) return fail(IBFS_ERR_IO, "%s: short write", beg_path);
  f = fopen(csr_path, "wb");
  if (!f) return fail(IBFS_ERR_IO, "%s cannot open for writing", csr_path);
  ok = dump(f, [&](int64_t i) { return h->csr.adj[i]; }, h->csr.E);
  fclose(f);
  if (!ok) return fail(IBFS_ERR_IO, "%s: short write", csr_path);
  return IBFS_OK;
}

int ibfs_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                  uint64_t seed, int32_t permute, int32_t nthreads, ibfs_hostcsr_t *out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0)
    return fail(IBFS_ERR_ARG, "ibfs_gen_rmat: bad arguments");
  const int64_t V = 1ll << scale;
  const int64_t M = (int64_t)edge_factor << scale;
  const int nt = default_threads(nthreads);
  std::vector<int32_t> src((size_t)M), dst((size_t)M);
  // seeded vertex permutation (Fisher-Yates over one stream)
  std::vector<int32_t> perm((size_t)V);
  for (int64_t v = 0; v < V; ++v) perm[v] = (int32_t)v;
  if (permute) {
    uint64_t st = stream_seed(seed, 0xFFFFFFFFFFull);
    for (int64_t i = V - 1; i > 0; --i) {
      int64_t j = (int64_t)(splitmix_next(st) % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
  }
  struct Ctx {
    int64_t M;
    int scale;
    double a, ab, abc;
    uint64_t seed;
    int32_t *s, *d;
    const int32_t *perm;
    int nt;
  } cx{M, scale, a, a + b, a + b + c, seed, src.data(), dst.data(), perm.data(), nt};
  parallel_for_threads(nt, [](int t, void *p) {
    Ctx &c = *(Ctx *)p;
    int64_t lo = c.M * t / c.nt, hi = c.M * (t + 1) / c.nt;
    for (int64_t e = lo; e < hi; ++e) {
      uint64_t st = stream_seed(c.seed, (uint64_t)e);
      int64_t u = 0, v = 0;
      for (int l = 0; l < c.scale; ++l) {
        double r = u01(st);
        int bu = 0, bv = 0;
        if (r < c.a) {
        } else if (r < c.ab) {
          bv = 1;
        } else if (r < c.abc) {
          bu = 1;
        } else {
          bu = 1;
          bv = 1;
        }
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
      c.s[e] = c.perm[u];
      c.d[e] = c.perm[v];
    }
  }, &cx);
  auto *h = new ibfs_hostcsr_s;
  build_undirected(V, src, dst, nt, h->csr);
  *out = h;
  return IBFS_OK;
}

int ibfs_gen_chung_lu(int64_t V, double gamma, double mean_degree, double max_degree,
                      uint64_t seed, int32_t nthreads, ibfs_hostcsr_t *out) {
  if (!out || V < 2 || V > 0x7fffffffLL || gamma <= 1.0 || mean_degree <= 0 ||
      max_degree < mean_degree)
    return fail(IBFS_ERR_ARG, "ibfs_gen_chung_lu: bad arguments");
  const int nt = default_threads(nthreads);
  // expected degrees: Pareto(x_min=1, exponent gamma) per vertex, then
  // alternately capped at max_degree and rescaled to mean_degree
  std::vector<double> w((size_t)V);
  {
    uint64_t st = stream_seed(seed, 0xC0FFEEull);
    for (int64_t i = 0; i < V; ++i) {
      double u = 1.0 - u01(st);  // (0,1]
      w[i] = std::pow(u, -1.0 / (gamma - 1.0));
    }
  }
  for (int it = 0; it < 64; ++it) {
    double s = 0;
    for (double x : w) s += x;
    double f = mean_degree * (double)V / s;
    bool capped = false;
    for (double &x : w) {
      x *= f;
      if (x > max_degree) {
        x = max_degree;
        capped = true;
      }
    }
    if (!capped) break;
  }
  double total = 0;
  for (double x : w) total += x;
  const int64_t M = (int64_t)std::llround(total / 2.0);
  // Vose alias table over weights
  std::vector<double> prob((size_t)V);
  std::vector<int32_t> alias((size_t)V);
  {
    std::vector<int32_t> small, large;
    small.reserve(V);
    large.reserve(V);
    std::vector<double> q((size_t)V);
    for (int64_t i = 0; i < V; ++i) {
      q[i] = w[i] * (double)V / total;
      (q[i] < 1.0 ? small : large).push_back((int32_t)i);
    }
    while (!small.empty() && !large.empty()) {
      int32_t s = small.back(), l = large.back();
      small.pop_back();
      prob[s] = q[s];
      alias[s] = l;
      q[l] = (q[l] + q[s]) - 1.0;
      if (q[l] < 1.0) {
        large.pop_back();
        small.push_back(l);
      }
    }
    for (int32_t l : large) { prob[l] = 1.0; alias[l] = l; }
    for (int32_t s : small) { prob[s] = 1.0; alias[s] = s; }
  }
  std::vector<int32_t> src((size_t)M), dst((size_t)M);
  struct Ctx {
    int64_t M, V;
    uint64_t seed;
    const double *prob;
    const int32_t *alias;
    int32_t *s, *d;
    int nt;
  } cx{M, V, seed, prob.data(), alias.data(), src.data(), dst.data(), nt};
  parallel_for_threads(nt, [](int t, void *p) {
    Ctx &c = *(Ctx *)p;
    int64_t lo = c.M * t / c.nt, hi = c.M * (t + 1) / c.nt;
    for (int64_t e = lo; e < hi; ++e) {
      uint64_t st = stream_seed(c.seed, (uint64_t)e);
      int32_t ends[2];
      for (int k = 0; k < 2; ++k) {
        uint64_t r = splitmix_next(st);
        int64_t i = (int64_t)((r >> 32) * (uint64_t)c.V >> 32);
        double f = (double)(r & 0xFFFFFFFFull) * (1.0 / 4294967296.0);
        ends[k] = f < c.prob[i] ? (int32_t)i : c.alias[i];
      }
      c.s[e] = ends[0];
      c.d[e] = ends[1];
    }
  }, &cx);
  auto *h = new ibfs_hostcsr_s;
  build_undirected(V, src, dst, nt, h->csr);
  *out = h;
  return IBFS_OK;
}

int ibfs_gen_sources(ibfs_hostcsr_t h, int64_t count, uint64_t seed, int32_t *out) {
  if (!h || !out || count < 0) return fail(IBFS_ERR_ARG, "ibfs_gen_sources: bad arguments");
  std::vector<int32_t> cand;
  for (int64_t v = 0; v < h->csr.V; ++v)
    if (h->csr.beg[v + 1] > h->csr.beg[v]) cand.push_back((int32_t)v);
  if ((int64_t)cand.size() < count)
    return fail(IBFS_ERR_ARG, "only %zu vertices with degree > 0, %lld sources requested",
                cand.size(), (long long)count);
  uint64_t st = stream_seed(seed, 0x5050ull);
  const int64_t n = (int64_t)cand.size();
  for (int64_t i = 0; i < count; ++i) {
    int64_t j = i + (int64_t)(splitmix_next(st) % (uint64_t)(n - i));
    std::swap(cand[i], cand[j]);
    out[i] = cand[i];
  }
  return IBFS_OK;
}

// Converter with the rules of tuple_text_to_bin.cpp:56-272: whitespace
// separated integer tokens after `skip_lines` lines; V = max - min + 1 over
// all tokens, ids shifted by the minimum; token pairs are (src, dst) in file
// order; reverse=1 adds (dst, src); adjacency keeps file order; no dedup.
int ibfs_convert_text(const char *path, int32_t reverse, int64_t skip_lines,
                      ibfs_hostcsr_t *out) {
  if (!path || !out || skip_lines < 0) return fail(IBFS_ERR_ARG, "ibfs_convert_text: bad arguments");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(IBFS_ERR_IO, "%s open error", path);
  std::string text;
  {
    char buf[1 << 16];
    size_t r;
    while ((r = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, r);
    fclose(f);
  }
  size_t off = 0;
  for (int64_t s = 0; s < skip_lines; ++s) {
    size_t nl = text.find('\n', off);
    if (nl == std::string::npos) return fail(IBFS_ERR_FORMAT, "skip more lines than the file has");
    off = nl + 1;
  }
  std::vector<long long> tok;
  const char *p = text.c_str() + off, *end = text.c_str() + text.size();
  while (p < end) {
    while (p < end && std::isspace((unsigned char)*p)) ++p;
    if (p >= end) break;
    char *q;
    long long x = std::strtoll(p, &q, 10);
    if (q == p) return fail(IBFS_ERR_FORMAT, "non-integer token at byte %lld", (long long)(p - text.c_str()));
    tok.push_back(x);
    p = q;
  }
  if (tok.size() < 2) return fail(IBFS_ERR_FORMAT, "fewer than two tokens");
  long long vmin = *std::min_element(tok.begin(), tok.end());
  long long vmax = *std::max_element(tok.begin(), tok.end());
  if (vmax - vmin + 1 > 0x7fffffffLL) return fail(IBFS_ERR_FORMAT, "id range exceeds int32");
  const int64_t V = vmax - vmin + 1, lines = (int64_t)tok.size() / 2;
  const int64_t E = reverse ? 2 * lines : lines;
  auto *h = new ibfs_hostcsr_s;
  h->csr.V = V;
  h->csr.E = E;
  h->csr.beg.assign((size_t)V + 1, 0);
  h->csr.adj.assign((size_t)E, 0);
  std::vector<int64_t> deg((size_t)V, 0);
  for (int64_t l = 0; l < lines; ++l) {
    deg[tok[2 * l] - vmin]++;
    if (reverse) deg[tok[2 * l + 1] - vmin]++;
  }
  for (int64_t v = 0; v < V; ++v) h->csr.beg[v + 1] = h->csr.beg[v] + deg[v];
  std::fill(deg.begin(), deg.end(), 0);
  for (int64_t l = 0; l < lines; ++l) {
    int64_t s = tok[2 * l] - vmin, d = tok[2 * l + 1] - vmin;
    // Both writes of a line target their slot before either degree moves
    // (tuple_text_to_bin.cpp:245-258), so a reversed self-loop fills one
    // slot and leaves the next one 0.
    h->csr.adj[h->csr.beg[s] + deg[s]] = (int32_t)d;
    if (reverse) h->csr.adj[h->csr.beg[d] + deg[d]] = (int32_t)s;
    deg[s]++;
    if (reverse) deg[d]++;
  }
  *out = h;
  return IBFS_OK;
}

// The converter's five output files (tuple_text_to_bin.cpp:56-310):
// <prefix>_beg_pos.bin (V+1), _csr.bin, _head.bin, _weight.bin (E each) and
// _deg.bin (V), little-endian int64.  E counts tokens (reverse=1) or token
// pairs (reverse=0) as the reference does, so with reverse=1 and an odd
// trailing token the last slot stays 0.  head[p] = source of the forward edge
// written at slot p (slots of reversed edges stay 0); weight = rand() % 63 + 1
// per input line from the C library generator seeded with 1 (the reference
// calls rand() unseeded), written at the forward slot and, with reverse=1,
// at the reversed slot.  The files are created fresh (the reference keeps a
// pre-existing file's bytes in slots it does not write).
static int write_i64_file(const std::string &path, const int64_t *v, size_t n) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return fail(IBFS_ERR_IO, "%s open error", path.c_str());
  const size_t w = n ? fwrite(v, sizeof(int64_t), n, f) : 0;
  const int c = fclose(f);
  if (w != n || c != 0) return fail(IBFS_ERR_IO, "%s write error", path.c_str());
  return IBFS_OK;
}

int ibfs_convert_text_files(const char *path, int32_t reverse, int64_t skip_lines,
                            const char *out_prefix, int64_t *vert_count, int64_t *edge_count) {
  if (!path || skip_lines < 0) return fail(IBFS_ERR_ARG, "ibfs_convert_text_files: bad arguments");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(IBFS_ERR_IO, "%s open error", path);
  std::string text;
  {
    char buf[1 << 16];
    size_t r;
    while ((r = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, r);
    fclose(f);
  }
  size_t off = 0;
  for (int64_t k = 0; k < skip_lines; ++k) {
    size_t nl = text.find('\n', off);
    if (nl == std::string::npos) return fail(IBFS_ERR_FORMAT, "skip more lines than the file has");
    off = nl + 1;
  }
  std::vector<long long> tok;
  const char *p = text.c_str() + off, *end = text.c_str() + text.size();
  while (p < end) {
    while (p < end && std::isspace((unsigned char)*p)) ++p;
    if (p >= end) break;
    char *q;
    long long x = std::strtoll(p, &q, 10);
    if (q == p) return fail(IBFS_ERR_FORMAT, "non-integer token at byte %lld", (long long)(p - text.c_str()));
    tok.push_back(x);
    p = q;
  }
  if (tok.size() < 2) return fail(IBFS_ERR_FORMAT, "fewer than two tokens");
  const long long vmin = *std::min_element(tok.begin(), tok.end());
  const long long vmax = *std::max_element(tok.begin(), tok.end());
  const int64_t V = vmax - vmin + 1, lines = (int64_t)tok.size() / 2;
  const int64_t E = reverse ? (int64_t)tok.size() : lines;
  std::vector<int64_t> beg((size_t)V + 1, 0), deg((size_t)V, 0), adj((size_t)E, 0),
      head((size_t)E, 0), weight((size_t)E, 0);
  for (int64_t l = 0; l < lines; ++l) {
    deg[tok[2 * l] - vmin]++;
    if (reverse) deg[tok[2 * l + 1] - vmin]++;
  }
  for (int64_t v = 1; v < V; ++v) beg[v] = beg[v - 1] + deg[v - 1];
  beg[V] = E;
  std::fill(deg.begin(), deg.end(), 0);
  std::srand(1);
  for (int64_t l = 0; l < lines; ++l) {
    const int64_t s = tok[2 * l] - vmin, d = tok[2 * l + 1] - vmin;
    // both writes of a line target their slot before either degree moves
    // (a reversed self-loop fills one slot and leaves the next one 0)
    adj[beg[s] + deg[s]] = d;
    if (reverse) adj[beg[d] + deg[d]] = s;
    const int64_t wt = std::rand() % 63 + 1;
    weight[beg[s] + deg[s]] = wt;
    if (reverse) weight[beg[d] + deg[d]] = wt;
    head[beg[s] + deg[s]] = s;
    deg[s]++;
    if (reverse) deg[d]++;
  }
  const std::string pre = out_prefix ? out_prefix : path;
  int rc = write_i64_file(pre + "_beg_pos.bin", beg.data(), beg.size());
  if (!rc) rc = write_i64_file(pre + "_csr.bin", adj.data(), adj.size());
  if (!rc) rc = write_i64_file(pre + "_head.bin", head.data(), head.size());
  if (!rc) rc = write_i64_file(pre + "_weight.bin", weight.data(), weight.size());
  if (!rc) rc = write_i64_file(pre + "_deg.bin", deg.data(), deg.size());
  if (rc) return rc;
  if (vert_count) *vert_count = V;
  if (edge_count) *edge_count = E;
  return IBFS_OK;
}

// TEPS numerator of validate.h:59-76 (report()): sum of degrees of reached
// vertices.
int ibfs_traversed_edges(ibfs_hostcsr_t h, const int32_t *depth, int64_t *edges) {
  if (!h || !depth || !edges) return fail(IBFS_ERR_ARG, "ibfs_traversed_edges: null argument");
  int64_t e = 0;
  for (int64_t v = 0; v < h->csr.V; ++v)
    if (depth[v] >= 0) e += h->csr.beg[v + 1] - h->csr.beg[v];
  *edges = e;
  return IBFS_OK;
}

}  // extern "C"
