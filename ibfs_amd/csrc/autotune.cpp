// Autotuner: coordinate descent over the engine configuration, scored by
// aggregate GTEPS of the caller's sample sources on the caller's graph.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

using namespace ibfs;

namespace {

struct Param {
  const char *name;
  int32_t ibfs_config::*field;
  std::vector<int32_t> values;
};

struct Trial {
  ibfs_config cfg;
  double gteps;
  double ms;
};

// GTEPS of cfg on the sample (best of `reps` runs after one warm-up).
int evaluate(ibfs_graph_t g, const ibfs_config &cfg, const int32_t *src, int64_t n, int32_t *dout,
             int reps, double *gteps, double *ms) {
  ibfs_engine_t e = nullptr;
  int rc = ibfs_engine_create(g, &cfg, &e);
  if (rc) return rc;
  ibfs_stats st{};
  rc = ibfs_run_groups(e, src, n, dout, IBFS_OUT_DEVICE, &st);
  double best = 1e300;
  int64_t edges = st.traversed_edges;
  for (int r = 0; r < reps && rc == IBFS_OK; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    rc = ibfs_run_groups(e, src, n, dout, IBFS_OUT_DEVICE, &st);
    const double t =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    best = std::min(best, t);
    edges = st.traversed_edges;
  }
  ibfs_engine_destroy(e);
  if (rc) return rc;
  *ms = best;
  *gteps = (double)edges / (best * 1e-3) / 1e9;
  return IBFS_OK;
}

void append_cfg(std::string &s, const ibfs_config &c) {
  char buf[512];
  std::snprintf(buf, sizeof buf,
                "{\"lane_bits\": %d, \"sml_shed\": %d, \"lrg_shed\": %d, \"policy\": %d, "
                "\"sw_level\": %d, \"alpha\": %d, \"beta\": %d, \"td_chunk\": %d, "
                "\"emit_mode\": %d}",
                c.lane_bits, c.sml_shed, c.lrg_shed, c.policy, c.sw_level, c.alpha, c.beta,
                c.td_chunk, c.emit_mode);
  s += buf;
}

}  // namespace

extern "C" {

int ibfs_autotune(ibfs_graph_t g, const int32_t *src, int64_t n, const ibfs_config *start,
                  int32_t passes, int32_t reps, ibfs_config *best, char *report,
                  size_t report_len) {
  if (!g || !src || n < 1 || !best || passes < 1 || reps < 1)
    return fail(IBFS_ERR_ARG, "ibfs_autotune: bad arguments");
  ibfs_config cur;
  if (start) cur = *start;
  else ibfs_config_default(&cur);
  cur.kernel_timing = 0;
  cur.work_model = 0;
  cur.policy = IBFS_POLICY_ALPHA_BETA;
  int64_t V = 0;
  int rc = ibfs_graph_info(g, &V, nullptr);
  if (rc) return rc;
  int32_t *dout = nullptr;
  rc = ibfs_device_alloc(cur.device, sizeof(int32_t) * (size_t)n * (size_t)V, (void **)&dout);
  if (rc) return rc;
  const std::vector<Param> params = {
      {"lane_bits", &ibfs_config::lane_bits, {64, 128, 256, 512}},
      {"alpha", &ibfs_config::alpha, {1, 2, 3, 4, 6, 8, 14}},
      {"beta", &ibfs_config::beta, {4, 8, 16, 24, 48, 96}},
      {"sml_shed", &ibfs_config::sml_shed, {8, 16, 32, 64}},
      {"lrg_shed", &ibfs_config::lrg_shed, {128, 256, 512, 1024, 4096}},
      {"td_chunk", &ibfs_config::td_chunk, {1024, 2048, 4096, 8192}},
      {"emit_mode", &ibfs_config::emit_mode, {0, 1, 2}},
      {"inspect_vpt", &ibfs_config::inspect_vpt, {0, 4, 8, 16}},
  };
  std::vector<Trial> trials;
  double best_g = 0, best_ms = 0;
  rc = evaluate(g, cur, src, n, dout, reps, &best_g, &best_ms);
  if (rc) {
    ibfs_device_free(dout);
    return rc;
  }
  trials.push_back({cur, best_g, best_ms});
  for (int p = 0; p < passes; ++p) {
    bool improved = false;
    for (const auto &par : params) {
      for (int32_t v : par.values) {
        if (cur.*par.field == v) continue;
        ibfs_config c = cur;
        c.*par.field = v;
        if (c.lrg_shed < c.sml_shed) continue;
        double gt = 0, ms = 0;
        rc = evaluate(g, c, src, n, dout, reps, &gt, &ms);
        if (rc) {
          ibfs_device_free(dout);
          return rc;
        }
        trials.push_back({c, gt, ms});
        if (gt > best_g * 1.01) {  // keep a change only for a clear gain
          best_g = gt;
          best_ms = ms;
          cur = c;
          improved = true;
        }
      }
    }
    if (!improved) break;
  }
  ibfs_device_free(dout);
  *best = cur;
  if (report && report_len) {
    std::string s = "{\"best\": ";
    append_cfg(s, cur);
    char buf[128];
    std::snprintf(buf, sizeof buf, ", \"best_gteps\": %.3f, \"best_ms\": %.4f, \"trials\": [",
                  best_g, best_ms);
    s += buf;
    for (size_t i = 0; i < trials.size(); ++i) {
      if (i) s += ", ";
      s += "{\"cfg\": ";
      append_cfg(s, trials[i].cfg);
      std::snprintf(buf, sizeof buf, ", \"gteps\": %.3f, \"ms\": %.4f}", trials[i].gteps,
                    trials[i].ms);
      s += buf;
    }
    s += "]}";
    std::snprintf(report, report_len, "%s", s.c_str());
    if (s.size() >= report_len) return fail(IBFS_ERR_ARG, "report buffer too small (%zu bytes needed)", s.size() + 1);
  }
  return IBFS_OK;
}

}  // extern "C"
