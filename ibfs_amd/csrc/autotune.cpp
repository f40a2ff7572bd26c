// Autotuner: coordinate descent over the traversal configuration, scored by
// aggregate GTEPS of the caller's sample sources on the caller's graph,
// either through the engine (groups one after another) or through the
// batched launcher (which adds its grid, context count and burst length).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

using namespace ibfs;

namespace {

// A tuned setting: the configuration plus the batched launcher's own knobs.
struct Setting {
  ibfs_config cfg;
  int32_t contexts = 2;
  int32_t burst = 8;
};

struct Param {
  const char *name;
  int32_t ibfs_config::*field;  // null: `slot` below
  int32_t Setting::*slot;
  std::vector<int32_t> values;
};

int32_t &ref(Setting &s, const Param &p) { return p.field ? s.cfg.*p.field : s.*p.slot; }

struct Trial {
  Setting set;
  double gteps;
  double ms;
};

// GTEPS of a setting on the sample (best of `reps` runs after one warm-up).
int evaluate(ibfs_graph_t g, const Setting &set, bool batch, const int32_t *src, int64_t n,
             int32_t *dout, int reps, double *gteps, double *ms) {
  ibfs_engine_t e = nullptr;
  ibfs_batch_t b = nullptr;
  int rc = batch ? ibfs_batch_create(g, &set.cfg, set.contexts, set.burst, &b)
                 : ibfs_engine_create(g, &set.cfg, &e);
  if (rc) return rc;
  auto run = [&](ibfs_stats *st) {
    return batch ? ibfs_batch_run(b, src, n, dout, IBFS_OUT_DEVICE, st)
                 : ibfs_run_groups(e, src, n, dout, IBFS_OUT_DEVICE, st);
  };
  ibfs_stats st{};
  rc = run(&st);
  double best = 1e300;
  int64_t edges = st.traversed_edges;
  for (int r = 0; r < reps && rc == IBFS_OK; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    rc = run(&st);
    const double t =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    best = std::min(best, t);
    edges = st.traversed_edges;
  }
  if (batch) ibfs_batch_destroy(b);
  else ibfs_engine_destroy(e);
  if (rc) return rc;
  *ms = best;
  *gteps = (double)edges / (best * 1e-3) / 1e9;
  return IBFS_OK;
}

void append_setting(std::string &s, const Setting &x, bool batch) {
  const ibfs_config &c = x.cfg;
  char buf[640];
  std::snprintf(buf, sizeof buf,
                "{\"lane_bits\": %d, \"sml_shed\": %d, \"lrg_shed\": %d, \"policy\": %d, "
                "\"sw_level\": %d, \"alpha\": %d, \"beta\": %d, \"td_chunk\": %d, "
                "\"emit_mode\": %d, \"inspect_vpt\": %d, \"wg_size\": %d, \"hybrid_gamma\": %d, "
                "\"hybrid_hubs\": %d",
                c.lane_bits, c.sml_shed, c.lrg_shed, c.policy, c.sw_level, c.alpha, c.beta,
                c.td_chunk, c.emit_mode, c.inspect_vpt, c.wg_size, c.hybrid_gamma, c.hybrid_hubs);
  s += buf;
  if (batch) {
    std::snprintf(buf, sizeof buf,
                  ", \"dd_grid\": %d, \"pack_sources\": %d, \"phase_offsets\": %d, "
                  "\"tail_div\": %d, \"contexts\": %d, \"burst\": %d",
                  c.dd_grid, c.pack_sources, c.phase_offsets, c.tail_div, x.contexts, x.burst);
    s += buf;
  }
  s += "}";
}

int tune(ibfs_graph_t g, const int32_t *src, int64_t n, const ibfs_config *start, bool batch,
         int32_t passes, int32_t reps, Setting *out, char *report, size_t report_len) {
  if (!g || !src || n < 1 || !out || passes < 1 || reps < 1)
    return fail(IBFS_ERR_ARG, "ibfs_autotune: bad arguments");
  Setting cur;
  if (start) cur.cfg = *start;
  else ibfs_config_default(&cur.cfg);
  cur.contexts = out->contexts;
  cur.burst = out->burst;
  cur.cfg.kernel_timing = 0;
  cur.cfg.work_model = 0;
  cur.cfg.policy = IBFS_POLICY_ALPHA_BETA;
  int64_t V = 0;
  int rc = ibfs_graph_info(g, &V, nullptr);
  if (rc) return rc;
  int32_t *dout = nullptr;
  rc = ibfs_device_alloc(cur.cfg.device, sizeof(int32_t) * (size_t)n * (size_t)V, (void **)&dout);
  if (rc) return rc;
  std::vector<Param> params = {
      {"lane_bits", &ibfs_config::lane_bits, nullptr, {64, 128, 256, 512}},
      {"alpha", &ibfs_config::alpha, nullptr, {1, 2, 3, 4, 6, 8, 14}},
      {"beta", &ibfs_config::beta, nullptr, {4, 8, 16, 24, 48, 96}},
      {"sml_shed", &ibfs_config::sml_shed, nullptr, {8, 16, 32, 64}},
      {"lrg_shed", &ibfs_config::lrg_shed, nullptr, {128, 256, 512, 1024, 4096}},
      {"td_chunk", &ibfs_config::td_chunk, nullptr, {1024, 2048, 4096, 8192}},
      {"emit_mode", &ibfs_config::emit_mode, nullptr, {0, 1, 2, 3}},
      {"inspect_vpt", &ibfs_config::inspect_vpt, nullptr, {0, 1, 2, 4, 8}},
      {"wg_size", &ibfs_config::wg_size, nullptr, {256, 512, 1024}},
      {"hybrid_gamma", &ibfs_config::hybrid_gamma, nullptr, {0, 8, 16, 24, 32}},
      {"hybrid_hubs", &ibfs_config::hybrid_hubs, nullptr, {1024, 4096, 16384, 0}},
  };
  if (batch) {
    params.push_back({"dd_grid", &ibfs_config::dd_grid, nullptr, {0, 2048, 4096, 16384}});
    params.push_back({"pack_sources", &ibfs_config::pack_sources, nullptr, {0, 1}});
    params.push_back({"phase_offsets", &ibfs_config::phase_offsets, nullptr, {0, 1, 2, 3}});
    params.push_back({"tail_div", &ibfs_config::tail_div, nullptr, {0, 8, 32, 256}});
    params.push_back({"contexts", nullptr, &Setting::contexts, {1, 2, 3, 4}});
    params.push_back({"burst", nullptr, &Setting::burst, {4, 8, 16}});
  }
  std::vector<Trial> trials;
  double best_g = 0, best_ms = 0;
  rc = evaluate(g, cur, batch, src, n, dout, reps, &best_g, &best_ms);
  if (rc) {
    ibfs_device_free(dout);
    return rc;
  }
  trials.push_back({cur, best_g, best_ms});
  for (int p = 0; p < passes; ++p) {
    bool improved = false;
    for (const auto &par : params) {
      for (int32_t v : par.values) {
        if (ref(cur, par) == v) continue;
        Setting c = cur;
        ref(c, par) = v;
        if (c.cfg.lrg_shed < c.cfg.sml_shed) continue;
        double gt = 0, ms = 0;
        rc = evaluate(g, c, batch, src, n, dout, reps, &gt, &ms);
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
  *out = cur;
  if (report && report_len) {
    std::string s = "{\"mode\": ";
    s += batch ? "\"batch\"" : "\"engine\"";
    s += ", \"best\": ";
    append_setting(s, cur, batch);
    char buf[128];
    std::snprintf(buf, sizeof buf, ", \"best_gteps\": %.3f, \"best_ms\": %.4f, \"trials\": [",
                  best_g, best_ms);
    s += buf;
    for (size_t i = 0; i < trials.size(); ++i) {
      if (i) s += ", ";
      s += "{\"cfg\": ";
      append_setting(s, trials[i].set, batch);
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

}  // namespace

extern "C" {

int ibfs_autotune(ibfs_graph_t g, const int32_t *src, int64_t n, const ibfs_config *start,
                  int32_t passes, int32_t reps, ibfs_config *best, char *report,
                  size_t report_len) {
  if (!best) return fail(IBFS_ERR_ARG, "ibfs_autotune: null result");
  Setting out;
  int rc = tune(g, src, n, start, false, passes, reps, &out, report, report_len);
  if (rc) return rc;
  *best = out.cfg;
  return IBFS_OK;
}

int ibfs_autotune_batch(ibfs_graph_t g, const int32_t *src, int64_t n, const ibfs_config *start,
                        int32_t contexts, int32_t burst, int32_t passes, int32_t reps,
                        ibfs_config *best, int32_t *best_contexts, int32_t *best_burst,
                        char *report, size_t report_len) {
  if (!best || !best_contexts || !best_burst)
    return fail(IBFS_ERR_ARG, "ibfs_autotune_batch: null result");
  if (contexts < 1 || contexts > 16 || burst < 1 || burst > 64)
    return fail(IBFS_ERR_ARG, "ibfs_autotune_batch: contexts in [1,16], burst in [1,64]");
  Setting out;
  out.contexts = contexts;
  out.burst = burst;
  int rc = tune(g, src, n, start, true, passes, reps, &out, report, report_len);
  if (rc) return rc;
  *best = out.cfg;
  *best_contexts = out.contexts;
  *best_burst = out.burst;
  return IBFS_OK;
}

}  // extern "C"
