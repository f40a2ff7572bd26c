// gpu-ibfs: command line front end with the reference's five positional
// arguments (main.cu:41-50) over the C ABI:
//   gpu-ibfs <beg_pos.bin> <csr.bin> <concurrent-count> <switch-level> <bfs_count>
// Loads the CSR, draws bfs_count seeded sources of degree > 0, orders them
// with GroupBy, and runs them as groups of concurrent-count sources with the
// fixed switch level (top-down while level <= switch-level), printing the
// reference's per-group and total times.  Depth arrays stay on the device.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "../../include/ibfs.h"

static double wtime() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int die(const char *what) {
  char err[1024];
  ibfs_last_error(err, sizeof err);
  std::cout << what << ": " << err << "\n";
  return -1;
}

int main(int argc, char *argv[]) {
  std::cout << "Input: /path/to/exe /path/to/beg /path/to/csr concurrent-count "
            << "switch-level bfs_count\n";
  if (argc != 6) {
    std::cout << "Wrong input\n";
    return -1;
  }
  const char *beg_file = argv[1], *csr_file = argv[2];
  const int concurr = std::atoi(argv[3]);
  const int sw_level = std::atoi(argv[4]);
  const long src_count = std::atol(argv[5]);
  if (concurr != 64 && concurr != 128 && concurr != 256) {
    std::cout << "concurrent-count must be 64, 128 or 256\n";
    return -1;
  }
  double tm = wtime();
  ibfs_hostcsr_t h = nullptr;
  if (ibfs_hostcsr_load(beg_file, csr_file, &h)) return die("graph load");
  int64_t V = 0, E = 0;
  ibfs_hostcsr_info(h, &V, &E);
  std::cout << "Graph load (success): " << V << " verts, " << E << " edges " << wtime() - tm
            << " second(s)\n";
  std::vector<int32_t> src(src_count), grouped(src_count);
  if (ibfs_gen_sources(h, src_count, 2, src.data())) return die("gen_src");
  ibfs_graph_t g = nullptr;
  if (ibfs_graph_from_host(h, 0, &g)) return die("graph upload");
  if (ibfs_groupby(g, src.data(), src_count, grouped.data())) return die("groupby");
  ibfs_config cfg;
  ibfs_config_default(&cfg);
  cfg.lane_bits = concurr;
  cfg.policy = IBFS_POLICY_FIXED;
  cfg.sw_level = sw_level;
  ibfs_engine_t e = nullptr;
  if (ibfs_engine_create(g, &cfg, &e)) return die("engine");
  int32_t *depth = nullptr;
  if (ibfs_device_alloc(0, sizeof(int32_t) * (size_t)concurr * (size_t)V, (void **)&depth))
    return die("device alloc");
  std::cout << "Start bfs ...\n";
  const double t_all = wtime();
  int64_t edges = 0;
  for (long i = 0; i < src_count / concurr; ++i) {
    ibfs_stats st;
    const double t = wtime();
    if (ibfs_traversal(e, grouped.data() + i * concurr, concurr, depth, IBFS_OUT_DEVICE, &st))
      return die("traversal");
    std::cout << "Traversal-time: " << wtime() - t << " second(s)\n";
    edges += st.traversed_edges;
  }
  const double total = wtime() - t_all;
  std::cout << "Total time: " << total << "\n";
  std::cout << "Traversed edges: " << edges << " GTEPS: " << (double)edges / total / 1e9 << "\n";
  ibfs_device_free(depth);
  ibfs_engine_destroy(e);
  ibfs_graph_destroy(g);
  ibfs_hostcsr_free(h);
  return 0;
}
