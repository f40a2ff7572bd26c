// text-to-bin: the converter's command line (tuple_text_to_bin.cpp:61-67)
// over the C ABI:
//   text-to-bin <tuple_file> <reverse 1/0> <lines_to_skip>
// Writes <tuple_file>_beg_pos.bin, _csr.bin, _head.bin, _weight.bin and
// _deg.bin (little-endian int64) with the converter's rules and prints the
// counts and the first vertices' adjacency, as the reference does.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../../include/ibfs.h"

int main(int argc, char *argv[]) {
  std::cout << "Input: ./exe tuple_file(text) reverse_the_edge(1 reverse, 0 not reverse) "
            << "lines_to_skip\n";
  if (argc < 4) {
    std::cout << "Wrong input\n";
    return -1;
  }
  const int reverse = std::atol(argv[2]) == 1;
  const long skip = std::atol(argv[3]);
  int64_t V = 0, E = 0;
  if (ibfs_convert_text_files(argv[1], reverse, skip, nullptr, &V, &E)) {
    char err[1024];
    ibfs_last_error(err, sizeof err);
    std::cout << err << "\n";
    return -1;
  }
  std::cout << "edge count: " << E << "\n"
            << "vert count: " << V << "\n";
  const std::string pre = argv[1];
  ibfs_hostcsr_t h = nullptr;
  if (ibfs_hostcsr_load((pre + "_beg_pos.bin").c_str(), (pre + "_csr.bin").c_str(), &h) == 0) {
    const int64_t *beg = nullptr;
    const int32_t *adj = nullptr;
    ibfs_hostcsr_data(h, &beg, &adj);
    for (int64_t i = 0; i < (V < 8 ? V : 8); ++i) {
      std::cout << i << " " << beg[i + 1] - beg[i] << " ";
      for (int64_t j = beg[i]; j < beg[i + 1]; ++j) std::cout << adj[j] << " ";
      std::cout << "\n";
    }
    ibfs_hostcsr_free(h);
  }
  return 0;
}
