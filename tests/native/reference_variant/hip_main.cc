// The reference driver's flow with the HIP variant selected: option parsing
// as diffusion3d/main.cc:84-127 (--count N, --size N, --dump, --help; other
// options ignored) and the `bmk = new Diffusion3DHIP(nx, nx, nx, argc, argv)`
// branch a maintainer adds beside main.cc:163 (the Physis branch), then
// RunBenchmark(count, dump) (main.cc:181).  Built against the reference's
// own headers and its diffusion3d.cc / baseline.cc, and linked to the engine
// library, by oracle/Makefile (target `ref`) and tests/test_reference_binding.py.
#include <getopt.h>
#include <stdlib.h>

#include <iostream>

#include "diffusion3d/diffusion3d_hip.h"

using namespace diffusion3d;

int main(int argc, char *argv[]) {
  int nx = 256;
  int count = 100;
  bool dump = false;
  static struct option long_options[] = {
      {"count", 1, 0, 0}, {"size", 1, 0, 0}, {"dump", 0, 0, 0}, {"help", 0, 0, 0}, {0, 0, 0, 0}};
  while (1) {
    int option_index = 0;
    const int c = getopt_long(argc, argv, "", long_options, &option_index);
    if (c == -1) break;
    if (c != 0) continue;
    switch (option_index) {
      case 0: count = atoi(optarg); break;
      case 1: nx = atoi(optarg); break;
      case 2: dump = true; break;
      case 3:
        std::cerr << "Usage: " << argv[0] << " [--count N] [--size N] [--dump] [--help]\n";
        return EXIT_SUCCESS;
      default: break;
    }
  }
  Diffusion3D *bmk = new Diffusion3DHIP(nx, nx, nx, argc, argv);
  bmk->RunBenchmark(count, dump);
  delete bmk;
  return 0;
}
