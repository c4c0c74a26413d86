// diffusion3d_hip: the reference's benchmark driver (diffusion3d/main.cc:84-184)
// with the HIP engine as the variant.  Same options (--count N, --size N,
// --dump, --help; unknown options are ignored like the reference's
// getopt loop), plus --warmup and --nd N (diffusion/main.cc:86-129; --nd 2
// runs the diffusion/ family's 2D benchmark) and --nx/--ny/--nz for
// non-cubic grids.  Engine knobs come from the environment:
// D3H_TEMPORAL_DEPTH (1, 2, 4, 8; 0 = tuned), D3H_ARITH (0 exact, 1 fma),
// D3H_USE_GRAPH (0/1), D3H_DEVICE, D3H_TUNE_FILE, D3H_AUTOTUNE.
#include <getopt.h>
#include <stdlib.h>

#include <iostream>

#include "diffusion2d.hpp"
#include "diffusion3d.hpp"

using namespace diffusion3d;

static void PrintUsage(std::ostream &os, const char *prog) {
  os << "Usage: " << prog << " [options] [benchmarks]\n\n";
  os << "Options\n"
     << "\t--nd N      Number of dimensions, 2 or 3 (default: 3)\n"
     << "\t--count N   Number of iterations\n"
     << "\t--size N    Size of each dimension\n"
     << "\t--nx N      Size of x (overrides --size)\n"
     << "\t--ny N      Size of y (overrides --size)\n"
     << "\t--nz N      Size of z (overrides --size)\n"
     << "\t--dump      Dump the final data to file\n"
     << "\t--dump-raw  Also write the final field bit-exact (d3h_raw_write format)\n"
     << "\t--warmup    Enable warming-up runs\n"
     << "\t--help      Display this help message\n";
}

int main(int argc, char *argv[]) {
  int size = 256, count = 100, nx = 0, ny = 0, nz = 0, nd = 3;
  bool dump = false, warmup = false, dump_raw = false;
  static struct option long_options[] = {
      {"count", 1, 0, 0}, {"size", 1, 0, 0}, {"dump", 0, 0, 0}, {"help", 0, 0, 0},
      {"warmup", 0, 0, 0}, {"nx", 1, 0, 0}, {"ny", 1, 0, 0},   {"nz", 1, 0, 0},
      {"nd", 1, 0, 0},     {"dump-raw", 0, 0, 0}, {0, 0, 0, 0}};
  while (1) {
    int option_index = 0;
    int c = getopt_long(argc, argv, "", long_options, &option_index);
    if (c == -1) break;
    if (c != 0) continue;
    switch (option_index) {
      case 0: count = atoi(optarg); break;
      case 1: size = atoi(optarg); break;
      case 2: dump = true; break;
      case 3: PrintUsage(std::cerr, argv[0]); return EXIT_SUCCESS;
      case 4: warmup = true; break;
      case 5: nx = atoi(optarg); break;
      case 6: ny = atoi(optarg); break;
      case 7: nz = atoi(optarg); break;
      case 8: nd = atoi(optarg); break;
      case 9: dump_raw = true; break;
      default: break;
    }
  }
  if (nx <= 0) nx = size;
  if (ny <= 0) ny = size;
  if (nz <= 0) nz = size;
  if (nx <= 0 || ny <= 0 || nz <= 0 || count < 0 || (nd != 2 && nd != 3)) {
    std::cerr << "invalid size, count or number of dimensions\n";
    return EXIT_FAILURE;
  }
  if (nd == 2) {
    diffusion::Diffusion2DHIP bmk(nx, ny);
    bmk.SetRawDump(dump_raw);
    bmk.RunBenchmark(count, dump, warmup);
    return 0;
  }
  Diffusion3D *bmk = new Diffusion3DHIP(nx, ny, nz, argc, argv);
  bmk->SetRawDump(dump_raw);
  bmk->RunBenchmark(count, dump, warmup);
  delete bmk;
  return 0;
}
