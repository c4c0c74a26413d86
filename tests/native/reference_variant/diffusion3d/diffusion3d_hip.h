// The HIP engine as one more variant of the reference's diffusion3d
// benchmark: the file a maintainer adds as src/diffusion3d/diffusion3d_hip.h
// of naoyam/benchmarks.  It follows the Physis variant's pattern
// (diffusion3d/diffusion3d_physis.h:21-43): a Baseline subclass whose
// InitializeBenchmark / RunKernel / FinalizeBenchmark call the engine's C ABI,
// plus the CUDA variant's "Kernel-only performance" block
// (diffusion3d/diffusion3d_cuda.cu:373-384).
//
// The engine's declarations are included with angle brackets: this file has
// the same name as the engine header, and a quoted include would find this
// file again in its own directory.
#ifndef DIFFUSION3D_DIFFUSION3D_HIP_H_
#define DIFFUSION3D_DIFFUSION3D_HIP_H_

#include <stdio.h>
#include <stdlib.h>

#include "diffusion3d/diffusion3d.h"
#include "diffusion3d/baseline.h"
#include <diffusion3d_hip.h>  // the engine's include/ directory

namespace diffusion3d {

inline void HipCheck(int rc, const char *what) {
  if (rc != D3H_OK) {
    fprintf(stderr, "%s: %s (%s)\n", what, d3h_status_string(rc), d3h_last_error());
    exit(EXIT_FAILURE);
  }
}

class Diffusion3DHIP: public Baseline {
 public:
  Diffusion3DHIP(int nx, int ny, int nz, int argc, char **argv):
      Baseline(nx, ny, nz) {
    HipCheck(initialize_hip(argc, argv, nx, ny, nz), "initialize_hip");
  }
  virtual std::string GetName() const { return std::string("hip"); }
  virtual void InitializeBenchmark() {
    Baseline::InitializeBenchmark();
    HipCheck(initialize_benchmark_hip(nx_, ny_, nz_), "initialize_benchmark_hip");
  }
  virtual void FinalizeBenchmark() {
    HipCheck(finalize_benchmark_hip(), "finalize_benchmark_hip");
    Baseline::FinalizeBenchmark();
  }
  virtual void RunKernel(int count) {
    HipCheck(run_kernel_hip(count, f1_, nx_, ny_, nz_,
                            ce_, cw_, cn_, cs_, ct_, cb_, cc_), "run_kernel_hip");
  }
  virtual void DisplayResult(int count, float time) {
    Baseline::DisplayResult(count, time);
    float t = 0;
    HipCheck(kernel_time_hip(&t), "kernel_time_hip");
    printf("Kernel-only performance:\n");
    printf("Elapsed time : %.3f (s)\n", t);
    printf("FLOPS        : %.3f (GFLOPS)\n", GetGFLOPS(count, t));
    printf("Throughput   : %.3f (GB/s)\n", GetThroughput(count, t));
  }
};

}

#endif /* DIFFUSION3D_DIFFUSION3D_HIP_H_ */
