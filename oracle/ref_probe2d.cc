// Probe over the reference's diffusion/ family CPU path (2D and 3D) -- TEST
// INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile with the family's own
// sources where they lie (src/diffusion/diffusion_openmp.cc and the headers
// diffusion.h / baseline.h) and the reference's flags; output goes to
// oracle/_ref/ only.
#include <string.h>

#include <chrono>

#include "diffusion/diffusion.h"
#include "diffusion/baseline.h"
#include "diffusion/diffusion_openmp.h"

using namespace diffusion;

namespace {

template <class Variant>
class Probe : public Variant {
 public:
  Probe(int nd, const int *dims) : Variant(nd, dims) {}
  void Run(const float *init, int count, float *out, float *accuracy, float *constants,
           double *seconds) {
    this->Setup();
    this->InitializeInput();
    const size_t n = (size_t)this->nx_ * this->ny_ * this->nz_;
    if (init) memcpy(this->f1_, init, sizeof(float) * n);
    const auto t0 = std::chrono::steady_clock::now();
    this->RunKernel(count);
    if (seconds)
      *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (out) memcpy(out, this->f1_, sizeof(float) * n);
    if (accuracy) *accuracy = this->GetAccuracy(count);
    if (constants) {
      const float v[15] = {this->kappa_, this->dx_, this->dy_, this->dz_, this->kx_, this->ky_,
                           this->kz_, this->dt_, this->ce_, this->cw_, this->cn_, this->cs_,
                           this->ct_, this->cb_, this->cc_};
      memcpy(constants, v, sizeof v);
    }
    this->FinalizeBenchmark();
  }
};

}  // namespace

extern "C" {

// variant 0 = Baseline (serial), 1 = DiffusionOpenMP; nd = 2 or 3.
// seconds (optional): wall clock of RunKernel(count).
int ref2d_run(int variant, int nd, int nx, int ny, int nz, int count, const float *init, float *out,
              float *accuracy, float *constants, double *seconds) {
  const int dims[3] = {nx, ny, nz};
  if (variant == 0) {
    Probe<Baseline> p(nd, dims);
    p.Run(init, count, out, accuracy, constants, seconds);
  } else {
    Probe<DiffusionOpenMP> p(nd, dims);
    p.Run(init, count, out, accuracy, constants, seconds);
  }
  return 0;
}

}  // extern "C"
