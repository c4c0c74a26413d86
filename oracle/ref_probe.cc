// Probe over the reference's own CPU path -- TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile together with the reference's sources where
// they lie (src/diffusion3d/diffusion3d.cc, baseline.cc,
// diffusion3d_openmp.cc, diffusion3d_openmp_temporal_blocking.cc) with the
// reference's flags (common/common.mk:17, common/gcc.mk:22,27).  It exposes
// the reference classes through a C ABI so tests/golden/gen_golden.py can
// record raw float32 output bits and bench.py can time the reference OpenMP
// RunKernel as the CPU baseline.  Output goes to oracle/_ref/ only.
#include <sys/time.h>
#include <string.h>

#include "diffusion3d/diffusion3d.h"
#include "diffusion3d/baseline.h"
#include "diffusion3d/diffusion3d_openmp.h"
#include "diffusion3d/diffusion3d_openmp_temporal_blocking.h"

using namespace diffusion3d;

namespace {

// Opens the protected interface of a reference variant (diffusion3d/diffusion3d.h:91-95).
template <class Variant>
class Probe : public Variant {
 public:
  Probe(int nx, int ny, int nz) : Variant(nx, ny, nz) {}
  void Init() { this->InitializeBenchmark(); }
  void Run(int count) { this->RunKernel(count); }
  float Accuracy(int count) { return this->GetAccuracy(count); }
  void Fini() { this->FinalizeBenchmark(); }
  const float *Field() const { return this->f1_; }
  float *MutableField() { return this->f1_; }
  void Constants(float *o) const {
    o[0] = this->kappa_; o[1] = this->dx_; o[2] = this->dy_; o[3] = this->dz_;
    o[4] = this->kx_; o[5] = this->ky_; o[6] = this->kz_; o[7] = this->dt_;
    o[8] = this->ce_; o[9] = this->cw_; o[10] = this->cn_; o[11] = this->cs_;
    o[12] = this->ct_; o[13] = this->cb_; o[14] = this->cc_;
  }
  void Analytical(float *out, int count) const {
    float *f = this->GetCorrectAnswer(count);
    memcpy(out, f, sizeof(float) * (size_t)this->nx_ * this->ny_ * this->nz_);
    free(f);
  }
};

double Now() {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

template <class Variant>
int RunVariant(int nx, int ny, int nz, int count, const float *init, float *out,
               float *accuracy, float *constants, double *seconds) {
  Probe<Variant> p(nx, ny, nz);
  p.Init();
  if (init) memcpy(p.MutableField(), init, sizeof(float) * (size_t)nx * ny * nz);
  double t0 = Now();
  p.Run(count);
  double t1 = Now();
  if (seconds) *seconds = t1 - t0;
  if (out) memcpy(out, p.Field(), sizeof(float) * (size_t)nx * ny * nz);
  if (accuracy) *accuracy = p.Accuracy(count);
  if (constants) p.Constants(constants);
  p.Fini();
  return 0;
}

}  // namespace

extern "C" {

// variant: 0 = Baseline (serial), 1 = Diffusion3DOpenMP,
//          2 = Diffusion3DOpenMPTemporalBlocking.
// init: optional initial field (NULL = the reference's own Initialize(t=0)).
int ref_run(int variant, int nx, int ny, int nz, int count, const float *init,
            float *out, float *accuracy, float *constants, double *seconds) {
  switch (variant) {
    case 0: return RunVariant<Baseline>(nx, ny, nz, count, init, out, accuracy, constants, seconds);
    case 1: return RunVariant<Diffusion3DOpenMP>(nx, ny, nz, count, init, out, accuracy, constants, seconds);
    case 2: return RunVariant<Diffusion3DOpenMPTemporalBlocking>(nx, ny, nz, count, init, out, accuracy, constants, seconds);
    default: return -1;
  }
}

// The constructor's constants only (diffusion3d/diffusion3d.h:54-66).
int ref_constants(int nx, int ny, int nz, float *constants) {
  Probe<Baseline> p(nx, ny, nz);
  p.Constants(constants);
  return 0;
}

// The reference's analytical field at t = count * dt (GetCorrectAnswer).
int ref_analytical(int nx, int ny, int nz, int count, float *out) {
  Probe<Baseline> p(nx, ny, nz);
  p.Analytical(out, count);
  return 0;
}

}  // extern "C"
