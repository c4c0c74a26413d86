// Host-side benchmark classes over the HIP engine's C ABI.
//
// Mirrors the reference's variant interface so the HIP engine drops in as one
// more variant:
//   Diffusion3D   : constants, RunBenchmark, metrics, DisplayResult,
//                   GetCorrectAnswer          (diffusion3d/diffusion3d.h:43-121)
//   Baseline      : host buffers f1_/f2_, GetAccuracy, Dump
//                                             (diffusion3d/baseline.{h,cc})
//   Diffusion3DHIP: InitializeBenchmark / RunKernel / FinalizeBenchmark over
//                   initialize_hip / initialize_benchmark_hip / run_kernel_hip /
//                   finalize_benchmark_hip, the same shape as the Physis
//                   variant (diffusion3d/diffusion3d_physis.h:21-43), plus the
//                   "Kernel-only performance" block of the CUDA variant
//                   (diffusion3d/diffusion3d_cuda.cu:373-384).
// Errors from the ABI are reported with the library's message and end the
// program with EXIT_FAILURE (the reference asserts instead).
#pragma once

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>

#include <iostream>
#include <string>

#include "diffusion3d_hip.h"

#define REAL float

namespace diffusion3d {

inline void CheckStatus(int rc, const char *what) {
  if (rc != D3H_OK) {
    fprintf(stderr, "%s failed: %s (%s)\n", what, d3h_status_string(rc), d3h_last_error());
    exit(EXIT_FAILURE);
  }
}

// Wall clock around RunKernel (common/stopwatch.h:25-45).
struct Stopwatch {
  struct timeval tv;
};
inline void StopwatchStart(Stopwatch *w) { gettimeofday(&w->tv, NULL); }
inline float StopwatchStop(Stopwatch *w) {
  struct timeval now;
  gettimeofday(&now, NULL);
  return (now.tv_sec - w->tv.tv_sec) + (now.tv_usec - w->tv.tv_usec) * 1.0e-06;
}

// diffusion3d::Initialize (diffusion3d/diffusion3d.cc:14-39), computed by the
// engine library's host code with the same promotions.
inline void Initialize(REAL *buff, int nx, int ny, int nz, const float constants[15], REAL time) {
  CheckStatus(d3h_initialize_f32(buff, nx, ny, nz, constants, time), "d3h_initialize_f32");
}

class Diffusion3D {
 protected:
  int nx_, ny_, nz_;
  REAL kappa_;
  REAL dx_, dy_, dz_;
  REAL kx_, ky_, kz_;
  REAL dt_;
  REAL ce_, cw_, cn_, cs_, ct_, cb_, cc_;
  float constants_[15];

 public:
  Diffusion3D(int nx, int ny, int nz) : nx_(nx), ny_(ny), nz_(nz) {
    CheckStatus(d3h_constants_f32(nx, ny, nz, constants_), "d3h_constants_f32");
    kappa_ = constants_[0];
    dx_ = constants_[1]; dy_ = constants_[2]; dz_ = constants_[3];
    kx_ = constants_[4]; ky_ = constants_[5]; kz_ = constants_[6];
    dt_ = constants_[7];
    ce_ = constants_[8]; cw_ = constants_[9]; cn_ = constants_[10]; cs_ = constants_[11];
    ct_ = constants_[12]; cb_ = constants_[13]; cc_ = constants_[14];
  }
  virtual ~Diffusion3D() {}
  virtual std::string GetName() const = 0;
  // Also write the final field bit-exact (d3h_raw_write format) after Dump.
  void SetRawDump(bool on) { raw_dump_ = on; }

  // The reference's RunBenchmark (diffusion3d/diffusion3d.h:69-84) with the
  // warm-up of the diffusion/ family (diffusion/diffusion.h:126-131): run,
  // then re-initialise the input.
  void RunBenchmark(int count, bool dump, bool warmup = false) {
    std::cout << "Initializing benchmark input...\n";
    InitializeBenchmark();
    std::cout << "Running diffusion3d/" << GetName() << "\n";
    std::cout << "Iteration count: " << count << "\n";
    std::cout << "Grid size: " << nx_ << "x" << ny_ << "x" << nz_ << "\n";
    if (warmup) {
      std::cout << "Warming up the kernel...\n";
      WarmingUp();
      std::cout << "Reinitializing benchmark input...\n";
      ReinitializeInput();
    }
    Stopwatch st;
    StopwatchStart(&st);
    RunKernel(count);
    float elapsed_time = StopwatchStop(&st);
    std::cout << "Benchmarking finished.\n";
    DisplayResult(count, elapsed_time);
    if (dump) Dump();
    if (raw_dump_) DumpRaw(count);
    FinalizeBenchmark();
  }

 protected:
  bool raw_dump_ = false;
  std::string GetDumpPath() const { return std::string("diffusion3d_result.") + GetName() + ".out"; }
  virtual void DumpRaw(int count) const = 0;
  virtual void InitializeBenchmark() = 0;
  virtual void ReinitializeInput() = 0;
  virtual void WarmingUp() { RunKernel(1); }
  virtual void RunKernel(int count) = 0;
  virtual void Dump() const = 0;
  virtual REAL GetAccuracy(int count) = 0;
  virtual void FinalizeBenchmark() = 0;

  // GB/s with one read and one write per cell per step (diffusion3d/diffusion3d.h:97-100).
  float GetThroughput(int count, float time) {
    return (double)nx_ * ny_ * nz_ * sizeof(REAL) * 2.0 * ((float)count) / time * 1.0e-09;
  }
  // 13 flops per cell per step (diffusion3d/diffusion3d.h:101-104).
  float GetGFLOPS(int count, float time) {
    return (double)nx_ * ny_ * nz_ * 13.0 * (float)(count) / time * 1.0e-09;
  }
  virtual void DisplayResult(int count, float time) {
    printf("Elapsed time : %.3f (s)\n", time);
    printf("FLOPS        : %.3f (GFLOPS)\n", GetGFLOPS(count, time));
    printf("Throughput   : %.3f (GB/s)\n", GetThroughput(count, time));
    printf("Accuracy     : %e\n", GetAccuracy(count));
  }
  // Analytical field at t = count * dt (diffusion3d/diffusion3d.h:113-120).
  REAL *GetCorrectAnswer(int count) const {
    REAL *f = (REAL *)malloc(sizeof(REAL) * (size_t)nx_ * ny_ * nz_);
    if (!f) {
      fprintf(stderr, "out of host memory\n");
      exit(EXIT_FAILURE);
    }
    Initialize(f, nx_, ny_, nz_, constants_, count * dt_);
    return f;
  }
};

class Baseline : public Diffusion3D {
 protected:
  REAL *f1_ = nullptr;
  REAL *f2_ = nullptr;

 public:
  Baseline(int nx, int ny, int nz) : Diffusion3D(nx, ny, nz) {}

 protected:
  // RMS vs the analytical field with a float accumulator in serial order
  // (diffusion3d/baseline.cc:51-60); the answer buffer is freed here.
  REAL GetAccuracy(int count) override {
    REAL *ref = GetCorrectAnswer(count);
    REAL err = 0.0;
    long len = (long)nx_ * ny_ * nz_;
    for (long i = 0; i < len; i++) {
      REAL diff = ref[i] - f1_[i];
      err += diff * diff;
    }
    free(ref);
    return (REAL)sqrtf(err / len);
  }
  // "%.7E\n" per cell (diffusion3d/baseline.cc:62-70).
  // diffusion3d_result.<name>.raw: the host field bit-exact.
  void DumpRaw(int count) const override {
    const std::string path = std::string("diffusion3d_result.") + GetName() + ".raw";
    CheckStatus(d3h_raw_write(path.c_str(), f1_, D3H_F32, nx_, ny_, nz_, (uint64_t)count, 0),
                "d3h_raw_write");
  }
  void Dump() const override {
    FILE *out = fopen(GetDumpPath().c_str(), "w");
    if (!out) {
      fprintf(stderr, "cannot open %s\n", GetDumpPath().c_str());
      exit(EXIT_FAILURE);
    }
    long nitems = (long)nx_ * ny_ * nz_;
    for (long i = 0; i < nitems; ++i) fprintf(out, "%.7E\n", f1_[i]);
    fclose(out);
  }
};

class Diffusion3DHIP : public Baseline {
 public:
  Diffusion3DHIP(int nx, int ny, int nz, int argc = 0, char **argv = nullptr) : Baseline(nx, ny, nz) {
    CheckStatus(initialize_hip(argc, argv, nx, ny, nz), "initialize_hip");
  }
  std::string GetName() const override { return "hip"; }

 protected:
  // Pinned host field initialised on the host, as the CUDA variant does
  // (diffusion3d/diffusion3d_cuda.cu:329-341), then the device field.
  void InitializeBenchmark() override {
    void *p = nullptr;
    CheckStatus(d3h_host_alloc(sizeof(REAL) * (size_t)nx_ * ny_ * nz_, &p), "d3h_host_alloc");
    f1_ = (REAL *)p;
    Initialize(f1_, nx_, ny_, nz_, constants_, 0.0f);
    CheckStatus(initialize_benchmark_hip(nx_, ny_, nz_), "initialize_benchmark_hip");
  }
  void ReinitializeInput() override { Initialize(f1_, nx_, ny_, nz_, constants_, 0.0f); }
  void RunKernel(int count) override {
    CheckStatus(run_kernel_hip(count, f1_, nx_, ny_, nz_, ce_, cw_, cn_, cs_, ct_, cb_, cc_),
                "run_kernel_hip");
  }
  void FinalizeBenchmark() override {
    CheckStatus(finalize_benchmark_hip(), "finalize_benchmark_hip");
    CheckStatus(d3h_host_free(f1_), "d3h_host_free");
    f1_ = nullptr;
  }
  void DisplayResult(int count, float time) override {
    Baseline::DisplayResult(count, time);
    float time_wo_pci = 0.f;
    CheckStatus(kernel_time_hip(&time_wo_pci), "kernel_time_hip");
    printf("Kernel-only performance:\n");
    printf("Elapsed time : %.3f (s)\n", time_wo_pci);
    printf("FLOPS        : %.3f (GFLOPS)\n", GetGFLOPS(count, time_wo_pci));
    printf("Throughput   : %.3f (GB/s)\n", GetThroughput(count, time_wo_pci));
  }
};

}  // namespace diffusion3d
