// The diffusion/ family's 2D benchmark over the HIP engine's C ABI.
//
// Mirrors that family's framework for ndim = 2 (diffusion/diffusion.h:80-180,
// diffusion/baseline.h:10-110): constructor constants, RunBenchmark with its
// printed lines and warm-up, 13-flop GFLOPS and 2-access throughput, float
// accumulated GetAccuracy, "%.7E" Dump to diffusion_result.<name>.out.  The
// step is diffusion/diffusion_openmp.cc:59-82 (s = y-1 summed before n = y+1),
// computed by the engine on an nx x ny x 1 device field whose rows are stored
// in reverse order (d3h_field_set_y_reversed) with ct = cb = 0, which gives
// the family's bits.  Engine knobs: D3H_TEMPORAL_DEPTH, D3H_ARITH,
// D3H_USE_GRAPH, D3H_DEVICE, D3H_TUNE_FILE, D3H_AUTOTUNE, as for the 3D
// driver.
#pragma once

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <fstream>
#include <iostream>
#include <string>

#include "diffusion3d.hpp"

namespace diffusion {

using diffusion3d::CheckStatus;
using diffusion3d::Stopwatch;
using diffusion3d::StopwatchStart;
using diffusion3d::StopwatchStop;

inline int EnvInt(const char *name, int dflt) {
  const char *v = getenv(name);
  return (v && *v) ? atoi(v) : dflt;
}

class Diffusion2DHIP {
 public:
  Diffusion2DHIP(int nx, int ny) : nx_(nx), ny_(ny) {
    CheckStatus(d3h_constants2d_f32(nx, ny, constants_), "d3h_constants2d_f32");
    dt_ = constants_[7];
    ce_ = constants_[8]; cw_ = constants_[9]; cn_ = constants_[10]; cs_ = constants_[11];
    cc_ = constants_[14];
  }
  ~Diffusion2DHIP() {
    if (field_) d3h_field_destroy(field_);
    if (ctx_) d3h_context_destroy(ctx_);
    if (f1_) d3h_host_free(f1_);
  }
  std::string GetName() const { return "hip"; }
  void SetRawDump(bool on) { raw_dump_ = on; }
  std::string GetDescription() const { return "HIP engine, 5-point operator on reversed rows"; }

  // diffusion/diffusion.h:112-141 with ndim = 2.
  void RunBenchmark(int count, bool dump, bool warmup) {
    std::cout << "*** Diffusion Benchmark ***\n";
    std::cout << "Benchmark: " << GetName() << " (" << GetDescription() << ")\n";
    Setup();
    std::cout << "Initializing benchmark input...\n";
    InitializeInput();
    std::cout << "Iteration count: " << count << "\n";
    std::cout << "Grid size: " << nx_ << "x" << ny_ << "\n";
    if (warmup) {
      std::cout << "Warming up the kernel...\n";
      RunKernel(1);
      std::cout << "Reinitializing benchmark input...\n";
      InitializeInput();
    }
    std::cout << "Running the kernel...\n";
    Stopwatch st;
    StopwatchStart(&st);
    RunKernel(count);
    const float elapsed_time = StopwatchStop(&st);
    std::cout << "Benchmarking finished.\n";
    DisplayResult(count, elapsed_time);
    if (dump) Dump();
    if (raw_dump_) {
      const std::string path = std::string("diffusion_result.") + GetName() + ".raw";
      CheckStatus(d3h_raw_write(path.c_str(), f1_, D3H_F32, nx_, ny_, 1, (uint64_t)count, 0),
                  "d3h_raw_write");
    }
  }

 private:
  int nx_, ny_;
  float constants_[15];
  float dt_, ce_, cw_, cn_, cs_, cc_;
  float *f1_ = nullptr;
  d3h_context *ctx_ = nullptr;
  d3h_field *field_ = nullptr;
  float kernel_seconds_ = 0.f;
  bool raw_dump_ = false;

  size_t Cells() const { return (size_t)nx_ * ny_; }

  void Setup() {
    void *p = nullptr;
    CheckStatus(d3h_host_alloc(sizeof(float) * Cells(), &p), "d3h_host_alloc");
    f1_ = static_cast<float *>(p);
    CheckStatus(d3h_context_create(EnvInt("D3H_DEVICE", 0), &ctx_), "d3h_context_create");
    CheckStatus(d3h_field_create(ctx_, nx_, ny_, 1, D3H_F32, &field_), "d3h_field_create");
    CheckStatus(d3h_field_set_y_reversed(field_, 1), "d3h_field_set_y_reversed");
    const char *tf = getenv("D3H_TUNE_FILE");
    if (tf && *tf && std::ifstream(tf))
      CheckStatus(d3h_autotune_load(ctx_, tf), "d3h_autotune_load");
    if (EnvInt("D3H_AUTOTUNE", 0)) {
      CheckStatus(d3h_autotune(ctx_, nx_, ny_, 1, D3H_F32, EnvInt("D3H_TEMPORAL_DEPTH", 1),
                               EnvInt("D3H_ARITH", D3H_ARITH_EXACT), nullptr, nullptr, nullptr),
                  "d3h_autotune");
      if (tf && *tf) CheckStatus(d3h_autotune_save(ctx_, tf), "d3h_autotune_save");
    }
  }
  void InitializeInput() {
    CheckStatus(d3h_initialize2d_f32(f1_, nx_, ny_, constants_, 0.0f), "d3h_initialize2d_f32");
  }
  void RunKernel(int count) {
    const double coef[7] = {ce_, cw_, cn_, cs_, 0.0, 0.0, cc_};
    d3h_run_config cfg;
    d3h_run_config_default(&cfg);
    cfg.temporal_depth = EnvInt("D3H_TEMPORAL_DEPTH", 1);
    cfg.arith = EnvInt("D3H_ARITH", D3H_ARITH_EXACT);
    cfg.use_graph = EnvInt("D3H_USE_GRAPH", 0);
    CheckStatus(d3h_field_upload(field_, f1_, D3H_F32), "d3h_field_upload");
    d3h_timing t{};
    CheckStatus(d3h_field_run(field_, count, coef, &cfg, &t), "d3h_field_run");
    kernel_seconds_ = (float)(t.kernel_ms * 1e-3);
    CheckStatus(d3h_field_download(field_, f1_, D3H_F32), "d3h_field_download");
  }
  // diffusion/diffusion.h:154-161 (nz = 1).
  float GetThroughput(int count, float time) const {
    return (double)Cells() * sizeof(float) * 2.0 * ((float)count) / time * 1.0e-09;
  }
  float GetGFLOPS(int count, float time) const {
    return (double)Cells() * 13.0 * (float)(count) / time * 1.0e-09;
  }
  // diffusion/baseline.h:86-95: float accumulator, serial order; the
  // analytical field at t = count * dt (diffusion/diffusion.h:172-179).
  float GetAccuracy(int count) const {
    float *ref = static_cast<float *>(malloc(sizeof(float) * Cells()));
    if (!ref) {
      fprintf(stderr, "out of host memory\n");
      exit(EXIT_FAILURE);
    }
    CheckStatus(d3h_initialize2d_f32(ref, nx_, ny_, constants_, count * dt_), "d3h_initialize2d_f32");
    float err = 0.0f;
    const long len = (long)Cells();
    for (long i = 0; i < len; i++) {
      const float diff = ref[i] - f1_[i];
      err += diff * diff;
    }
    free(ref);
    return sqrtf(err / len);
  }
  void DisplayResult(int count, float time) {
    printf("Elapsed time : %.3f (s)\n", time);
    printf("FLOPS        : %.3f (GFLOPS)\n", GetGFLOPS(count, time));
    printf("Throughput   : %.3f (GB/s)\n", GetThroughput(count, time));
    printf("Accuracy     : %e\n", GetAccuracy(count));
    printf("Kernel-only performance:\n");
    printf("Elapsed time : %.3f (s)\n", kernel_seconds_);
    printf("FLOPS        : %.3f (GFLOPS)\n", GetGFLOPS(count, kernel_seconds_));
    printf("Throughput   : %.3f (GB/s)\n", GetThroughput(count, kernel_seconds_));
  }
  // diffusion/baseline.h:97-105, path diffusion/diffusion.h:144-147.
  void Dump() const {
    const std::string path = std::string("diffusion_result.") + GetName() + ".out";
    FILE *out = fopen(path.c_str(), "w");
    if (!out) {
      fprintf(stderr, "cannot open %s\n", path.c_str());
      exit(EXIT_FAILURE);
    }
    for (size_t i = 0; i < Cells(); ++i) fprintf(out, "%.7E\n", f1_[i]);
    fclose(out);
  }
};

}  // namespace diffusion
