// Drives the engine's host-side functions (constants, analytical fields,
// noise) on small, odd-sized grids; built by tests/test_sanitizers.py with
// g++ -fsanitize=address,undefined together with
// benchmarks_amd/host/initialize.cpp.  Test infrastructure only.
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "diffusion3d_hip.h"

int main() {
  const int nx = 13, ny = 7, nz = 5;
  const size_t n = (size_t)nx * ny * nz;
  float k[15];
  double kd[15];
  std::vector<float> a(n);
  std::vector<double> ad(n);
  if (d3h_constants_f32(nx, ny, nz, k) || d3h_constants_f64(nx, ny, nz, kd)) return 1;
  if (d3h_initialize_f32(a.data(), nx, ny, nz, k, 3.0f * k[7])) return 2;
  if (d3h_initialize_f64(ad.data(), nx, ny, nz, kd, 3.0 * kd[7])) return 3;
  if (d3h_add_noise_f32(a.data(), n, 11, 1e-3f)) return 4;
  float k2[15];
  if (d3h_constants2d_f32(nx, ny, k2) || d3h_initialize2d_f32(a.data(), nx, ny, k2, 0.f)) return 5;
  if (d3h_constants_f32(0, 1, 1, k) != D3H_ERR_INVALID) return 6;
  if (d3h_initialize2d_f32(nullptr, nx, ny, k2, 0.f) != D3H_ERR_INVALID) return 7;
  printf("%g %g\n", a[n / 2], ad[n / 3]);
  return 0;
}
