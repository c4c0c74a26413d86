/* Drives every oracle function once on small, odd-sized grids; built by
 * tests/test_sanitizers.py with gcc -fsanitize=address,undefined so that out
 * of bounds accesses, leaks and undefined behaviour in the oracle fail a CPU
 * test.  Test infrastructure only. */
#include <stdio.h>
#include <stdlib.h>
#include "oracle.h"
int main(void) {
  int nx = 13, ny = 7, nz = 5; long n = (long)nx * ny * nz;
  float k[ORACLE_NCONST]; double kd[ORACLE_NCONST];
  oracle_constants_f32(nx, ny, nz, k); oracle_constants_f64(nx, ny, nz, kd);
  float *a = malloc(sizeof(float) * n), *b = malloc(sizeof(float) * n);
  double *ad = malloc(sizeof(double) * n), *bd = malloc(sizeof(double) * n);
  uint16_t *h = malloc(2 * n), *h2 = malloc(2 * n);
  oracle_initialize_f32(a, nx, ny, nz, k, 0.f); oracle_initialize_f64(ad, nx, ny, nz, kd, 0.0);
  oracle_add_noise_f32(a, n, 3, 1e-3f);
  oracle_run_f32(a, b, nx, ny, nz, k + 8, 7); oracle_run_f64(ad, bd, nx, ny, nz, kd + 8, 7);
  oracle_step_f32(a, b, nx, ny, nz, k + 8); oracle_step_f64(ad, bd, nx, ny, nz, kd + 8);
  for (int kind = 0; kind < 2; ++kind) {
    oracle_f32_to_lp(a, h, n, kind); oracle_run_lp(h, h2, nx, ny, nz, k + 8, 5, kind);
    oracle_run_lp_sr(h, nx, ny, nz, k + 8, 5, kind, 7, 3); oracle_lp_to_f32(h, b, n, kind);
  }
  oracle_run_bf16_sr(h, nx, ny, nz, k + 8, 3, 1, 0);
  /* the rounding rules at their edges: zeros, subnormals, overflow, inf, nan */
  const uint32_t edge[] = {0u, 0x80000000u, 1u, 0x007fffffu, 0x387fffffu, 0x33000000u, 0x477ff000u,
                           0x7f7fffffu, 0x7f800000u, 0xff800000u, 0x7fc00001u, 0xffbfffffu};
  unsigned acc = 0;
  for (int i = 0; i < (int)(sizeof edge / sizeof edge[0]); ++i) {
    union { uint32_t u; float f; } v = {edge[i]};
    for (int kind = 0; kind < 2; ++kind)
      acc += oracle_sr_round(v.f, 0xffffu, kind) + oracle_sr_round(v.f, 0u, kind) +
             oracle_sr_hash(9u, (uint32_t)i, 0xffffffffu, 0xffffffffu);
  }
  printf("%u\n", acc);
  printf("%g %g %g\n", oracle_accuracy_f32(a, nx, ny, nz, k, 7),
         oracle_rms_vs_analytical_f32(a, nx, ny, nz, k, 7), oracle_rms_vs_analytical_f64(ad, nx, ny, nz, kd, 7));
  double o[3]; oracle_diff_norms_f32(a, b, n, o); oracle_diff_norms_f64(ad, bd, n, o);
  float k2[ORACLE_NCONST]; oracle_constants2d_f32(nx, ny, k2);
  oracle_initialize2d_f32(a, nx, ny, k2, 0.f); oracle_run2d_f32(a, b, nx, ny, k2 + 8, 9);
  printf("%g\n", oracle_accuracy2d_f32(a, nx, ny, k2, 9));
  free(a); free(b); free(ad); free(bd); free(h); free(h2);
  return 0;
}
