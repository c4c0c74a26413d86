/*
 * CPU oracle for the 3D 7-point diffusion stencil -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference CPU path of naoyam/benchmarks
 * (src/diffusion3d).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline; the
 * product path (benchmarks_amd) never calls it and fails loudly without its
 * HIP library.
 *
 * Every function cites the reference file:line it restates.  Pins (see
 * tests/test_oracle.py): bitwise equality against fixtures produced by the
 * reference's own CPU build (oracle/_ref, tests/golden), closed forms,
 * conservation, maximum principle, brute force on tiny grids.
 */
#ifndef DIFFUSION3D_ORACLE_H_
#define DIFFUSION3D_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Problem constants, in this order (diffusion3d/diffusion3d.h:47-53):
 * kappa, dx, dy, dz, kx, ky, kz, dt, ce, cw, cn, cs, ct, cb, cc */
enum { ORACLE_NCONST = 15 };
/* Stencil coefficient order used by every step function: ce cw cn cs ct cb cc
 * (the argument order of run_kernel_physis, diffusion3d/diffusion3d_physis.h:13-15). */
enum { ORACLE_NCOEF = 7 };

void oracle_constants_f32(int nx, int ny, int nz, float out[ORACLE_NCONST]);
void oracle_constants_f64(int nx, int ny, int nz, double out[ORACLE_NCONST]);

/* Analytical cosine-product field at time t (diffusion3d/diffusion3d.cc:14-39). */
void oracle_initialize_f32(float *buf, int nx, int ny, int nz,
                           const float k[ORACLE_NCONST], float time);
void oracle_initialize_f64(double *buf, int nx, int ny, int nz,
                           const double k[ORACLE_NCONST], double time);

/* One 7-point step f1 -> f2, OpenMP over z (diffusion3d/diffusion3d_openmp.cc:43-71). */
void oracle_step_f32(const float *f1, float *f2, int nx, int ny, int nz,
                     const float c[ORACLE_NCOEF]);
void oracle_step_f64(const double *f1, double *f2, int nx, int ny, int nz,
                     const double c[ORACLE_NCOEF]);

/* `count` steps with ping-pong; the result is left in f1 (the reference keeps
 * it in f1_ by swapping pointers, diffusion3d/baseline.cc:44-46). */
void oracle_run_f32(float *f1, float *f2, int nx, int ny, int nz,
                    const float c[ORACLE_NCOEF], int count);
void oracle_run_f64(double *f1, double *f2, int nx, int ny, int nz,
                    const double c[ORACLE_NCOEF], int count);

/* Reduced-precision storage with FP32 arithmetic in the reference's term order
 * (no reference counterpart; same step as above, rounded to storage after each
 * step, round-to-nearest-even).  kind: 0 = bfloat16, 1 = IEEE binary16. */
void oracle_run_lp(uint16_t *f1, uint16_t *f2, int nx, int ny, int nz,
                   const float c[ORACLE_NCOEF], int count, int kind);
void oracle_f32_to_lp(const float *in, uint16_t *out, long n, int kind);
/* 16-bit storage with deterministic stochastic rounding (no reference
 * counterpart).  After step t (numbered from step0) the FP32 value u of cell
 * (x, y, z) is rounded with 16 random bits r taken from
 *   base = lowbias32(row * 0x9E3779B1 + t * 0x85EBCA77 + seed),  row = z*ny + y
 *   h    = base + (x / 2) * 0x9E3779B9              (oracle_sr_hash)
 * r = the high half of h for even x, the low half for odd x.
 * kind 0 (BF16): finite u -> (bits(u) + r) >> 16, i.e. r added below the BF16
 *   mantissa and truncated (can carry to inf); inf and nan round to nearest
 *   (nan quieted).
 * kind 1 (FP16): finite u -> |u| + (r >> 3) * ulp / 2^13 truncated to the FP16
 *   grid (ulp = the FP16 ulp at |u|, 2^-24 below 2^-14, so subnormals round
 *   stochastically too; results >= 2^16 become inf); inf and nan round to
 *   nearest (nan quieted, payload kept).
 * These are the rounding rules of gfx950's v_cvt_sr_bf16_f32 and
 * v_cvt_sr_f16_f32 given the random operand h (even x) or h << 16 (odd x). */
void oracle_run_bf16_sr(uint16_t *f1, int nx, int ny, int nz, const float c[ORACLE_NCOEF],
                        int count, uint32_t seed, int step0);
void oracle_run_lp_sr(uint16_t *f1, int nx, int ny, int nz, const float c[ORACLE_NCOEF], int count,
                      int kind, uint32_t seed, int step0);
uint32_t oracle_sr_hash(uint32_t seed, uint32_t row, uint32_t pair, uint32_t step);
/* One value rounded as above with the 16 random bits r (kind 0 BF16, 1 FP16). */
uint16_t oracle_sr_round(float u, uint32_t r, int kind);
void oracle_lp_to_f32(const uint16_t *in, float *out, long n, int kind);

/* Legacy accuracy: RMS vs the analytical field with a float accumulator in
 * serial order (diffusion3d/baseline.cc:51-60, diffusion3d/diffusion3d.h:113-120). */
float oracle_accuracy_f32(const float *f, int nx, int ny, int nz,
                          const float k[ORACLE_NCONST], int count);
/* Same definition with an FP64 accumulator (metric form). */
double oracle_rms_vs_analytical_f32(const float *f, int nx, int ny, int nz,
                                    const float k[ORACLE_NCONST], int count);
double oracle_rms_vs_analytical_f64(const double *f, int nx, int ny, int nz,
                                    const double k[ORACLE_NCONST], int count);

/* Seeded additive noise, counter-based (SplitMix64 finaliser of seed and the
 * linear cell index), uniform in [-amp, amp); added in float.  No reference
 * counterpart (the reference initialises with the cosine product only). */
void oracle_add_noise_f32(float *buf, long n, uint64_t seed, float amp);

/* 2D 5-point operator of the diffusion/ family (nz = 1, no t/b terms):
 * constants of Diffusion(2, {nx, ny}) (diffusion/diffusion.h:93-108, ct = cb
 * reported as 0), the 2D cosine product (diffusion/diffusion.h:48-77), the
 * step with that family's y naming, s = y-1 summed before n = y+1
 * (diffusion/diffusion_openmp.cc:59-82), and its accuracy (diffusion/baseline.h
 * GetAccuracy, float accumulator). */
void oracle_constants2d_f32(int nx, int ny, float out[ORACLE_NCONST]);
void oracle_initialize2d_f32(float *buf, int nx, int ny, const float k[ORACLE_NCONST], float time);
void oracle_run2d_f32(float *f1, float *f2, int nx, int ny, const float c[ORACLE_NCOEF], int count);
float oracle_accuracy2d_f32(const float *f, int nx, int ny, const float k[ORACLE_NCONST], int count);

/* Relative L2 and max-abs difference in FP64 (metric helpers). */
void oracle_diff_norms_f32(const float *a, const float *b, long n, double out[3]);
void oracle_diff_norms_f64(const double *a, const double *b, long n, double out[3]);

#ifdef __cplusplus
}
#endif

#endif
