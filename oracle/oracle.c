/*
 * CPU oracle for the 3D 7-point diffusion stencil -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU path (naoyam/benchmarks,
 * src/diffusion3d).  Only the test suite, the smoke check and the CPU-baseline
 * leg of bench.py load it.  It is never part of the product path.
 *
 * Arithmetic follows the reference's C++ promotions exactly (REAL = float):
 * float libm (cosf/expf) where the reference passes float arguments, double
 * products where it multiplies by double literals, and the 7-term left-to-right
 * sum with diffusion3d's y order (s = y+1 before n = y-1).  Build with
 * -ffp-contract=off so no multiply-add is fused (the reference's g++ -O3 build
 * for baseline x86-64 has no FMA instruction to fuse into).
 *
 * The double instantiation is the same body at REAL = double; the reference
 * never builds it (diffusion3d/diffusion3d.h:13 hard-wires float), so it is
 * pinned through the float instantiation's bitwise match plus the closed-form
 * and invariant checks in tests/test_oracle.py.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------------------
 * a1: Diffusion3D::Diffusion3D (diffusion3d/diffusion3d.h:54-66)
 * ------------------------------------------------------------------------- */
void oracle_constants_f32(int nx, int ny, int nz, float o[ORACLE_NCONST]) {
  float kappa = (float)0.1;                  /* kappa_(0.1)            :55 */
  float l = 1.0f;                            /* REAL l = 1.0           :56 */
  float dx = l / (float)nx;                  /* dx_ = l / nx           :57 */
  float dy = l / (float)ny;
  float dz = l / (float)nz;
  float kx = (float)(2.0 * M_PI);            /* kx_ = 2.0 * M_PI       :60 */
  /* dt_ = 0.1 * dx_ * dx_ / kappa_ evaluates in double          :61 */
  float dt = (float)(0.1 * (double)dx * (double)dx / (double)kappa);
  float ce = kappa * dt / (dx * dx);         /* :62 */
  float cn = kappa * dt / (dy * dy);         /* :63 */
  float ct = kappa * dt / (dz * dz);         /* :64 */
  /* cc_ = 1.0 - (ce_ + cw_ + cn_ + cs_ + ct_ + cb_): float sum, double
   * subtraction, stored to float                                   :65 */
  float s = ce + ce;
  s = s + cn;
  s = s + cn;
  s = s + ct;
  s = s + ct;
  float cc = (float)(1.0 - (double)s);
  o[0] = kappa; o[1] = dx; o[2] = dy; o[3] = dz;
  o[4] = kx; o[5] = kx; o[6] = kx; o[7] = dt;
  o[8] = ce; o[9] = ce; o[10] = cn; o[11] = cn; o[12] = ct; o[13] = ct;
  o[14] = cc;
}

void oracle_constants_f64(int nx, int ny, int nz, double o[ORACLE_NCONST]) {
  double kappa = 0.1;
  double l = 1.0;
  double dx = l / nx, dy = l / ny, dz = l / nz;
  double kx = 2.0 * M_PI;
  double dt = 0.1 * dx * dx / kappa;
  double ce = kappa * dt / (dx * dx);
  double cn = kappa * dt / (dy * dy);
  double ct = kappa * dt / (dz * dz);
  double cc = 1.0 - (ce + ce + cn + cn + ct + ct);
  o[0] = kappa; o[1] = dx; o[2] = dy; o[3] = dz;
  o[4] = kx; o[5] = kx; o[6] = kx; o[7] = dt;
  o[8] = ce; o[9] = ce; o[10] = cn; o[11] = cn; o[12] = ct; o[13] = ct;
  o[14] = cc;
}

/* ---------------------------------------------------------------------------
 * a2: diffusion3d::Initialize (diffusion3d/diffusion3d.cc:14-39)
 * f = 0.125 * (1 - ax cos(kx x)) (1 - ay cos(ky y)) (1 - az cos(kz z)),
 * x = dx * (REAL)(j + 0.5).
 * ------------------------------------------------------------------------- */
void oracle_initialize_f32(float *buf, int nx, int ny, int nz,
                           const float k[ORACLE_NCONST], float time) {
  const float kappa = k[0], dx = k[1], dy = k[2], dz = k[3];
  const float kx = k[4], ky = k[5], kz = k[6];
  /* exp(-kappa*time*(kx*kx)) with float arguments -> expf       :18-20 */
  const float ax = expf(-kappa * time * (kx * kx));
  const float ay = expf(-kappa * time * (ky * ky));
  const float az = expf(-kappa * time * (kz * kz));
#pragma omp parallel for
  for (int jz = 0; jz < nz; jz++) {
    for (int jy = 0; jy < ny; jy++) {
      for (int jx = 0; jx < nx; jx++) {
        long j = (long)jz * nx * ny + (long)jy * nx + jx;
        float x = dx * (float)((double)jx + 0.5);     /* :28 */
        float y = dy * (float)((double)jy + 0.5);     /* :29 */
        float z = dz * (float)((double)jz + 0.5);     /* :30 */
        /* (REAL)0.125 * (1.0 - ax*cos(kx*x)) * ... : ax*cosf is a float
         * product, the rest is double, stored to float       :31-34 */
        double f0 = (double)0.125f * (1.0 - (double)(ax * cosf(kx * x)));
        f0 = f0 * (1.0 - (double)(ay * cosf(ky * y)));
        f0 = f0 * (1.0 - (double)(az * cosf(kz * z)));
        buf[j] = (float)f0;
      }
    }
  }
}

void oracle_initialize_f64(double *buf, int nx, int ny, int nz,
                           const double k[ORACLE_NCONST], double time) {
  const double kappa = k[0], dx = k[1], dy = k[2], dz = k[3];
  const double kx = k[4], ky = k[5], kz = k[6];
  const double ax = exp(-kappa * time * (kx * kx));
  const double ay = exp(-kappa * time * (ky * ky));
  const double az = exp(-kappa * time * (kz * kz));
#pragma omp parallel for
  for (int jz = 0; jz < nz; jz++) {
    for (int jy = 0; jy < ny; jy++) {
      for (int jx = 0; jx < nx; jx++) {
        long j = (long)jz * nx * ny + (long)jy * nx + jx;
        double x = dx * ((double)jx + 0.5);
        double y = dy * ((double)jy + 0.5);
        double z = dz * ((double)jz + 0.5);
        double f0 = 0.125 * (1.0 - ax * cos(kx * x));
        f0 = f0 * (1.0 - ay * cos(ky * y));
        f0 = f0 * (1.0 - az * cos(kz * z));
        buf[j] = f0;
      }
    }
  }
}

/* ---------------------------------------------------------------------------
 * a4/a5: one step (diffusion3d/baseline.cc:22-49, the OpenMP form at
 * diffusion3d/diffusion3d_openmp.cc:43-71).  Clamped neighbours (:33-38),
 * sum order cc c + cw w + ce e + cs s + cn n + cb b + ct t (:39-40).
 * ------------------------------------------------------------------------- */
#define ORACLE_STEP_BODY(REALT)                                              \
  const REALT ce = c[0], cw = c[1], cn = c[2], cs = c[3];                   \
  const REALT ct = c[4], cb = c[5], cc = c[6];                              \
  const long sxy = (long)nx * ny;                                           \
  _Pragma("omp parallel for")                                               \
  for (int z = 0; z < nz; z++) {                                            \
    for (int y = 0; y < ny; y++) {                                          \
      for (int x = 0; x < nx; x++) {                                        \
        long ci = x + (long)y * nx + (long)z * sxy;                         \
        long w = (x == 0) ? ci : ci - 1;                                    \
        long e = (x == nx - 1) ? ci : ci + 1;                               \
        long n = (y == 0) ? ci : ci - nx;                                   \
        long s = (y == ny - 1) ? ci : ci + nx;                              \
        long b = (z == 0) ? ci : ci - sxy;                                  \
        long t = (z == nz - 1) ? ci : ci + sxy;                             \
        f2[ci] = cc * f1[ci] + cw * f1[w] + ce * f1[e] + cs * f1[s] +       \
                 cn * f1[n] + cb * f1[b] + ct * f1[t];                      \
      }                                                                     \
    }                                                                       \
  }

void oracle_step_f32(const float *f1, float *f2, int nx, int ny, int nz,
                     const float c[ORACLE_NCOEF]) {
  ORACLE_STEP_BODY(float)
}

void oracle_step_f64(const double *f1, double *f2, int nx, int ny, int nz,
                     const double c[ORACLE_NCOEF]) {
  ORACLE_STEP_BODY(double)
}

void oracle_run_f32(float *f1, float *f2, int nx, int ny, int nz,
                    const float c[ORACLE_NCOEF], int count) {
  float *a = f1, *b = f2;
  for (int i = 0; i < count; ++i) {
    oracle_step_f32(a, b, nx, ny, nz, c);
    float *t = a; a = b; b = t;       /* diffusion3d/baseline.cc:44-46 */
  }
  if (a != f1) memcpy(f1, a, sizeof(float) * (size_t)nx * ny * nz);
}

void oracle_run_f64(double *f1, double *f2, int nx, int ny, int nz,
                    const double c[ORACLE_NCOEF], int count) {
  double *a = f1, *b = f2;
  for (int i = 0; i < count; ++i) {
    oracle_step_f64(a, b, nx, ny, nz, c);
    double *t = a; a = b; b = t;
  }
  if (a != f1) memcpy(f1, a, sizeof(double) * (size_t)nx * ny * nz);
}

/* ---------------------------------------------------------------------------
 * Reduced-precision storage (no reference counterpart).
 * bfloat16: round-to-nearest-even on the upper 16 bits of the float.
 * binary16: round-to-nearest-even including subnormals, by the magic-number
 * method (align the 10 mantissa bits by a float addition for subnormal
 * results; add the rounding bias 0xfff + lsb for normal ones).
 * ------------------------------------------------------------------------- */
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static inline uint16_t to_bf16(float f) {
  uint32_t u = f2u(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
static inline float from_bf16(uint16_t h) { return u2f((uint32_t)h << 16); }

static inline uint16_t to_f16(float fin) {
  uint32_t f = f2u(fin);
  uint32_t sign = f & 0x80000000u;
  uint16_t o;
  f ^= sign;
  if (f >= (uint32_t)(127 + 16) << 23) {              /* overflow, inf, nan */
    /* nan: quieted, payload kept (the conversion of gfx950 and IEEE 754) */
    o = (f > 0x7f800000u) ? (uint16_t)(0x7e00u | ((f & 0x7fffffu) >> 13)) : 0x7c00;
  } else if (f < (uint32_t)113 << 23) {                /* half subnormal / 0 */
    const uint32_t magic = (uint32_t)((127 - 15) + (23 - 10) + 1) << 23;
    float r = u2f(f) + u2f(magic);
    o = (uint16_t)(f2u(r) - magic);
  } else {
    uint32_t odd = (f >> 13) & 1u;
    f += ((uint32_t)(15 - 127) << 23) + 0xfffu;
    f += odd;
    o = (uint16_t)(f >> 13);
  }
  return (uint16_t)(o | (sign >> 16));
}
static inline float from_f16(uint16_t h) {
  uint32_t s = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  if (e == 0) {
    float v = (float)m * 0x1p-24f;                      /* exact */
    return s ? -v : v;
  }
  if (e == 31) return u2f(s | 0x7f800000u | (m << 13));
  return u2f(s | ((e + 112u) << 23) | (m << 13));
}

void oracle_f32_to_lp(const float *in, uint16_t *out, long n, int kind) {
  for (long i = 0; i < n; ++i) out[i] = kind == 0 ? to_bf16(in[i]) : to_f16(in[i]);
}
void oracle_lp_to_f32(const uint16_t *in, float *out, long n, int kind) {
  for (long i = 0; i < n; ++i) out[i] = kind == 0 ? from_bf16(in[i]) : from_f16(in[i]);
}

void oracle_run_lp(uint16_t *f1, uint16_t *f2, int nx, int ny, int nz,
                   const float c[ORACLE_NCOEF], int count, int kind) {
  const long n = (long)nx * ny * nz;
  float *a = (float *)malloc(sizeof(float) * (size_t)n);
  float *b = (float *)malloc(sizeof(float) * (size_t)n);
  oracle_lp_to_f32(f1, a, n, kind);
  for (int i = 0; i < count; ++i) {
    oracle_step_f32(a, b, nx, ny, nz, c);
    /* store to the reduced-precision field and read it back */
    for (long j = 0; j < n; ++j) {
      uint16_t h = kind == 0 ? to_bf16(b[j]) : to_f16(b[j]);
      a[j] = kind == 0 ? from_bf16(h) : from_f16(h);
    }
  }
  oracle_f32_to_lp(a, f1, n, kind);
  (void)f2;
  free(a);
  free(b);
}

/* Stochastic rounding of 16-bit storage (no reference counterpart).  The
 * random bits of a row and step come from the "lowbias32" integer mixer
 * (C. Wellons, hash-prospector) of a Weyl-style combination of the seed, the
 * row index and the step number; the cells of the row add a Weyl increment
 * per x pair to it. */
static inline uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

uint32_t oracle_sr_hash(uint32_t seed, uint32_t row, uint32_t pair, uint32_t step) {
  return lowbias32(row * 0x9E3779B1u + step * 0x85EBCA77u + seed) + pair * 0x9E3779B9u;
}

uint16_t oracle_sr_round(float fin, uint32_t r, int kind) {
  const uint32_t u = f2u(fin);
  const uint32_t a = u & 0x7fffffffu;
  if (kind == 0) {
    if (a >= 0x7f800000u) return to_bf16(fin);           /* inf, nan: nearest */
    return (uint16_t)((u + (r & 0xffffu)) >> 16);
  }
  const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
  if (a >= 0x7f800000u) return to_f16(fin);              /* inf, nan: nearest */
  const uint32_t rr = (r & 0xffffu) >> 3;                /* 13 bits: units of ulp / 2^13 */
  const uint32_t e = a >> 23;
  if (e >= 113u) {                                       /* FP16 normal range and above */
    const uint32_t t = (a + rr) & 0xffffe000u;           /* an FP16 value, or >= 2^16 */
    if (t >= 0x47800000u) return (uint16_t)(sign | 0x7c00u);
    return (uint16_t)(sign | to_f16(u2f(t)));
  }
  /* |u| below 2^-14: in units of 2^-37 (= 2^-24 / 2^13) it is m * 2^(e-113);
   * the FP16 subnormal count is floor((that + rr) / 2^13), and only the
   * integer part of the scaled value can matter */
  const uint32_t m = e ? ((a & 0x7fffffu) | 0x800000u) : (a & 0x7fffffu);
  const uint32_t sh = 113u - (e ? e : 1u);
  const uint32_t scaled = sh >= 32u ? 0u : m >> sh;
  return (uint16_t)(sign | ((scaled + rr) >> 13));       /* 0x400: the least normal */
}

void oracle_run_lp_sr(uint16_t *f1, int nx, int ny, int nz, const float c[ORACLE_NCOEF], int count,
                      int kind, uint32_t seed, int step0) {
  const long n = (long)nx * ny * nz;
  float *a = (float *)malloc(sizeof(float) * (size_t)n);
  float *b = (float *)malloc(sizeof(float) * (size_t)n);
  oracle_lp_to_f32(f1, a, n, kind);
  for (int i = 0; i < count; ++i) {
    oracle_step_f32(a, b, nx, ny, nz, c);
    const uint32_t step = (uint32_t)(step0 + i);
    for (long j = 0; j < n; ++j) {
      const uint32_t x = (uint32_t)(j % nx), row = (uint32_t)(j / nx);
      const uint32_t h = oracle_sr_hash(seed, row, x / 2u, step);
      const uint32_t r = (x & 1u) ? (h & 0xffffu) : (h >> 16);
      const uint16_t v = oracle_sr_round(b[j], r, kind);
      a[j] = kind == 0 ? from_bf16(v) : from_f16(v);
    }
  }
  oracle_f32_to_lp(a, f1, n, kind);
  free(a);
  free(b);
}

void oracle_run_bf16_sr(uint16_t *f1, int nx, int ny, int nz, const float c[ORACLE_NCOEF],
                        int count, uint32_t seed, int step0) {
  oracle_run_lp_sr(f1, nx, ny, nz, c, count, 0, seed, step0);
}

/* ---------------------------------------------------------------------------
 * a7: Baseline::GetAccuracy + GetCorrectAnswer (diffusion3d/baseline.cc:51-60,
 * diffusion3d/diffusion3d.h:113-120).  time = count * dt_ in float.
 * ------------------------------------------------------------------------- */
float oracle_accuracy_f32(const float *f, int nx, int ny, int nz,
                          const float k[ORACLE_NCONST], int count) {
  const long len = (long)nx * ny * nz;
  float *ref = (float *)malloc(sizeof(float) * (size_t)len);
  oracle_initialize_f32(ref, nx, ny, nz, k, (float)count * k[7]);
  float err = 0.0f;
  for (long i = 0; i < len; i++) {
    float diff = ref[i] - f[i];
    err += diff * diff;
  }
  free(ref);
  return sqrtf(err / (float)len);
}

double oracle_rms_vs_analytical_f32(const float *f, int nx, int ny, int nz,
                                    const float k[ORACLE_NCONST], int count) {
  const long len = (long)nx * ny * nz;
  float *ref = (float *)malloc(sizeof(float) * (size_t)len);
  oracle_initialize_f32(ref, nx, ny, nz, k, (float)count * k[7]);
  double err = 0.0;
  for (long i = 0; i < len; i++) {
    double d = (double)ref[i] - (double)f[i];
    err += d * d;
  }
  free(ref);
  return sqrt(err / (double)len);
}

double oracle_rms_vs_analytical_f64(const double *f, int nx, int ny, int nz,
                                    const double k[ORACLE_NCONST], int count) {
  const long len = (long)nx * ny * nz;
  double *ref = (double *)malloc(sizeof(double) * (size_t)len);
  oracle_initialize_f64(ref, nx, ny, nz, k, (double)count * k[7]);
  double err = 0.0;
  for (long i = 0; i < len; i++) {
    double d = ref[i] - f[i];
    err += d * d;
  }
  free(ref);
  return sqrt(err / (double)len);
}

/* ---------------------------------------------------------------------------
 * 2D operator of the diffusion/ family
 * ------------------------------------------------------------------------- */
void oracle_constants2d_f32(int nx, int ny, float o[ORACLE_NCONST]) {
  /* diffusion/diffusion.h:93-108 with ndim = 2 (nz = 1) */
  float kappa = (float)0.1;
  float l = 1.0f;
  float dx = l / (float)nx, dy = l / (float)ny, dz = l / 1.0f;
  float kx = (float)(2.0 * M_PI);
  float dt = (float)(0.1 * (double)dx * (double)dx / (double)kappa);
  float ce = kappa * dt / (dx * dx);
  float cn = kappa * dt / (dy * dy);
  /* 1.0 - (ce + cw + cn + cs + (ndim == 3 ? (ct + cb) : 0)) */
  float s = ce + ce;
  s = s + cn;
  s = s + cn;
  s = s + 0.0f;
  float cc = (float)(1.0 - (double)s);
  o[0] = kappa; o[1] = dx; o[2] = dy; o[3] = dz;
  o[4] = kx; o[5] = kx; o[6] = kx; o[7] = dt;
  o[8] = ce; o[9] = ce; o[10] = cn; o[11] = cn; o[12] = 0.0f; o[13] = 0.0f;
  o[14] = cc;
}

void oracle_initialize2d_f32(float *buf, int nx, int ny, const float k[ORACLE_NCONST], float time) {
  /* diffusion/diffusion.h:48-77, ndim = 2 */
  const float kappa = k[0], dx = k[1], dy = k[2], kx = k[4], ky = k[5];
  const float ax = expf(-kappa * time * (kx * kx));
  const float ay = expf(-kappa * time * (ky * ky));
  for (int jy = 0; jy < ny; jy++) {
    for (int jx = 0; jx < nx; jx++) {
      float x = dx * (float)((double)jx + 0.5);
      float y = dy * (float)((double)jy + 0.5);
      double f0 = (double)0.125f * (1.0 - (double)(ax * cosf(kx * x)));
      f0 = f0 * (1.0 - (double)(ay * cosf(ky * y)));
      buf[(long)jy * nx + jx] = (float)f0;
    }
  }
}

void oracle_run2d_f32(float *f1, float *f2, int nx, int ny, const float c[ORACLE_NCOEF], int count) {
  /* diffusion/diffusion_openmp.cc:59-82: s = y-1, n = y+1 */
  const float ce = c[0], cw = c[1], cn = c[2], cs = c[3], cc = c[6];
  float *a = f1, *b = f2;
  for (int it = 0; it < count; ++it) {
#pragma omp parallel for
    for (int y = 0; y < ny; y++) {
      for (int x = 0; x < nx; x++) {
        long ci = x + (long)y * nx;
        long w = (x == 0) ? ci : ci - 1;
        long e = (x == nx - 1) ? ci : ci + 1;
        long s = (y == 0) ? ci : ci - nx;
        long n = (y == ny - 1) ? ci : ci + nx;
        b[ci] = cc * a[ci] + cw * a[w] + ce * a[e] + cs * a[s] + cn * a[n];
      }
    }
    float *t = a; a = b; b = t;
  }
  if (a != f1) memcpy(f1, a, sizeof(float) * (size_t)nx * ny);
}

float oracle_accuracy2d_f32(const float *f, int nx, int ny, const float k[ORACLE_NCONST], int count) {
  const long len = (long)nx * ny;
  float *ref = (float *)malloc(sizeof(float) * (size_t)len);
  oracle_initialize2d_f32(ref, nx, ny, k, (float)count * k[7]);
  float err = 0.0f;
  for (long i = 0; i < len; i++) {
    float diff = ref[i] - f[i];
    err += diff * diff;
  }
  free(ref);
  return sqrtf(err / (float)len);
}

/* ---------------------------------------------------------------------------
 * Seeded noise (counter-based SplitMix64 finaliser keyed by seed and index).
 * ------------------------------------------------------------------------- */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void oracle_add_noise_f32(float *buf, long n, uint64_t seed, float amp) {
  for (long i = 0; i < n; ++i) {
    uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)i + 1ull);
    float u = (float)(uint32_t)(h >> 40) * 0x1p-24f;   /* [0,1), exact */
    buf[i] = buf[i] + amp * (2.0f * u - 1.0f);
  }
}

void oracle_diff_norms_f32(const float *a, const float *b, long n, double out[3]) {
  double d2 = 0.0, b2 = 0.0, mx = 0.0;
  for (long i = 0; i < n; ++i) {
    double d = (double)a[i] - (double)b[i];
    d2 += d * d;
    b2 += (double)b[i] * (double)b[i];
    if (fabs(d) > mx) mx = fabs(d);
  }
  out[0] = b2 > 0 ? sqrt(d2 / b2) : sqrt(d2);
  out[1] = mx;
  out[2] = sqrt(d2);
}

void oracle_diff_norms_f64(const double *a, const double *b, long n, double out[3]) {
  double d2 = 0.0, b2 = 0.0, mx = 0.0;
  for (long i = 0; i < n; ++i) {
    double d = a[i] - b[i];
    d2 += d * d;
    b2 += b[i] * b[i];
    if (fabs(d) > mx) mx = fabs(d);
  }
  out[0] = b2 > 0 ? sqrt(d2 / b2) : sqrt(d2);
  out[1] = mx;
  out[2] = sqrt(d2);
}
