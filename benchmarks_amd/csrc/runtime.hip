// Runtime of the diffusion engine: contexts, device-resident fields, the step
// loop (direct launches or a captured hipGraph), FP64 error norms, dtype
// conversion, the HBM probe, the tile table and autotuner, and the
// Physis-shaped drop-in entry points.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "runtime.hpp"

namespace d3h {

static thread_local std::string t_err;

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_err = buf;
  return code;
}

size_t dtype_bytes(int dtype) {
  switch (dtype) {
    case D3H_F32: return 4;
    case D3H_F64: return 8;
    case D3H_BF16:
    case D3H_F16: return 2;
    default: return 0;
  }
}
bool dtype_valid(int dtype) { return dtype >= D3H_F32 && dtype <= D3H_F16; }
int pitch_elems(int nx, int dtype) {
  const int v = 16 / (int)dtype_bytes(dtype);
  return (nx + v - 1) / v * v;
}

const std::vector<KernelEntry> &registry() {
  static const std::vector<KernelEntry> all = [] {
    std::vector<KernelEntry> v;
    const KernelEntry *(*tables[])(int *) = {kernels_f32_plain, kernels_f32_tb, kernels_f64,
                                              kernels_bf16, kernels_f16, kernels_f32_rows};
    for (auto fn : tables) {
      int n = 0;
      const KernelEntry *t = fn(&n);
      v.insert(v.end(), t, t + n);
    }
    return v;
  }();
  return all;
}

const KernelEntry *default_entry(int dtype, int T, int fma, bool rows) {
  for (const auto &e : registry())
    if (e.dtype == dtype && e.T == T && e.fma == fma && (e.rows != 0) == rows) return &e;
  return nullptr;
}

int occupancy_of(const KernelEntry &e) {
  static std::mutex mu;
  static std::map<const void *, int> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(e.func);
  if (it != cache.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, e.func, 64 * e.NW, 0) != hipSuccess) nb = 0;
  cache[e.func] = nb;
  return nb;
}

int make_geometry(const KernelEntry &e, int nx, int ny, int nz, int pitch, int zchunks,
                  int num_cus, int occupancy, Geometry *g, unsigned *blocks) {
  const int TX = e.tile_x(), TY = e.tile_y(), T = e.T;
  if ((long long)pitch * ny * (long long)dtype_bytes(e.dtype) >= 0x7ffffff0LL)
    return fail(D3H_ERR_UNSUPPORTED, "one xy plane of %d x %d exceeds 2 GiB", pitch, ny);
  Geometry q{};
  q.nx = nx; q.ny = ny; q.nz = nz; q.pitch = pitch;
  q.plane = (long long)pitch * ny;
  if (nx <= TX) {
    q.ntx = 1; q.ox = nx; q.hx = 0;
  } else {
    q.hx = (T + e.vec - 1) / e.vec * e.vec;
    q.ox = TX - 2 * q.hx;
    if (q.ox <= 0) return fail(D3H_ERR_UNSUPPORTED, "tile %d wide cannot hold depth %d", TX, T);
    q.ntx = (nx + q.ox - 1) / q.ox;
  }
  if (ny <= TY) {
    q.nty = 1; q.oy = ny;
  } else {
    q.oy = TY - 2 * T;
    if (q.oy <= 0) return fail(D3H_ERR_UNSUPPORTED, "tile %d rows cannot hold depth %d", TY, T);
    q.nty = (ny + q.oy - 1) / q.oy;
  }
  const long long tiles = (long long)q.ntx * q.nty;
  int nzc = zchunks;
  if (nzc <= 0) {
    // One full wave of resident workgroups, but keep chunks at least
    // max(4T, 8) planes deep so the 2T halo planes stay a small overhead.
    const long long slots = (long long)std::max(occupancy, 1) * std::max(num_cus, 1);
    nzc = (int)std::max(1LL, slots / tiles);
    nzc = std::min(nzc, std::max(1, nz / std::max(4 * T, 8)));
  }
  nzc = std::max(1, std::min(nzc, nz));
  q.oz = (nz + nzc - 1) / nzc;
  if (nzc > 1) {
    // A chunk runs its z loop in whole groups of NS iterations (NS = the
    // register-slot count, PF + 3, or 3 for the ring), oz + T + H + max(PF, 1)
    // of them for an interior chunk with H halo planes (T, or ceil(T / RPW)
    // for the 2D kernels' row blocks): round oz up so that is a multiple of
    // NS and no interior chunk computes planes past its end.
    const int ns = e.pf > 0 ? e.pf + 3 : 3;
    const int halo = e.rows ? (T + e.RPW - 1) / e.RPW : T;
    const int lead = T + halo + std::max(e.pf, 1);
    while ((q.oz + lead) % ns != 0) ++q.oz;
  }
  q.nzc = (nz + q.oz - 1) / q.oz;
  const long long nb = tiles * q.nzc;
  if (nb <= 0 || nb > INT_MAX) return fail(D3H_ERR_UNSUPPORTED, "grid too large");
  *g = q;
  *blocks = (unsigned)nb;
  return D3H_OK;
}

static int ilog2ceil(int v) {
  int r = 0;
  while ((1 << r) < v) ++r;
  return r;
}

TuneKey tune_key(int dtype, int T, int fma, int nx, int ny, int nz, int batch) {
  return TuneKey{dtype, T, fma, ilog2ceil(nx), ilog2ceil(ny), ilog2ceil(nz), batch};
}

const KernelEntry *find_entry(int dtype, int T, int fma, int vec, int G, int RPW, int NW, int pf,
                              int rows) {
  for (const auto &e : registry())
    if (e.dtype == dtype && e.T == T && e.fma == fma && e.vec == vec && e.G == G && e.RPW == RPW &&
        e.NW == NW && e.pf == pf && (e.rows != 0) == (rows != 0))
      return &e;
  return nullptr;
}

const KernelEntry *find_entry(int dtype, int fma, const TuneVal &v) {
  return find_entry(dtype, v.T, fma, v.vec, v.G, v.RPW, v.NW, v.pf, v.rows);
}

bool rows_ok(const d3h_field *f, const double coef[7]) {
  return f->nz == 1 && coef[4] == 0.0 && coef[5] == 0.0;
}

// 2D kernels.  An nx x ny x 1 field with ct == cb == 0 is run as a stream over
// its rows in blocks of RPW rows: the tile runner sees planes of RPW rows
// (ny' = RPW, nz' = ceil(ny / RPW), plane stride RPW * pitch, the last block
// clipped to the field by rows2d) and evaluates the 5-point update
//   cc c + cw w + ce e + cs f[y+1] + cn f[y-1]
// with y neighbours taken across block boundaries from the adjacent blocks.
// This is the 7-point step of the nz == 1 field without its two clamped z
// terms (0 c each when ct == cb == 0), which can differ from it at most in
// the sign of an exactly zero result.
int field_geometry(const KernelEntry &e, const d3h_field *f, int zchunks, Geometry *g,
                   unsigned *blocks) {
  const int occ = occupancy_of(e), cus = f->ctx->num_cus;
  if (!e.rows) return make_geometry(e, f->nx, f->ny, f->nz, f->pitch, zchunks, cus, occ, g, blocks);
  if (f->nz != 1) return fail(D3H_ERR_INVALID, "2D tile on a field with nz = %d", f->nz);
  const int planes = (f->ny + e.RPW - 1) / e.RPW;
  if (zchunks <= 0) {
    // one full wave of resident workgroups, chunks of at least max(16T, 64)
    // rows (a row is cheap, so the halo blocks must stay a small share)
    D3H_TRY(make_geometry(e, f->nx, e.RPW, planes, f->pitch, 1, cus, occ, g, blocks));
    const long long slots = (long long)std::max(occ, 1) * std::max(cus, 1);
    const long long min_planes = (std::max(16 * e.T, 64) + e.RPW - 1) / e.RPW;
    zchunks = (int)std::max(1LL, std::min(slots / g->ntx, (long long)planes / min_planes));
  }
  D3H_TRY(make_geometry(e, f->nx, e.RPW, planes, f->pitch, zchunks, cus, occ, g, blocks));
  g->rows2d = f->ny;
  return D3H_OK;
}

static void launch_entry(const KernelEntry &e, const void *src, void *dst, const Geometry &g,
                         const double coef[7], unsigned blocks, hipStream_t stream) {
  e.launch(src, dst, g, coef, blocks, stream);
}

int tuned_depth(d3h_context *ctx, int dtype, int fma, int nx, int ny, int nz, int dflt) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->tuned.find(tune_key(dtype, 0, fma, nx, ny, nz));
  return it != ctx->tuned.end() ? it->second.T : dflt;
}

// A schedule slower than this per z iteration is refused: a plane takes
// microseconds, and a wait ends at its bound long before.
constexpr int kMaxPaceNs = 1000000;

int pace_ticks(const d3h_context *ctx, const KernelEntry &e, int pace_ns) {
  if (pace_ns <= 0 || e.T != 1 || ctx->rtc_khz <= 0) return 0;
  const long long t = (long long)pace_ns * ctx->rtc_khz * 256 / 1000000;
  return (int)std::min<long long>(t, INT_MAX);
}

int resolve_entry(d3h_context *ctx, int nx, int ny, int nz, int dtype, const d3h_run_config &cfg,
                  bool rows, const KernelEntry **entry, int *zchunks, int *pace_ns) {
  // an explicit schedule (> 0) or none (< 0) overrides the tuned one
  if (cfg.pace_ns > kMaxPaceNs)
    return fail(D3H_ERR_INVALID, "plane schedule of %d ns per iteration (at most %d)", cfg.pace_ns,
                kMaxPaceNs);
  *pace_ns = cfg.pace_ns > 0 ? cfg.pace_ns : 0;
  int T = cfg.temporal_depth;
  const int fma = cfg.arith == D3H_ARITH_FMA ? 1 : 0;
  if (cfg.arith != D3H_ARITH_EXACT && cfg.arith != D3H_ARITH_FMA)
    return fail(D3H_ERR_INVALID, "unknown arith mode %d", cfg.arith);
  if (T != 0 && T != 1 && T != 2 && T != 4 && T != 8)
    return fail(D3H_ERR_INVALID, "temporal depth %d (0 = tuned, 1, 2, 4 or 8)", T);
  const auto &reg = registry();
  if (cfg.tile >= 0) {
    if (cfg.tile >= (int)reg.size()) return fail(D3H_ERR_INVALID, "tile index %d out of range", cfg.tile);
    const KernelEntry &e = reg[cfg.tile];
    if (T == 0) T = e.T;  // an explicit tile brings its own depth
    if (e.dtype != dtype || e.T != T || e.fma != fma)
      return fail(D3H_ERR_INVALID, "tile %d is dtype %d depth %d fma %d, run wants %d/%d/%d", cfg.tile,
                  e.dtype, e.T, e.fma, dtype, T, fma);
    if (e.rows && !rows)
      return fail(D3H_ERR_INVALID, "tile %d streams rows: it needs nz == 1 and ct == cb == 0", cfg.tile);
    *entry = &e;
    *zchunks = cfg.zchunks;
    return D3H_OK;
  }
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (T == 0) {
      // the depth the autotuner chose for this size class, else plain steps
      auto it = ctx->tuned.find(tune_key(dtype, 0, fma, nx, ny, nz));
      if (it != ctx->tuned.end()) {
        const TuneVal &v = it->second;
        const KernelEntry *e = find_entry(dtype, fma, v);
        if (e && (!e->rows || rows)) {
          *entry = e;
          *zchunks = cfg.zchunks > 0 ? cfg.zchunks : v.zchunks;
          if (cfg.pace_ns == 0 && cfg.zchunks <= 0) *pace_ns = v.pace_ns;
          return D3H_OK;
        }
      }
      T = 1;
    }
    auto it = ctx->tuned.find(tune_key(dtype, T, fma, nx, ny, nz));
    if (it != ctx->tuned.end()) {
      const KernelEntry *e = find_entry(dtype, fma, it->second);
      if (e && e->T == T && (!e->rows || rows)) {
        *entry = e;
        *zchunks = cfg.zchunks > 0 ? cfg.zchunks : it->second.zchunks;
        if (cfg.pace_ns == 0 && cfg.zchunks <= 0) *pace_ns = it->second.pace_ns;
        return D3H_OK;
      }
    }
  }
  const KernelEntry *e = rows ? default_entry(dtype, T, fma, true) : nullptr;
  if (!e) e = default_entry(dtype, T, fma);
  if (!e) return fail(D3H_ERR_UNSUPPORTED, "no kernel for dtype %d depth %d fma %d", dtype, T, fma);
  *entry = e;
  *zchunks = cfg.zchunks;
  return D3H_OK;
}

int enqueue_steps(d3h_field *f, int count, const double coef[7], const KernelEntry &e, int zchunks,
                  int pace_ns, int sr, unsigned seed, hipStream_t stream, int *launches,
                  unsigned *blocks_out) {
  const int T = e.T;
  const int nfull = count / T, rem = count % T;
  Geometry g{}, g1{};
  unsigned nb = 0, nb1 = 0;
  if (e.rows && !rows_ok(f, coef))
    return fail(D3H_ERR_INVALID, "2D tile needs nz == 1 and ct == cb == 0");
  D3H_TRY(field_geometry(e, f, zchunks, &g, &nb));
  const KernelEntry *plain = nullptr;
  if (rem) {
    plain = default_entry(f->dtype, 1, e.fma, e.rows != 0);
    if (!plain) return fail(D3H_ERR_UNSUPPORTED, "no plain kernel for dtype %d", f->dtype);
    D3H_TRY(field_geometry(*plain, f, 0, &g1, &nb1));
  }
  g.sr = g1.sr = sr;
  g.seed = g1.seed = seed;
  g.pace = pace_ticks(f->ctx, e, pace_ns);  // the plain tail runs unscheduled
  // step numbers matter only under stochastic rounding; keeping them 0
  // otherwise leaves captured graphs reusable across calls
  const long long base = sr ? f->steps_done : 0;
  int n = 0;
  for (int i = 0; i < nfull; ++i, ++n) {
    g.step0 = (int)(base + (long long)i * T);
    launch_entry(e, f->buf[f->cur], f->buf[f->cur ^ 1], g, coef, nb, stream);
    f->cur ^= 1;
  }
  for (int i = 0; i < rem; ++i, ++n) {
    g1.step0 = (int)(base + (long long)nfull * T + i);
    launch_entry(*plain, f->buf[f->cur], f->buf[f->cur ^ 1], g1, coef, nb1, stream);
    f->cur ^= 1;
  }
  D3H_HIP(hipGetLastError());
  if (launches) *launches = n;
  if (blocks_out) *blocks_out = nb;
  return D3H_OK;
}

// ---- conversion ------------------------------------------------------------
__device__ __forceinline__ double load_elem(const void *p, int dtype, size_t i) {
  switch (dtype) {
    case D3H_F32: return (double)static_cast<const float *>(p)[i];
    case D3H_F64: return static_cast<const double *>(p)[i];
    case D3H_BF16: return (double)__uint_as_float((uint32_t) static_cast<const uint16_t *>(p)[i] << 16);
    default: return (double)f16_bits_to_f32(static_cast<const uint16_t *>(p)[i]);
  }
}
__device__ __forceinline__ void store_elem(void *p, int dtype, size_t i, double v) {
  switch (dtype) {
    case D3H_F32: static_cast<float *>(p)[i] = (float)v; break;
    case D3H_F64: static_cast<double *>(p)[i] = v; break;
    case D3H_BF16: static_cast<uint16_t *>(p)[i] = (uint16_t)to_bf16((float)v); break;
    default: static_cast<uint16_t *>(p)[i] = (uint16_t)to_f16((float)v); break;
  }
}

// Copies rows of nx cells; with ny > 0 the destination's rows of every plane
// are in reverse order (row y of plane z goes to row ny-1-y).
__global__ void convert_kernel(const void *src, int sdt, int sp, void *dst, int ddt, int dp, int nx,
                               long long rows, int ny_rev) {
  const long long n = rows * nx;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long row = i / nx, x = i - row * nx;
    long long drow = row;
    if (ny_rev > 0) {
      const long long z = row / ny_rev, y = row - z * ny_rev;
      drow = z * ny_rev + (ny_rev - 1 - y);
    }
    store_elem(dst, ddt, (size_t)(drow * dp + x), load_elem(src, sdt, (size_t)(row * sp + x)));
  }
}

int launch_convert(const void *src, int src_dtype, int src_pitch, void *dst, int dst_dtype,
                   int dst_pitch, int nx, int ny, int nz, hipStream_t stream, bool yrev) {
  const long long rows = (long long)ny * nz;
  const long long n = rows * nx;
  const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(convert_kernel, dim3(blocks), dim3(256), 0, stream, src, src_dtype,
                     src_pitch ? src_pitch : nx, dst, dst_dtype, dst_pitch ? dst_pitch : nx, nx, rows,
                     yrev ? ny : 0);
  D3H_HIP(hipGetLastError());
  return D3H_OK;
}

bool needs_staging(int host_dtype, int dtype, int pitch, int nx, bool yrev) {
  return host_dtype != dtype || pitch != nx || yrev;
}

int copy_in(const void *host, int host_dtype, void *dst, int dtype, int pitch, int nx, int ny, int nz,
            void *staging, hipStream_t stream, bool yrev) {
  const size_t n = (size_t)nx * ny * nz;
  if (!needs_staging(host_dtype, dtype, pitch, nx, yrev)) {
    D3H_HIP(hipMemcpyAsync(dst, host, n * dtype_bytes(dtype), hipMemcpyHostToDevice, stream));
    return D3H_OK;
  }
  if (!staging) return fail(D3H_ERR_STATE, "copy_in needs a staging buffer");
  D3H_HIP(hipMemcpyAsync(staging, host, n * dtype_bytes(host_dtype), hipMemcpyHostToDevice, stream));
  return launch_convert(staging, host_dtype, 0, dst, dtype, pitch, nx, ny, nz, stream, yrev);
}

int copy_out(const void *src, int dtype, int pitch, void *host, int host_dtype, int nx, int ny, int nz,
             void *staging, hipStream_t stream, bool yrev) {
  const size_t n = (size_t)nx * ny * nz;
  if (!needs_staging(host_dtype, dtype, pitch, nx, yrev)) {
    D3H_HIP(hipMemcpyAsync(host, src, n * dtype_bytes(dtype), hipMemcpyDeviceToHost, stream));
    return D3H_OK;
  }
  if (!staging) return fail(D3H_ERR_STATE, "copy_out needs a staging buffer");
  D3H_TRY(launch_convert(src, dtype, pitch, staging, host_dtype, 0, nx, ny, nz, stream, yrev));
  D3H_HIP(hipMemcpyAsync(host, staging, n * dtype_bytes(host_dtype), hipMemcpyDeviceToHost, stream));
  return D3H_OK;
}

// ---- FP64 reductions --------------------------------------------------------
constexpr int kRedThreads = 256;

template <int K>
__device__ __forceinline__ void block_reduce(double (&v)[K], double *part, bool last_is_max) {
  __shared__ double sh[K][kRedThreads];
#pragma unroll
  for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double o = sh[k][threadIdx.x + s];
        if (last_is_max && k == K - 1) sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], o);
        else sh[k][threadIdx.x] += o;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) part[blockIdx.x * K + k] = sh[k][0];
}

__global__ void __launch_bounds__(kRedThreads)
compare_kernel(const void *a, int adt, int ap, const void *b, int bdt, int bp, int nx, long long rows,
               double *part) {
  double v[3] = {0.0, 0.0, 0.0};  // sum d^2, sum b^2, max |d|
  const long long n = rows * nx;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long row = i / nx, x = i - row * nx;
    const double av = load_elem(a, adt, (size_t)(row * ap + x));
    const double bv = load_elem(b, bdt, (size_t)(row * bp + x));
    const double d = av - bv;
    v[0] += d * d;
    v[1] += bv * bv;
    v[2] = fmax(v[2], fabs(d));
  }
  block_reduce<3>(v, part, true);
}

// fx[i] = 1 - ax cos(kx x_i) etc. are precomputed per axis in FP64 on the host.
__global__ void __launch_bounds__(kRedThreads)
analytical_kernel(const void *f, int dt, int p, int nx, int ny, long long rows, const double *tab,
                  double *part) {
  double v[2] = {0.0, 0.0};  // sum d^2, max |d|
  const long long n = rows * nx;
  const double *fx = tab, *fy = tab + nx, *fz = tab + nx + ny;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long row = i / nx, x = i - row * nx;
    const long long z = row / ny, y = row - z * ny;
    const double ref = 0.125 * fx[x] * fy[y] * fz[z];
    const double d = load_elem(f, dt, (size_t)(row * p + x)) - ref;
    v[0] += d * d;
    v[1] = fmax(v[1], fabs(d));
  }
  block_reduce<2>(v, part, true);
}

static int ensure_partials(d3h_context *ctx) {
  if (ctx->partials && ctx->host_partials) return D3H_OK;
  const int nb = std::max(1, ctx->num_cus) * 4;
  void *dev = nullptr, *host = nullptr;
  if (hipMalloc(&dev, sizeof(double) * 4 * nb) != hipSuccess)
    return fail(D3H_ERR_NOMEM, "reduction scratch");
  if (hipHostMalloc(&host, sizeof(double) * 4 * nb, hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(dev);
    return fail(D3H_ERR_NOMEM, "pinned reduction scratch");
  }
  ctx->partials = static_cast<double *>(dev);
  ctx->host_partials = static_cast<double *>(host);
  ctx->partial_blocks = nb;
  return D3H_OK;
}

// ---- allocation-free analytical error norms -------------------------------------
// d3h_field_error_analytical_noalloc keeps everything it needs beyond the
// field and the caller's out[2] in this file-scope device array: the per-axis
// factors 1 - a cos(k x) (nx + ny + nz <= kNoAllocTab), the per-block partial
// sums and the two results.  Each device holds its own copy until the process
// ends; its address is looked up once per device by d3h_context_create
// (hipGetSymbolAddress, with the device current) and kept in g_scratch_dev.
constexpr int kNoAllocTab = 1 << 16;
constexpr int kNoAllocBlocks = 2048;
constexpr int kNoAllocDoubles = kNoAllocTab + 2 * kNoAllocBlocks + 2;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices];

// Once per device, with that device current.
static int init_noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return D3H_OK;  // the entry point refuses such devices
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return D3H_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(D3H_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  g_scratch_dev[device] = static_cast<double *>(ptr);
  return D3H_OK;
}

// tab[i] for i in [0, nx + ny + nz): 1 - amp[a] cos(k[a] (d[a] (j + 0.5))) of
// axis a and index j, as the host computes it for d3h_field_error_analytical.
__global__ void __launch_bounds__(256)
axis_factor_kernel(double *tab, int nx, int ny, int nz, double ax, double ay, double az, double kx,
                   double ky, double kz, double dx, double dy, double dz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx + ny + nz) return;
  double amp, k, d;
  int j;
  if (i < nx) {
    amp = ax; k = kx; d = dx; j = i;
  } else if (i < nx + ny) {
    amp = ay; k = ky; d = dy; j = i - nx;
  } else {
    amp = az; k = kz; d = dz; j = i - nx - ny;
  }
  tab[i] = 1.0 - amp * cos(k * (d * (j + 0.5)));
}

// One workgroup: out[0] = sqrt(sum of part[2i] / n), out[1] = max of part[2i+1].
__global__ void __launch_bounds__(kRedThreads)
rms_max_final_kernel(const double *part, int nb, double n, double *out) {
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += kRedThreads) {
    v[0] += part[2 * i];
    v[1] = fmax(v[1], part[2 * i + 1]);
  }
  __shared__ double fin[2];
  block_reduce<2>(v, fin, true);  // blockIdx.x == 0: writes fin[0], fin[1]
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = sqrt(fin[0] / n);
    out[1] = fin[1];
  }
}

// A smooth field for the autotuner to time on, the kind its callers advance:
// 0.125 * prod (1 - cos(2 pi (i + 0.5) / n_i)) over the three axes (the
// benchmarks' initial field).  The deep fused kernels draw enough power that
// the clock drops on such data (on an MI355X, depth 4 at 768^3 measured 14 %
// slower per launch on the cosine field than on a zero field, and the 16-bit
// tiles rank differently on uniform random values than on smooth fields), so
// timing on zeros or noise would rank tiles at clocks no real run sees.
__global__ void tune_fill_kernel(void *dst, int dtype, int pitch, int nx, int ny, int nz) {
  const long long n = (long long)nx * ny * nz;
  const float tw = 6.28318530717958647692f;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long row = i / nx;
    const int x = (int)(i - row * nx), y = (int)(row % ny), z = (int)(row / ny);
    const float fx = 1.0f - __cosf(tw * (x + 0.5f) / nx);
    const float fy = 1.0f - __cosf(tw * (y + 0.5f) / ny);
    const float fz = nz > 1 ? 1.0f - __cosf(tw * (z + 0.5f) / nz) : 1.0f;
    store_elem(dst, dtype, (size_t)(row * pitch + x), 0.125 * fx * fy * fz);
  }
}

// ---- HBM probe ----------------------------------------------------------------
// One 16-byte vector per thread, non-temporal load and store: the fastest
// streaming form measured on MI355X (the reference's "copy" variant computes
// f2 = cc * f1 the same way, diffusion/diffusion_cuda_copy.cu:8-24).
__global__ void __launch_bounds__(256) scale_copy_kernel(const u32x4 *__restrict__ in,
                                                         u32x4 *__restrict__ out, long long nvec,
                                                         float cc) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nvec) return;
  const u32x4 v = __builtin_nontemporal_load(in + i);
  u32x4 o;
  o.x = __float_as_uint(cc * __uint_as_float(v.x));
  o.y = __float_as_uint(cc * __uint_as_float(v.y));
  o.z = __float_as_uint(cc * __uint_as_float(v.z));
  o.w = __float_as_uint(cc * __uint_as_float(v.w));
  __builtin_nontemporal_store(o, out + i);
}

}  // namespace d3h

using namespace d3h;

static double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

extern "C" {

const char *d3h_status_string(int status) {
  switch (status) {
    case D3H_OK: return "ok";
    case D3H_ERR_INVALID: return "invalid argument";
    case D3H_ERR_HIP: return "HIP runtime error";
    case D3H_ERR_NOMEM: return "out of memory";
    case D3H_ERR_UNSUPPORTED: return "unsupported configuration";
    case D3H_ERR_STATE: return "invalid state";
    case D3H_ERR_NOT_FOUND: return "not found";
    case D3H_ERR_UNSTABLE: return "unstable coefficients";
    default: return "unknown status";
  }
}

const char *d3h_last_error(void) { return t_err.c_str(); }

void d3h_run_config_default(d3h_run_config *cfg) {
  if (!cfg) return;
  cfg->temporal_depth = 1;
  cfg->arith = D3H_ARITH_EXACT;
  cfg->tile = -1;
  cfg->zchunks = 0;
  cfg->use_graph = 0;
  cfg->rounding = D3H_ROUND_NEAREST;
  cfg->seed = 0;
  cfg->pace_ns = 0;
}

int d3h_context_create(int device, d3h_context **out) {
  if (!out) return fail(D3H_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  D3H_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(D3H_ERR_INVALID, "device %d of %d", device, ndev);
  D3H_HIP(hipSetDevice(device));
  auto *c = new d3h_context();
  c->device = device;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete c;
    return fail(D3H_ERR_HIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
  }
  c->num_cus = prop.multiProcessorCount;
  c->lds_per_block = (int)prop.sharedMemPerBlock;
  if (hipDeviceGetAttribute(&c->rtc_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
      c->rtc_khz < 0)
    c->rtc_khz = 0;
  {
    const int rc = init_noalloc_scratch(device);
    if (rc != D3H_OK) {
      delete c;
      return rc;
    }
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    d3h_context_destroy(c);
    return fail(D3H_ERR_HIP, "stream/event creation failed");
  }
  *out = c;
  return D3H_OK;
}

int d3h_context_destroy(d3h_context *c) {
  if (!c) return D3H_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->copy) (void)hipStreamSynchronize(c->copy);
  if (c->partials) (void)hipFree(c->partials);
  if (c->host_partials) (void)hipHostFree(c->host_partials);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->copy) (void)hipStreamDestroy(c->copy);
  delete c;
  return D3H_OK;
}

int d3h_device_info(d3h_context *ctx, int *num_cus, size_t *free_bytes, size_t *total_bytes,
                    int *lds_bytes) {
  if (!ctx) return fail(D3H_ERR_INVALID, "ctx is NULL");
  D3H_HIP(hipSetDevice(ctx->device));
  size_t fr = 0, tot = 0;
  D3H_HIP(hipMemGetInfo(&fr, &tot));
  if (num_cus) *num_cus = ctx->num_cus;
  if (free_bytes) *free_bytes = fr;
  if (total_bytes) *total_bytes = tot;
  if (lds_bytes) *lds_bytes = ctx->lds_per_block;
  return D3H_OK;
}

int d3h_host_alloc(size_t bytes, void **ptr) {
  if (!ptr) return fail(D3H_ERR_INVALID, "ptr is NULL");
  *ptr = nullptr;
  if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess)
    return fail(D3H_ERR_NOMEM, "hipHostMalloc(%zu) failed", bytes);
  return D3H_OK;
}

int d3h_host_free(void *ptr) {
  if (ptr) D3H_HIP(hipHostFree(ptr));
  return D3H_OK;
}

static void release_graph(d3h_field *f) {
  if (f->gexec) (void)hipGraphExecDestroy(f->gexec);
  f->gexec = nullptr;
  f->gkey.clear();
}

int d3h_field_create(d3h_context *ctx, int nx, int ny, int nz, int dtype, d3h_field **out) {
  if (!ctx || !out) return fail(D3H_ERR_INVALID, "ctx/out is NULL");
  *out = nullptr;
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(D3H_ERR_INVALID, "dims %dx%dx%d", nx, ny, nz);
  if (!dtype_valid(dtype)) return fail(D3H_ERR_INVALID, "dtype %d", dtype);
  D3H_HIP(hipSetDevice(ctx->device));
  auto *f = new d3h_field();
  f->ctx = ctx;
  f->nx = nx; f->ny = ny; f->nz = nz; f->dtype = dtype;
  f->pitch = pitch_elems(nx, dtype);
  f->plane = (size_t)f->pitch * ny;
  f->elems = f->plane * nz;
  f->bytes = f->elems * dtype_bytes(dtype);
  const size_t bytes = f->bytes;
  for (int i = 0; i < 2; ++i) {
    if (hipMalloc(&f->buf[i], bytes) != hipSuccess) {
      d3h_field_destroy(f);
      return fail(D3H_ERR_NOMEM, "hipMalloc(%zu) for a %dx%dx%d field failed", bytes, nx, ny, nz);
    }
    if (hipMemsetAsync(f->buf[i], 0, bytes, ctx->stream) != hipSuccess) {
      d3h_field_destroy(f);
      return fail(D3H_ERR_HIP, "hipMemsetAsync failed");
    }
  }
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    d3h_field_destroy(f);
    return fail(D3H_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
  }
  *out = f;
  return D3H_OK;
}

int d3h_field_destroy(d3h_field *f) {
  if (!f) return D3H_OK;
  (void)hipSetDevice(f->ctx->device);
  (void)hipStreamSynchronize(f->ctx->stream);
  release_graph(f);
  for (int i = 0; i < 2; ++i)
    if (f->buf[i]) (void)hipFree(f->buf[i]);
  if (f->staging) (void)hipFree(f->staging);
  delete f;
  return D3H_OK;
}

int d3h_field_set_y_reversed(d3h_field *f, int reversed) {
  if (!f) return fail(D3H_ERR_INVALID, "field is NULL");
  f->yrev = reversed != 0;
  return D3H_OK;
}

int d3h_field_dims(const d3h_field *f, int *nx, int *ny, int *nz, int *dtype) {
  if (!f) return fail(D3H_ERR_INVALID, "field is NULL");
  if (nx) *nx = f->nx;
  if (ny) *ny = f->ny;
  if (nz) *nz = f->nz;
  if (dtype) *dtype = f->dtype;
  return D3H_OK;
}

static int ensure_staging(d3h_field *f, size_t bytes) {
  if (f->staging_bytes >= bytes) return D3H_OK;
  if (f->staging) (void)hipFree(f->staging);
  f->staging = nullptr;
  f->staging_bytes = 0;
  if (hipMalloc(&f->staging, bytes) != hipSuccess)
    return fail(D3H_ERR_NOMEM, "staging hipMalloc(%zu) failed", bytes);
  f->staging_bytes = bytes;
  return D3H_OK;
}

int d3h_field_upload(d3h_field *f, const void *host, int host_dtype) {
  if (!f || !host) return fail(D3H_ERR_INVALID, "field/host is NULL");
  if (!dtype_valid(host_dtype)) return fail(D3H_ERR_INVALID, "host dtype %d", host_dtype);
  d3h_context *ctx = f->ctx;
  D3H_HIP(hipSetDevice(ctx->device));
  const size_t n = (size_t)f->nx * f->ny * f->nz;
  if (needs_staging(host_dtype, f->dtype, f->pitch, f->nx, f->yrev))
    D3H_TRY(ensure_staging(f, n * dtype_bytes(host_dtype)));
  D3H_TRY(copy_in(host, host_dtype, f->buf[f->cur], f->dtype, f->pitch, f->nx, f->ny, f->nz, f->staging,
                  ctx->stream, f->yrev));
  D3H_HIP(hipStreamSynchronize(ctx->stream));
  f->steps_done = 0;
  return D3H_OK;
}

int d3h_field_download(d3h_field *f, void *host, int host_dtype) {
  if (!f || !host) return fail(D3H_ERR_INVALID, "field/host is NULL");
  if (!dtype_valid(host_dtype)) return fail(D3H_ERR_INVALID, "host dtype %d", host_dtype);
  d3h_context *ctx = f->ctx;
  D3H_HIP(hipSetDevice(ctx->device));
  const size_t n = (size_t)f->nx * f->ny * f->nz;
  if (needs_staging(host_dtype, f->dtype, f->pitch, f->nx, f->yrev))
    D3H_TRY(ensure_staging(f, n * dtype_bytes(host_dtype)));
  D3H_TRY(copy_out(f->buf[f->cur], f->dtype, f->pitch, host, host_dtype, f->nx, f->ny, f->nz, f->staging,
                   ctx->stream, f->yrev));
  D3H_HIP(hipStreamSynchronize(ctx->stream));
  return D3H_OK;
}

int d3h_field_copy(d3h_field *dst, const d3h_field *src) {
  if (!dst || !src) return fail(D3H_ERR_INVALID, "field is NULL");
  if (dst == src) return D3H_OK;
  if (dst->ctx != src->ctx) return fail(D3H_ERR_INVALID, "fields of different contexts");
  if (dst->nx != src->nx || dst->ny != src->ny || dst->nz != src->nz || dst->dtype != src->dtype)
    return fail(D3H_ERR_INVALID, "copy between a %dx%dx%d field of dtype %d and a %dx%dx%d field of "
                "dtype %d", src->nx, src->ny, src->nz, src->dtype, dst->nx, dst->ny, dst->nz, dst->dtype);
  if (dst->yrev != src->yrev) return fail(D3H_ERR_INVALID, "fields store their rows in different orders");
  d3h_context *ctx = dst->ctx;
  D3H_HIP(hipSetDevice(ctx->device));
  D3H_HIP(hipMemcpyAsync(dst->buf[dst->cur], src->buf[src->cur], src->bytes, hipMemcpyDeviceToDevice,
                         ctx->stream));
  D3H_HIP(hipStreamSynchronize(ctx->stream));
  dst->steps_done = src->steps_done;
  return D3H_OK;
}

// ---- raw field files ------------------------------------------------------------
// 40-byte little-endian header ("D3HFIELD", version 1, dtype, nx, ny, nz,
// flags, steps since upload), then the dense field in host row order.
namespace {
struct RawHeader {
  char magic[8];
  uint32_t version, dtype;
  int32_t nx, ny, nz;
  uint32_t flags;  // bit 0: the field was stored with reversed rows
  uint64_t steps;
};
static_assert(sizeof(RawHeader) == 40, "raw header layout");
constexpr char kRawMagic[8] = {'D', '3', 'H', 'F', 'I', 'E', 'L', 'D'};
}  // namespace

int d3h_raw_write(const char *path, const void *host, int dtype, int nx, int ny, int nz, uint64_t steps,
                  int flags) {
  if (!path || !host) return fail(D3H_ERR_INVALID, "path/host is NULL");
  if (!dtype_valid(dtype) || nx <= 0 || ny <= 0 || nz <= 0)
    return fail(D3H_ERR_INVALID, "dtype %d dims %dx%dx%d", dtype, nx, ny, nz);
  RawHeader h{};
  memcpy(h.magic, kRawMagic, 8);
  h.version = 1;
  h.dtype = (uint32_t)dtype;
  h.nx = nx; h.ny = ny; h.nz = nz;
  h.flags = (uint32_t)flags;
  h.steps = steps;
  const size_t bytes = (size_t)nx * ny * nz * dtype_bytes(dtype);
  FILE *out = fopen(path, "wb");
  if (!out) return fail(D3H_ERR_INVALID, "cannot write %s", path);
  const bool ok = fwrite(&h, sizeof h, 1, out) == 1 && fwrite(host, 1, bytes, out) == bytes;
  if (fclose(out) != 0 || !ok) return fail(D3H_ERR_INVALID, "write to %s failed", path);
  return D3H_OK;
}

int d3h_raw_read(const char *path, void *host, size_t capacity, int *dtype, int *nx, int *ny, int *nz,
                 uint64_t *steps, int *flags) {
  if (!path) return fail(D3H_ERR_INVALID, "path is NULL");
  FILE *in = fopen(path, "rb");
  if (!in) return fail(D3H_ERR_NOT_FOUND, "cannot read %s", path);
  RawHeader h{};
  if (fread(&h, sizeof h, 1, in) != 1 || memcmp(h.magic, kRawMagic, 8) != 0 || h.version != 1 ||
      !dtype_valid((int)h.dtype) || h.nx <= 0 || h.ny <= 0 || h.nz <= 0) {
    fclose(in);
    return fail(D3H_ERR_INVALID, "%s is not a field file", path);
  }
  if (dtype) *dtype = (int)h.dtype;
  if (nx) *nx = h.nx;
  if (ny) *ny = h.ny;
  if (nz) *nz = h.nz;
  if (steps) *steps = h.steps;
  if (flags) *flags = (int)h.flags;
  const size_t bytes = (size_t)h.nx * h.ny * h.nz * dtype_bytes((int)h.dtype);
  if (!host) {  // header only
    fclose(in);
    return D3H_OK;
  }
  if (capacity < bytes) {
    fclose(in);
    return fail(D3H_ERR_INVALID, "%s needs %zu bytes, buffer has %zu", path, bytes, capacity);
  }
  const bool ok = fread(host, 1, bytes, in) == bytes;
  fclose(in);
  return ok ? D3H_OK : fail(D3H_ERR_INVALID, "%s is truncated", path);
}

int d3h_field_save(d3h_field *f, const char *path) {
  if (!f || !path) return fail(D3H_ERR_INVALID, "field/path is NULL");
  std::vector<char> host((size_t)f->nx * f->ny * f->nz * dtype_bytes(f->dtype));
  D3H_TRY(d3h_field_download(f, host.data(), f->dtype));
  return d3h_raw_write(path, host.data(), f->dtype, f->nx, f->ny, f->nz, (uint64_t)f->steps_done,
                       f->yrev ? 1 : 0);
}

int d3h_field_load(d3h_field *f, const char *path) {
  if (!f || !path) return fail(D3H_ERR_INVALID, "field/path is NULL");
  int dt = 0, nx = 0, ny = 0, nz = 0;
  uint64_t steps = 0;
  D3H_TRY(d3h_raw_read(path, nullptr, 0, &dt, &nx, &ny, &nz, &steps, nullptr));
  if (nx != f->nx || ny != f->ny || nz != f->nz)
    return fail(D3H_ERR_INVALID, "%s holds %dx%dx%d, the field is %dx%dx%d", path, nx, ny, nz, f->nx,
                f->ny, f->nz);
  std::vector<char> host((size_t)nx * ny * nz * dtype_bytes(dt));
  D3H_TRY(d3h_raw_read(path, host.data(), host.size(), nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr));
  D3H_TRY(d3h_field_upload(f, host.data(), dt));
  f->steps_done = (long long)steps;  // stochastic rounding continues the step numbers
  return D3H_OK;
}

int d3h_field_run(d3h_field *f, int count, const double coef[7], const d3h_run_config *cfg_in,
                  d3h_timing *timing) {
  const double t0 = now_ms();
  if (!f || !coef) return fail(D3H_ERR_INVALID, "field/coef is NULL");
  if (count < 0) return fail(D3H_ERR_INVALID, "count %d < 0", count);
  d3h_run_config cfg;
  d3h_run_config_default(&cfg);
  if (cfg_in) cfg = *cfg_in;
  d3h_context *ctx = f->ctx;
  D3H_HIP(hipSetDevice(ctx->device));
  const KernelEntry *e = nullptr;
  int zchunks = 0;
  if (cfg.rounding != D3H_ROUND_NEAREST && cfg.rounding != D3H_ROUND_STOCHASTIC)
    return fail(D3H_ERR_INVALID, "rounding %d", cfg.rounding);
  const int sr = cfg.rounding == D3H_ROUND_STOCHASTIC;
  if (sr && f->dtype != D3H_BF16 && f->dtype != D3H_F16)
    return fail(D3H_ERR_UNSUPPORTED, "stochastic rounding is for 16-bit storage (field dtype %d)",
                f->dtype);
  if (sr && f->steps_done + count > INT_MAX)
    return fail(D3H_ERR_INVALID, "step numbers of stochastic rounding exceed 2^31");
  int pace_ns = 0;
  D3H_TRY(resolve_entry(ctx, f->nx, f->ny, f->nz, f->dtype, cfg, rows_ok(f, coef), &e, &zchunks,
                        &pace_ns));
  if (pace_ticks(ctx, *e, pace_ns) == 0) pace_ns = 0;
  int launches = 0;
  unsigned blocks = 0;
  float ms = 0.f;
  if (count > 0) {
    if (cfg.use_graph) {
      std::vector<long long> key = {count, (long long)(intptr_t)e, zchunks, pace_ns, f->cur, sr,
                                    sr ? (long long)cfg.seed : 0, sr ? f->steps_done : 0};
      for (int i = 0; i < 7; ++i) {
        long long bits;
        memcpy(&bits, &coef[i], 8);
        key.push_back(bits);
      }
      if (key != f->gkey) {
        release_graph(f);
        const int cur0 = f->cur;
        hipGraph_t graph = nullptr;
        D3H_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_steps(f, count, coef, *e, zchunks, pace_ns, sr, cfg.seed, ctx->stream,
                               &launches, &blocks);
        hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
        f->cur = cur0;
        if (rc != D3H_OK) {
          if (graph) (void)hipGraphDestroy(graph);
          return rc;
        }
        if (ce != hipSuccess) return fail(D3H_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
        hipError_t ie = hipGraphInstantiate(&f->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return fail(D3H_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
        key.push_back(launches);
        key.push_back(blocks);
        f->gkey = key;
      }
      launches = (int)f->gkey[f->gkey.size() - 2];
      blocks = (unsigned)f->gkey.back();
      D3H_HIP(hipEventRecord(ctx->ev0, ctx->stream));
      D3H_HIP(hipGraphLaunch(f->gexec, ctx->stream));
      D3H_HIP(hipEventRecord(ctx->ev1, ctx->stream));
      if (launches & 1) f->cur ^= 1;
    } else {
      D3H_HIP(hipEventRecord(ctx->ev0, ctx->stream));
      D3H_TRY(enqueue_steps(f, count, coef, *e, zchunks, pace_ns, sr, cfg.seed, ctx->stream, &launches,
                            &blocks));
      D3H_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    }
    D3H_HIP(hipEventSynchronize(ctx->ev1));
    D3H_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    f->steps_done += count;
  }
  if (timing) {
    timing->kernel_ms = ms;
    timing->launches = launches;
    timing->avg_launch_ms = launches ? ms / launches : 0.0;
    timing->steps = count;
    timing->temporal_depth = e->T;
    timing->tile = (int)(e - registry().data());
    timing->pace_ns = pace_ns;
    Geometry g{};
    unsigned nb = 0;
    if (field_geometry(*e, f, zchunks, &g, &nb) == D3H_OK) {
      timing->zchunks = g.nzc;
      timing->blocks = (int)nb;
    }
    timing->wall_ms = now_ms() - t0;
  }
  return D3H_OK;
}

int d3h_field_compare(const d3h_field *a, const d3h_field *b, double out[3]) {
  if (!a || !b || !out) return fail(D3H_ERR_INVALID, "NULL argument");
  if (a->nx != b->nx || a->ny != b->ny || a->nz != b->nz)
    return fail(D3H_ERR_INVALID, "dims differ");
  d3h_context *ctx = a->ctx;
  D3H_HIP(hipSetDevice(ctx->device));
  D3H_TRY(ensure_partials(ctx));
  const int nb = ctx->partial_blocks;
  hipLaunchKernelGGL(compare_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream, a->buf[a->cur],
                     a->dtype, a->pitch, b->buf[b->cur], b->dtype, b->pitch, a->nx,
                     (long long)a->ny * a->nz, ctx->partials);
  D3H_HIP(hipGetLastError());
  D3H_HIP(hipMemcpyAsync(ctx->host_partials, ctx->partials, sizeof(double) * 3 * nb,
                         hipMemcpyDeviceToHost, ctx->stream));
  D3H_HIP(hipStreamSynchronize(ctx->stream));
  double d2 = 0, b2 = 0, mx = 0;
  for (int i = 0; i < nb; ++i) {
    d2 += ctx->host_partials[3 * i];
    b2 += ctx->host_partials[3 * i + 1];
    mx = std::max(mx, ctx->host_partials[3 * i + 2]);
  }
  out[0] = b2 > 0 ? std::sqrt(d2 / b2) : std::sqrt(d2);
  out[1] = mx;
  out[2] = std::sqrt(d2);
  return D3H_OK;
}

int d3h_field_error_analytical(const d3h_field *f, double kappa, double time, const double k[3],
                               const double d[3], double out[2]) {
  if (!f || !k || !d || !out) return fail(D3H_ERR_INVALID, "NULL argument");
  d3h_context *ctx = f->ctx;
  D3H_HIP(hipSetDevice(ctx->device));
  D3H_TRY(ensure_partials(ctx));
  const int dims[3] = {f->nx, f->ny, f->nz};
  std::vector<double> tab((size_t)f->nx + f->ny + f->nz);
  size_t o = 0;
  for (int a = 0; a < 3; ++a) {
    const double amp = std::exp(-kappa * time * k[a] * k[a]);
    for (int j = 0; j < dims[a]; ++j) tab[o++] = 1.0 - amp * std::cos(k[a] * (d[a] * (j + 0.5)));
  }
  double *dtab = nullptr;
  D3H_HIP(hipMalloc(&dtab, sizeof(double) * tab.size()));
  hipError_t e = hipMemcpy(dtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(dtab);
    return fail(D3H_ERR_HIP, "hipMemcpy: %s", hipGetErrorString(e));
  }
  const int nb = ctx->partial_blocks;
  hipLaunchKernelGGL(analytical_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream, f->buf[f->cur],
                     f->dtype, f->pitch, f->nx, f->ny, (long long)f->ny * f->nz, dtab, ctx->partials);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(ctx->host_partials, ctx->partials, sizeof(double) * 2 * nb,
                       hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dtab);
  if (e != hipSuccess) return fail(D3H_ERR_HIP, "analytical norm: %s", hipGetErrorString(e));
  double d2 = 0, mx = 0;
  for (int i = 0; i < nb; ++i) {
    d2 += ctx->host_partials[2 * i];
    mx = std::max(mx, ctx->host_partials[2 * i + 1]);
  }
  out[0] = std::sqrt(d2 / ((double)f->nx * f->ny * f->nz));
  out[1] = mx;
  return D3H_OK;
}

int d3h_field_error_analytical_noalloc(const d3h_field *f, double kappa, double time, const double k[3],
                                       const double d[3], double out[2]) {
  if (!f || !k || !d || !out) return fail(D3H_ERR_INVALID, "NULL argument");
  if ((long long)f->nx + f->ny + f->nz > kNoAllocTab)
    return fail(D3H_ERR_INVALID, "nx + ny + nz = %lld exceeds the %d factors of the static scratch",
                (long long)f->nx + f->ny + f->nz, kNoAllocTab);
  d3h_context *ctx = f->ctx;
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return fail(D3H_ERR_UNSUPPORTED, "device %d: the static scratch serves devices 0-%d", ctx->device,
                kMaxDevices - 1);
  D3H_HIP(hipSetDevice(ctx->device));
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = g_scratch_dev[ctx->device];
  }
  if (!base) return fail(D3H_ERR_STATE, "static scratch of device %d not set up", ctx->device);
  double *tab = base, *part = base + kNoAllocTab, *res = part + 2 * kNoAllocBlocks;
  const int ntab = f->nx + f->ny + f->nz;
  double amp[3];
  for (int a = 0; a < 3; ++a) amp[a] = std::exp(-kappa * time * k[a] * k[a]);
  hipLaunchKernelGGL(axis_factor_kernel, dim3((ntab + 255) / 256), dim3(256), 0, ctx->stream, tab, f->nx,
                     f->ny, f->nz, amp[0], amp[1], amp[2], k[0], k[1], k[2], d[0], d[1], d[2]);
  const int nb = std::min(std::max(1, ctx->num_cus) * 4, kNoAllocBlocks);
  hipLaunchKernelGGL(analytical_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream, f->buf[f->cur],
                     f->dtype, f->pitch, f->nx, f->ny, (long long)f->ny * f->nz, (const double *)tab, part);
  hipLaunchKernelGGL(rms_max_final_kernel, dim3(1), dim3(kRedThreads), 0, ctx->stream, (const double *)part,
                     nb, (double)f->nx * f->ny * f->nz, res);
  D3H_HIP(hipGetLastError());
  D3H_HIP(hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  D3H_HIP(hipStreamSynchronize(ctx->stream));
  return D3H_OK;
}

int d3h_tile_count(void) { return (int)registry().size(); }

int d3h_tile_info(int index, d3h_tile_desc *out) {
  if (!out) return fail(D3H_ERR_INVALID, "out is NULL");
  const auto &reg = registry();
  if (index < 0 || index >= (int)reg.size()) return fail(D3H_ERR_NOT_FOUND, "tile %d", index);
  const KernelEntry &e = reg[index];
  out->dtype = e.dtype;
  out->temporal_depth = e.T;
  out->arith = e.fma ? D3H_ARITH_FMA : D3H_ARITH_EXACT;
  out->groups = e.G;
  out->rows_per_wave = e.RPW;
  out->waves = e.NW;
  out->tile_x = e.tile_x();
  out->tile_y = e.tile_y();
  out->prefetch = e.pf;
  out->row_stream = e.rows;
  return D3H_OK;
}

// Time every tile of depth T (and its z-chunk candidates) on field f; the
// fastest per step goes to *bi / *bz / *best (bi < 0 when none ran), with its
// plane schedule in *bp (depth 1; 0 = none).
static int tune_depth(d3h_context *ctx, d3h_field *f, int T, int fma, const double coef[7], int *bi,
                      int *bz, double *best, int *bp) {
  const auto &reg = registry();
  const bool rows = rows_ok(f, coef);
  const long long cells = (long long)f->nx * f->ny * f->nz;
  // enough launches for ~4 ms of work at ~1 Tcell/s, at least 4; every
  // candidate is timed in kPasses interleaved passes and keeps its fastest,
  // so one noisy measurement cannot pick a tile that is slower in steady
  // state (tiles of one depth are often within 1-2 % of each other)
  constexpr int kPasses = 2;
  const int reps = (int)std::max(4LL, std::min(200LL, 4000000000LL / std::max(cells, 1LL) / T));
  // a 2D field that 2D kernels can run skips the 3D tiles (one plane per
  // workgroup: several times slower, and long to time on a large field)
  const bool only_2d = rows && default_entry(f->dtype, T, fma, true) != nullptr;
  struct Cand {
    int entry;
    Geometry g;
    unsigned blocks;
    double best;
    int pace_ns;
  };
  std::vector<Cand> cands;
  for (int i = 0; i < (int)reg.size(); ++i) {
    const KernelEntry &e = reg[i];
    if (e.dtype != f->dtype || e.T != T || e.fma != fma || (e.rows && !rows)) continue;
    if (only_2d && !e.rows) continue;
    const int occ = occupancy_of(e);
    if (occ <= 0) continue;
    Geometry g{};
    unsigned nb = 0;
    if (field_geometry(e, f, 0, &g, &nb) != D3H_OK) continue;
    // planes (rows, row blocks) streamed
    const int nstream = e.rows ? (f->ny + e.RPW - 1) / e.RPW : f->nz;
    // few long streams (1-4 chunks) up to several waves of workgroups; 2D
    // fields also try a quarter and an eighth of the default count, where
    // plane schedules gained most (16384^2: 2-row blocks in 16-23 chunks)
    std::vector<int> zcs = {1, 2, 3, 4, g.nzc, g.nzc * 2, g.nzc * 4, std::max(1, g.nzc / 2)};
    if (only_2d) {
      zcs.push_back(std::max(1, g.nzc / 4));
      zcs.push_back(std::max(1, g.nzc / 8));
    }
    std::sort(zcs.begin(), zcs.end());
    zcs.erase(std::unique(zcs.begin(), zcs.end()), zcs.end());
    int last_nzc = -1;
    for (int zc : zcs) {
      if (zc > nstream) continue;
      Geometry g2{};
      unsigned nb2 = 0;
      if (field_geometry(e, f, zc, &g2, &nb2) != D3H_OK) continue;
      if (g2.nzc == last_nzc) continue;  // same launch as the previous count
      last_nzc = g2.nzc;
      cands.push_back(Cand{i, g2, nb2, 1e300, 0});
    }
  }
  // 16-bit storage is timed with stochastic rounding, the mode this problem
  // needs (nearest even stalls its small per-step changes) and the costlier one
  if (f->dtype == D3H_BF16 || f->dtype == D3H_F16)
    for (Cand &c : cands) c.g.sr = 1;
  // The second pass times only the candidates within `keep` of the first
  // pass's fastest, which shortens tuning: launches right after a long tuning
  // run are slower (512^3 FP32: 1-3 % over the first 20 launches after a
  // tuner that timed four times as many scheduled candidates).
  auto time_all = [&](std::vector<Cand> &cs, double keep) -> int {
    for (int pass = 0; pass < kPasses; ++pass) {
      double lead = 1e300;
      for (const Cand &c : cs) lead = std::min(lead, c.best);
      for (Cand &c : cs) {
        if (pass > 0 && c.best > lead * keep) continue;
        const KernelEntry &e = reg[c.entry];
        launch_entry(e, f->buf[f->cur], f->buf[f->cur ^ 1], c.g, coef, c.blocks, ctx->stream);  // warm-up
        f->cur ^= 1;
        (void)hipEventRecord(ctx->ev0, ctx->stream);
        for (int r = 0; r < reps; ++r) {
          launch_entry(e, f->buf[f->cur], f->buf[f->cur ^ 1], c.g, coef, c.blocks, ctx->stream);
          f->cur ^= 1;
        }
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        if (hipEventSynchronize(ctx->ev1) != hipSuccess || hipGetLastError() != hipSuccess)
          return fail(D3H_ERR_HIP, "autotune launch of tile %d failed", c.entry);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
        c.best = std::min(c.best, (double)ms / reps / T);
      }
    }
    return D3H_OK;
  };
  D3H_TRY(time_all(cands, 1.08));
  *bp = 0;
  for (const Cand &c : cands) {
    if (c.best < *best) {
      *best = c.best;
      *bi = c.entry;
      *bz = c.g.nzc;
    }
  }
  // Depth 1: plane schedules (Geometry::pace) for the six fastest
  // configurations, at 0.86-0.99 of each one's measured time per z iteration
  // (launch time / (iterations x rounds of resident workgroups)).
  // A schedule that only the fastest workgroups reach keeps the streams of all
  // workgroups on nearby planes; its optimum is narrow (on the 512^3 FP32 step
  // about 0.93 of the unscheduled time per iteration for the 4-wave tile: 4-5 %
  // faster; 0.97 for the 1-wave tile: under 1 %), and a slower schedule only
  // loses.  One is kept only when it beats the best unscheduled time by more
  // than 0.5 %.
  if (T == 1 && ctx->rtc_khz > 0 && !cands.empty()) {
    std::vector<Cand> top = cands;
    std::sort(top.begin(), top.end(), [](const Cand &a, const Cand &b) { return a.best < b.best; });
    if (top.size() > 6) top.resize(6);
    std::vector<Cand> paced;
    for (const Cand &c : top) {
      const KernelEntry &e = reg[c.entry];
      const int halo = e.rows ? (T + e.RPW - 1) / e.RPW : T;
      const int iters = c.g.oz + T + halo + std::max(e.pf, 1);
      // a launch with more workgroups than fit at once runs them in rounds
      const long long slots = (long long)std::max(occupancy_of(e), 1) * std::max(ctx->num_cus, 1);
      const long long rounds = std::max(1LL, ((long long)c.blocks + slots - 1) / slots);
      const double ns_iter = c.best * 1e6 / ((double)iters * rounds);
      for (int pct = 86; pct <= 99; ++pct) {
        Cand p = c;
        p.best = 1e300;
        const double ns = ns_iter * pct / 100.0;
        if (ns > kMaxPaceNs) break;  // a tune table refuses such schedules
        p.pace_ns = (int)ns;
        p.g.pace = pace_ticks(ctx, e, p.pace_ns);
        if (p.g.pace > 0) paced.push_back(p);
      }
    }
    D3H_TRY(time_all(paced, 1.03));
    const double unscheduled = *best;
    for (const Cand &p : paced) {
      if (p.best < unscheduled * 0.995 && p.best < *best) {
        *best = p.best;
        *bi = p.entry;
        *bz = p.g.nzc;
        *bp = p.pace_ns;
      }
    }
  }
  return D3H_OK;
}

int d3h_autotune(d3h_context *ctx, int nx, int ny, int nz, int dtype, int temporal_depth, int arith,
                 int *best_tile, int *best_zchunks, double *best_ms_per_step) {
  if (!ctx) return fail(D3H_ERR_INVALID, "ctx is NULL");
  const int D = temporal_depth;
  if (D != 0 && D != 1 && D != 2 && D != 4 && D != 8)
    return fail(D3H_ERR_INVALID, "temporal depth %d (0 = choose, 1, 2, 4 or 8)", D);
  const int fma = arith == D3H_ARITH_FMA ? 1 : 0;
  d3h_field *f = nullptr;
  D3H_TRY(d3h_field_create(ctx, nx, ny, nz, dtype, &f));
  {
    const long long rows = (long long)ny * nz, n = rows * nx;
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 65536);
    for (int b = 0; b < 2; ++b)
      hipLaunchKernelGGL(tune_fill_kernel, dim3(blocks), dim3(256), 0, ctx->stream, f->buf[b], dtype,
                         f->pitch, nx, ny, nz);
    if (hipGetLastError() != hipSuccess) {
      d3h_field_destroy(f);
      return fail(D3H_ERR_HIP, "autotune field fill failed");
    }
  }
  // nz == 1 fields are tuned for the 2D operator (ct = cb = 0), which also
  // admits the 2D tiles
  const bool rows = nz == 1;
  const double coef[7] = {0.1, 0.1, 0.1, 0.1, rows ? 0.0 : 0.1, rows ? 0.0 : 0.1, rows ? 0.6 : 0.4};
  const auto &reg = registry();
  std::vector<int> depths;
  // depth 8 is searched only for 2D fields: in 3D the register file fits
  // only small depth-8 tiles, whose 2 x 8 halo rows leave few stored rows
  // (768^3: 981 Gcell/s effective against about 1790 at depth 4 and 743 for
  // plain steps); explicit depth-8 runs stay available
  if (D == 0) depths = nz == 1 ? std::vector<int>{1, 2, 4, 8} : std::vector<int>{1, 2, 4};
  else depths = {D};
  double best_all = 1e300;
  int bi_all = -1, bz_all = 0, bp_all = 0;
  int rc = D3H_OK;
  for (int T : depths) {
    double best = 1e300;
    int bi = -1, bz = 0, bp = 0;
    rc = tune_depth(ctx, f, T, fma, coef, &bi, &bz, &best, &bp);
    if (rc != D3H_OK) break;
    if (bi < 0) continue;
    const KernelEntry &e = reg[bi];
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      ctx->tuned[tune_key(dtype, T, fma, nx, ny, nz)] =
          TuneVal{e.vec, e.G, e.RPW, e.NW, e.pf, bz, best, T, e.rows, bp};
    }
    if (best < best_all) {
      best_all = best;
      bi_all = bi;
      bz_all = bz;
      bp_all = bp;
    }
  }
  d3h_field_destroy(f);
  if (rc != D3H_OK) return rc;
  if (bi_all < 0) return fail(D3H_ERR_UNSUPPORTED, "no kernel for dtype %d depth %d", dtype, D);
  if (D == 0) {
    const KernelEntry &e = reg[bi_all];
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tuned[tune_key(dtype, 0, fma, nx, ny, nz)] =
        TuneVal{e.vec, e.G, e.RPW, e.NW, e.pf, bz_all, best_all, e.T, e.rows, bp_all};
  }
  if (best_tile) *best_tile = bi_all;
  if (best_zchunks) *best_zchunks = bz_all;
  if (best_ms_per_step) *best_ms_per_step = best_all;
  return D3H_OK;
}

int d3h_autotune_save(d3h_context *ctx, const char *path) {
  if (!ctx || !path) return fail(D3H_ERR_INVALID, "NULL argument");
  std::ofstream os(path);
  if (!os) return fail(D3H_ERR_INVALID, "cannot write %s", path);
  os << "# dtype depth fma log2nx log2ny log2nz vec groups rows_per_wave waves prefetch zchunks "
        "ms_per_step entry_depth batch rows pace_ns   (depth 0 = the depth chosen for the size "
        "class; batch 1 = chosen for batched sweeps; rows 1 = a 2D entry; pace_ns = the plane "
        "schedule of depth-1 launches, 0 = none)\n";
  std::lock_guard<std::mutex> lk(ctx->mu);
  for (const auto &kv : ctx->tuned) {
    const auto &k = kv.first;
    const auto &v = kv.second;
    os << k.dtype << ' ' << k.T << ' ' << k.fma << ' ' << k.cx << ' ' << k.cy << ' ' << k.cz << ' '
       << v.vec << ' ' << v.G << ' ' << v.RPW << ' ' << v.NW << ' ' << v.pf << ' ' << v.zchunks << ' '
       << v.ms_per_step << ' ' << v.T << ' ' << k.batch << ' ' << v.rows << ' ' << v.pace_ns << '\n';
  }
  return os ? D3H_OK : fail(D3H_ERR_INVALID, "write to %s failed", path);
}

int d3h_autotune_load(d3h_context *ctx, const char *path) {
  if (!ctx || !path) return fail(D3H_ERR_INVALID, "NULL argument");
  std::ifstream is(path);
  if (!is) return fail(D3H_ERR_NOT_FOUND, "cannot read %s", path);
  std::string line;
  std::lock_guard<std::mutex> lk(ctx->mu);
  while (std::getline(is, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    TuneKey k{};
    TuneVal v{};
    if (!(ss >> k.dtype >> k.T >> k.fma >> k.cx >> k.cy >> k.cz >> v.vec >> v.G >> v.RPW >> v.NW >>
          v.pf >> v.zchunks >> v.ms_per_step))
      return fail(D3H_ERR_INVALID, "malformed line in %s: %s", path, line.c_str());
    // entry depth (absent in tables without depth-0 lines: the key's depth),
    // the batch flag (absent in tables without batched choices: 0), the 2D
    // flag (absent in tables written before 2D entries were told apart: 0),
    // then the plane schedule (absent in tables written before schedules: 0)
    if (!(ss >> v.T)) v.T = k.T;
    else if (!(ss >> k.batch)) k.batch = 0;
    else if (!(ss >> v.rows)) v.rows = 0;
    else if (!(ss >> v.pace_ns)) v.pace_ns = 0;
    if (v.pace_ns < 0 || v.pace_ns > kMaxPaceNs)
      return fail(D3H_ERR_INVALID, "bad plane schedule in %s: %s", path, line.c_str());
    if (k.batch != 0 && k.batch != 1)
      return fail(D3H_ERR_INVALID, "bad batch flag in %s: %s", path, line.c_str());
    if (v.rows != 0 && v.rows != 1)
      return fail(D3H_ERR_INVALID, "bad 2D flag in %s: %s", path, line.c_str());
    if (v.T <= 0 || (k.T != 0 && v.T != k.T))
      return fail(D3H_ERR_INVALID, "bad entry depth in %s: %s", path, line.c_str());
    // entries no longer compiled are skipped
    if (find_entry(k.dtype, k.fma, v)) ctx->tuned[k] = v;
  }
  return D3H_OK;
}

int d3h_bandwidth_probe(d3h_context *ctx, size_t n, int dtype, int reps, double *gbps, double *avg_ms) {
  if (!ctx) return fail(D3H_ERR_INVALID, "ctx is NULL");
  if (!dtype_valid(dtype) || reps <= 0 || n == 0) return fail(D3H_ERR_INVALID, "bad probe arguments");
  D3H_HIP(hipSetDevice(ctx->device));
  const size_t bytes = (n * dtype_bytes(dtype) + 15) / 16 * 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return fail(D3H_ERR_NOMEM, "probe alloc");
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail(D3H_ERR_NOMEM, "probe alloc");
  }
  (void)hipMemsetAsync(a, 0, bytes, ctx->stream);
  const long long nvec = (long long)(bytes / 16);
  if ((nvec + 255) / 256 > INT_MAX) {
    (void)hipFree(a);
    (void)hipFree(b);
    return fail(D3H_ERR_INVALID, "probe size too large");
  }
  const unsigned blocks = (unsigned)((nvec + 255) / 256);
  hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks), dim3(256), 0, ctx->stream, (const u32x4 *)a,
                     (u32x4 *)b, nvec, 0.4f);
  (void)hipEventRecord(ctx->ev0, ctx->stream);
  for (int r = 0; r < reps; ++r) {
    hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                       (const u32x4 *)((r & 1) ? b : a), (u32x4 *)((r & 1) ? a : b), nvec, 0.4f);
  }
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  hipError_t e = hipEventSynchronize(ctx->ev1);
  if (e == hipSuccess) e = hipGetLastError();
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(D3H_ERR_HIP, "probe: %s", hipGetErrorString(e));
  const double avg = ms / reps;
  if (avg_ms) *avg_ms = avg;
  if (gbps) *gbps = 2.0 * (double)bytes / (avg * 1e-3) * 1e-9;
  return D3H_OK;
}

// ---- Physis-shaped drop-in ----------------------------------------------------
static struct DropIn {
  d3h_context *ctx = nullptr;
  d3h_field *field = nullptr;
  int nx = 0, ny = 0, nz = 0;
  float kernel_seconds = 0.f;
} g_drop;

static int env_int(const char *name, int dflt) {
  const char *v = getenv(name);
  return (v && *v) ? atoi(v) : dflt;
}

// Knobs of the drop-in path (environment):
//   D3H_DEVICE          device index (default 0)
//   D3H_TEMPORAL_DEPTH  1 (default), 2, 4, 8, or 0 = the tuned depth
//   D3H_ARITH           0 EXACT (default), 1 FMA
//   D3H_USE_GRAPH       1 = replay a captured hipGraph
//   D3H_TUNE_FILE       tuning table loaded at initialize_hip when it exists
//   D3H_AUTOTUNE        1 = tune this grid at initialize_benchmark_hip (for
//                       D3H_TEMPORAL_DEPTH, or every depth when it is 0) and
//                       append the result to D3H_TUNE_FILE when set
int initialize_hip(int argc, char **argv, int nx, int ny, int nz) {
  (void)argc;
  (void)argv;
  if (g_drop.ctx) return fail(D3H_ERR_STATE, "initialize_hip called twice");
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(D3H_ERR_INVALID, "dims %dx%dx%d", nx, ny, nz);
  D3H_TRY(d3h_context_create(env_int("D3H_DEVICE", 0), &g_drop.ctx));
  g_drop.nx = nx; g_drop.ny = ny; g_drop.nz = nz;
  const char *tf = getenv("D3H_TUNE_FILE");
  if (tf && *tf) {
    std::ifstream probe(tf);
    if (probe) D3H_TRY(d3h_autotune_load(g_drop.ctx, tf));
  }
  return D3H_OK;
}

int initialize_benchmark_hip(int nx, int ny, int nz) {
  if (!g_drop.ctx) return fail(D3H_ERR_STATE, "initialize_hip not called");
  if (nx != g_drop.nx || ny != g_drop.ny || nz != g_drop.nz)
    return fail(D3H_ERR_INVALID, "dims differ from initialize_hip");
  if (g_drop.field) return fail(D3H_ERR_STATE, "benchmark already initialised");
  if (env_int("D3H_AUTOTUNE", 0)) {
    D3H_TRY(d3h_autotune(g_drop.ctx, nx, ny, nz, D3H_F32, env_int("D3H_TEMPORAL_DEPTH", 1),
                         env_int("D3H_ARITH", D3H_ARITH_EXACT), nullptr, nullptr, nullptr));
    const char *tf = getenv("D3H_TUNE_FILE");
    if (tf && *tf) D3H_TRY(d3h_autotune_save(g_drop.ctx, tf));
  }
  return d3h_field_create(g_drop.ctx, nx, ny, nz, D3H_F32, &g_drop.field);
}

int run_kernel_hip(int count, float *f1_host, int nx, int ny, int nz, float ce, float cw, float cn,
                   float cs, float ct, float cb, float cc) {
  if (!g_drop.field) return fail(D3H_ERR_STATE, "initialize_benchmark_hip not called");
  if (!f1_host) return fail(D3H_ERR_INVALID, "f1_host is NULL");
  if (nx != g_drop.nx || ny != g_drop.ny || nz != g_drop.nz)
    return fail(D3H_ERR_INVALID, "dims differ from initialize_hip");
  const double coef[7] = {ce, cw, cn, cs, ct, cb, cc};
  d3h_run_config cfg;
  d3h_run_config_default(&cfg);
  cfg.temporal_depth = env_int("D3H_TEMPORAL_DEPTH", 1);
  cfg.arith = env_int("D3H_ARITH", D3H_ARITH_EXACT);
  cfg.use_graph = env_int("D3H_USE_GRAPH", 0);
  D3H_TRY(d3h_field_upload(g_drop.field, f1_host, D3H_F32));
  d3h_timing t{};
  D3H_TRY(d3h_field_run(g_drop.field, count, coef, &cfg, &t));
  g_drop.kernel_seconds = (float)(t.kernel_ms * 1e-3);
  return d3h_field_download(g_drop.field, f1_host, D3H_F32);
}

int finalize_benchmark_hip(void) {
  d3h_field_destroy(g_drop.field);
  g_drop.field = nullptr;
  d3h_context_destroy(g_drop.ctx);
  g_drop.ctx = nullptr;
  return D3H_OK;
}

int kernel_time_hip(float *seconds) {
  if (!seconds) return fail(D3H_ERR_INVALID, "seconds is NULL");
  *seconds = g_drop.kernel_seconds;
  return D3H_OK;
}

// By-reference form of the reference's Fortran variant.  One context and
// field are kept for the process and re-created when the dimensions change;
// calls must not overlap.
static struct ByRef {
  d3h_context *ctx = nullptr;
  d3h_field *field = nullptr;
  int nx = 0, ny = 0, nz = 0;
} g_byref;

int diffusion_hip_(float *f1, float *f2, int *nx, int *ny, int *nz, float *ce, float *cw, float *cn,
                   float *cs, float *ct, float *cb, float *cc, int *count) {
  (void)f2;
  if (!f1 || !nx || !ny || !nz || !ce || !cw || !cn || !cs || !ct || !cb || !cc || !count)
    return fail(D3H_ERR_INVALID, "NULL argument");
  if (*count < 0) return fail(D3H_ERR_INVALID, "count %d < 0", *count);
  if (!g_byref.ctx) D3H_TRY(d3h_context_create(env_int("D3H_DEVICE", 0), &g_byref.ctx));
  if (g_byref.field && (g_byref.nx != *nx || g_byref.ny != *ny || g_byref.nz != *nz)) {
    d3h_field_destroy(g_byref.field);
    g_byref.field = nullptr;
  }
  if (!g_byref.field) {
    D3H_TRY(d3h_field_create(g_byref.ctx, *nx, *ny, *nz, D3H_F32, &g_byref.field));
    g_byref.nx = *nx; g_byref.ny = *ny; g_byref.nz = *nz;
  }
  const double coef[7] = {*ce, *cw, *cn, *cs, *ct, *cb, *cc};
  d3h_run_config cfg;
  d3h_run_config_default(&cfg);
  cfg.temporal_depth = env_int("D3H_TEMPORAL_DEPTH", 1);
  cfg.arith = env_int("D3H_ARITH", D3H_ARITH_EXACT);
  cfg.use_graph = env_int("D3H_USE_GRAPH", 0);
  D3H_TRY(d3h_field_upload(g_byref.field, f1, D3H_F32));
  D3H_TRY(d3h_field_run(g_byref.field, *count, coef, &cfg, nullptr));
  return d3h_field_download(g_byref.field, f1, D3H_F32);
}

}  // extern "C"
