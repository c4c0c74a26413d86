// Internal runtime types shared by runtime.hip, batch.hip and server.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "diffusion3d_hip.h"
#include "kernels.hpp"

namespace d3h {

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define D3H_HIP(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::d3h::fail(D3H_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                              \
  } while (0)

#define D3H_TRY(expr)             \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != D3H_OK) return rc_; \
  } while (0)

size_t dtype_bytes(int dtype);
bool dtype_valid(int dtype);
// Row pitch in elements: nx rounded up to a multiple of 16 bytes.
int pitch_elems(int nx, int dtype);

// All compiled kernels.
const std::vector<KernelEntry> &registry();
// Default entry for (dtype, T, fma) (first in table order), or nullptr.
// rows selects the 2D entries (stencil2d_kernel) instead of the 3D tiles.
const KernelEntry *default_entry(int dtype, int T, int fma, bool rows = false);

// Geometry and block count of one launch of `e` over a grid; zchunks 0 =
// heuristic from occupancy and CU count.
int make_geometry(const KernelEntry &e, int nx, int ny, int nz, int pitch, int zchunks,
                  int num_cus, int occupancy, Geometry *g, unsigned *blocks);
int occupancy_of(const KernelEntry &e);

struct TuneKey {
  int dtype, T, fma, cx, cy, cz;
  int batch;  // 1: chosen for batched sweeps (d3h_batch_autotune), 0: single grids
  bool operator<(const TuneKey &o) const {
    return std::tie(dtype, T, fma, cx, cy, cz, batch) <
           std::tie(o.dtype, o.T, o.fma, o.cx, o.cy, o.cz, o.batch);
  }
};
struct TuneVal {
  int vec, G, RPW, NW, pf;  // entry identity (stable across builds)
  int zchunks;
  double ms_per_step;
  int T;                    // depth of the entry; under a depth-0 key, the tuned depth
  int rows;                 // 1: a 2D entry (stencil2d_kernel); 3D tiles and 2D
                            // entries can share every other parameter
  int pace_ns = 0;          // depth 1: plane schedule, ns per z iteration (0: none)
};
// Size class of a grid: ceil(log2) of each dimension.  Depth 0 keys hold the
// depth the autotuner chose for the class (d3h_autotune with depth 0).
TuneKey tune_key(int dtype, int T, int fma, int nx, int ny, int nz, int batch = 0);
// The compiled entry with these parameters, or nullptr.  rows selects 2D
// entries or 3D tiles (a 3D tile can have exactly a 2D entry's parameters).
const KernelEntry *find_entry(int dtype, int T, int fma, int vec, int G, int RPW, int NW, int pf,
                              int rows);
// The entry a tune-table value names (its depth, vec, G, RPW, NW, pf, rows).
const KernelEntry *find_entry(int dtype, int fma, const TuneVal &v);
// The tuned depth of a size class, or `dflt` when the class has none.
int tuned_depth(d3h_context *ctx, int dtype, int fma, int nx, int ny, int nz, int dflt);

// Host <-> pitched device field.  One dense 1-D copy when the dtypes match and
// rows are unpadded; otherwise a dense 1-D copy through `staging` (at least
// nx*ny*nz elements of the host dtype) and an on-device conversion/repitch.
// Strided 2-D copies from pageable memory are avoided: they run row by row.
// yrev: the device field holds the rows of every plane in reverse order (the
// host's row y is device row ny-1-y).
bool needs_staging(int host_dtype, int dtype, int pitch, int nx, bool yrev = false);
int copy_in(const void *host, int host_dtype, void *dst, int dtype, int pitch, int nx, int ny, int nz,
            void *staging, hipStream_t stream, bool yrev = false);
int copy_out(const void *src, int dtype, int pitch, void *host, int host_dtype, int nx, int ny, int nz,
             void *staging, hipStream_t stream, bool yrev = false);
}  // namespace d3h

struct d3h_context {
  int device = 0;
  hipStream_t stream = nullptr;  // compute
  hipStream_t copy = nullptr;    // host transfers (serving)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int num_cus = 0;
  int lds_per_block = 0;
  int rtc_khz = 0;                  // real-time counter rate (0: no plane schedules)
  double *partials = nullptr;       // device scratch for reductions
  double *host_partials = nullptr;  // pinned mirror
  int partial_blocks = 0;
  std::mutex mu;                    // guards `tuned`
  std::map<d3h::TuneKey, d3h::TuneVal> tuned;
};

struct d3h_field {
  d3h_context *ctx = nullptr;
  int nx = 0, ny = 0, nz = 0, dtype = 0, pitch = 0;
  size_t plane = 0, elems = 0, bytes = 0;
  void *buf[2] = {nullptr, nullptr};
  int cur = 0;
  void *staging = nullptr;
  size_t staging_bytes = 0;
  bool yrev = false;  // rows stored in reverse order (see d3h_field_set_y_reversed)
  long long steps_done = 0;  // steps since the last upload (stochastic rounding)
  // cached captured launch sequence
  hipGraphExec_t gexec = nullptr;
  std::vector<long long> gkey;
};

namespace d3h {
// True when a field and coefficients can run on the 2D entries: nz == 1 and
// ct == cb == 0 (the 2D 5-point operator).
bool rows_ok(const d3h_field *f, const double coef[7]);
// Resolve the kernel for a run configuration (tuned / explicit / default);
// 2D entries are chosen only when `rows` (see rows_ok).  *pace_ns: the plane
// schedule of depth-1 launches (cfg.pace_ns, else the tuned one; 0: none).
int resolve_entry(d3h_context *ctx, int nx, int ny, int nz, int dtype, const d3h_run_config &cfg,
                  bool rows, const KernelEntry **entry, int *zchunks, int *pace_ns);
// Geometry::pace of a plane schedule of pace_ns per z iteration on ctx's
// device (0 when pace_ns <= 0, the entry is fused or the device reports no
// real-time counter rate).
int pace_ticks(const d3h_context *ctx, const KernelEntry &e, int pace_ns);
// Launch geometry of entry `e` on field `f` (2D entries stream the rows of
// the nz == 1 field in blocks of RPW rows).
int field_geometry(const KernelEntry &e, const d3h_field *f, int zchunks, Geometry *g,
                   unsigned *blocks);
// Enqueue `count` steps from f->buf[f->cur] on `stream` (no sync); updates cur.
// sr / seed: stochastic rounding of BF16 storage, step numbers continuing from
// f->steps_done (the caller advances steps_done).
int enqueue_steps(d3h_field *f, int count, const double coef[7], const KernelEntry &e, int zchunks,
                  int pace_ns, int sr, unsigned seed, hipStream_t stream, int *launches,
                  unsigned *blocks);
// Convert nx*ny*nz cells between device arrays of two dtypes; a row pitch of
// 0 means dense (pitch = nx).  Narrowing rounds to nearest even (from f64 to
// bf16/f16 it rounds to f32 first).
int launch_convert(const void *src, int src_dtype, int src_pitch, void *dst, int dst_dtype,
                   int dst_pitch, int nx, int ny, int nz, hipStream_t stream, bool yrev = false);
}  // namespace d3h
