/*
 * diffusion3d_hip.h -- C ABI of the MI355X (gfx950) 3D 7-point diffusion engine.
 *
 * Plain C types only (no HIP or torch types), every entry point returns an int
 * status (D3H_OK == 0) and never asserts, exits or throws; d3h_last_error()
 * gives the calling thread's message for the last failure.
 *
 * Data layout seen by callers: a dense nx*ny*nz array, x fastest,
 * OFFSET3D(i,j,k) = i + j*nx + k*nx*ny (diffusion3d/diffusion3d.h:33-34 of the
 * reference).  Boundaries are clamped by index (diffusion3d/baseline.cc:33-38).
 * Coefficients are always given in the reference's run_kernel order
 * ce, cw, cn, cs, ct, cb, cc (diffusion3d/diffusion3d_physis.h:13-15) with
 * cn applied to y-1 and cs to y+1 (diffusion3d/baseline.cc:35-40).
 */
#ifndef DIFFUSION3D_HIP_H_
#define DIFFUSION3D_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
  D3H_OK = 0,
  D3H_ERR_INVALID = 1,      /* bad argument (dims, count, dtype, NULL, ...)   */
  D3H_ERR_HIP = 2,          /* a HIP runtime call failed                      */
  D3H_ERR_NOMEM = 3,        /* device or pinned host allocation failed        */
  D3H_ERR_UNSUPPORTED = 4,  /* no kernel for this dtype / depth / tile        */
  D3H_ERR_STATE = 5,        /* call out of order (not initialised, busy, ...) */
  D3H_ERR_NOT_FOUND = 6,    /* unknown request / handle id                    */
  D3H_ERR_UNSTABLE = 7      /* coefficients violate cc >= 0 (scheme unstable) */
};

/* Storage types.  FP32 and FP64 compute in their own precision; BF16 and
 * FP16 are storage formats with FP32 arithmetic, rounded to storage after
 * every step (D3H_ROUND_*). */
enum { D3H_F32 = 0, D3H_F64 = 1, D3H_BF16 = 2, D3H_F16 = 3 };

/* Arithmetic modes.  EXACT: the reference's 7-term expression with every
 * multiply and add rounded separately, in the reference's order -- bitwise
 * equal to the reference CPU build; the parity mode.  FMA: the same
 * expression contracted the way a contracting compiler does it,
 * fma(cc,c,cw*w) then a chain of fmas.  FMA is NOT a conforming mode: its
 * rounding drifts systematically from the reference CPU path (1.04e-6
 * relative L2 after 100 steps at 128^3, above the 1e-6 parity bound), though
 * it stays closer to the exact PDE solution. */
enum { D3H_ARITH_EXACT = 0, D3H_ARITH_FMA = 1 };

/* Rounding of 16-bit storage.  NEAREST: round to nearest even after every
 * step.  STOCHASTIC: round up or down with probability given by the distance
 * to the two neighbours -- unbiased, so the small per-step changes of this
 * problem are not lost to rounding, and reproducible bitwise.  Cell (x, y, z)
 * at step t uses 16 bits of h = lowbias32(row * 0x9E3779B1 + t * 0x85EBCA77 +
 * seed) + (x / 2) * 0x9E3779B9 with row = z*ny + y (the high half for even x,
 * the low half for odd x), applied by gfx950's v_cvt_sr_bf16_f32 /
 * v_cvt_sr_f16_f32 (FP16: subnormals round stochastically too; inf and nan
 * round to nearest).  Step numbers count from the field's last upload (or
 * the source's count after d3h_field_copy / d3h_field_load), from 0 in every
 * d3h_batch_run, and from a server request's first step. */
enum { D3H_ROUND_NEAREST = 0, D3H_ROUND_STOCHASTIC = 1 };

/* ---- drop-in variant ABI ---------------------------------------------------
 * The four Physis entry points of the reference
 * (diffusion3d/diffusion3d_physis.h:7-17), same argument meaning, plus a
 * kernel-only timer for DisplayResult's "Kernel-only performance" block
 * (diffusion3d/diffusion3d_cuda.cu:373-384).  run_kernel_hip advances f1_host
 * by exactly `count` steps (H2D, launches, D2H; synchronous), like
 * Diffusion3DCUDA::RunKernel (diffusion3d/diffusion3d_cuda.cu:353-371).
 * Environment: D3H_DEVICE, D3H_TEMPORAL_DEPTH (1, 2, 4, 8; 0 = tuned),
 * D3H_ARITH (0 EXACT, 1 FMA), D3H_USE_GRAPH (1 = hipGraph replay),
 * D3H_TUNE_FILE (tuning table loaded by initialize_hip when it exists) and
 * D3H_AUTOTUNE (1 = tune the grid in initialize_benchmark_hip and save the
 * table to D3H_TUNE_FILE when set). */
int initialize_hip(int argc, char **argv, int nx, int ny, int nz);
int initialize_benchmark_hip(int nx, int ny, int nz);
int run_kernel_hip(int count, float *f1_host, int nx, int ny, int nz,
                   float ce, float cw, float cn, float cs,
                   float ct, float cb, float cc);
int finalize_benchmark_hip(void);
/* Kernel-only seconds of the last run_kernel_hip (events around the launches). */
int kernel_time_hip(float *seconds);
/* By-reference entry in the shape of the reference's Fortran variant
 * (diffusion3d/diffusion3d_fortran.h:9-12; called as
 * diffusion_fortran_(f1_, f2_, &nx_, ..., &ce_, ..., &cc_, &count) from
 * Diffusion3DFortran::RunKernel).  f1 is advanced by exactly *count steps
 * (the Fortran variant runs 2 * (count / 2)); f2 is not touched.  The context
 * and device field are kept between calls (re-created when the dimensions
 * change); calls must not overlap.  Same environment knobs as run_kernel_hip. */
int diffusion_hip_(float *f1, float *f2, int *nx, int *ny, int *nz, float *ce, float *cw,
                   float *cn, float *cs, float *ct, float *cb, float *cc, int *count);

/* ---- errors ---------------------------------------------------------------- */
const char *d3h_status_string(int status);
const char *d3h_last_error(void);

/* ---- context ---------------------------------------------------------------
 * One context per host thread and device (the reference is single-threaded on
 * the default stream).  The context owns a compute stream, a copy stream,
 * events, the autotune table and scratch for reductions. */
typedef struct d3h_context d3h_context;
int d3h_context_create(int device, d3h_context **out);
int d3h_context_destroy(d3h_context *ctx);
/* Device facts: CU count, free and total HBM bytes, LDS bytes per workgroup. */
int d3h_device_info(d3h_context *ctx, int *num_cus, size_t *free_bytes,
                    size_t *total_bytes, int *lds_bytes);

/* Pinned host memory (hipHostMalloc).  The reference pins with cudaMallocHost
 * (diffusion3d/diffusion3d_cuda.cu:331). */
int d3h_host_alloc(size_t bytes, void **ptr);
int d3h_host_free(void *ptr);

/* ---- device-resident fields --------------------------------------------------
 * A field owns two device buffers (ping-pong), rows padded to 16 bytes. */
typedef struct d3h_field d3h_field;

typedef struct {
  int temporal_depth;  /* steps fused per launch: 1 (plain), 2, 4 or 8; 0 =
                          the depth d3h_autotune(..., 0, ...) chose for the
                          grid's size class, else 1                           */
  int arith;           /* D3H_ARITH_EXACT or D3H_ARITH_FMA                     */
  int tile;            /* kernel/tile index (d3h_tile_info); -1 = tuned choice */
  int zchunks;         /* z-chunks per launch; 0 = the tuned / default value  */
  int use_graph;       /* 1: capture the launch sequence in a hipGraph         */
  int rounding;        /* D3H_ROUND_*; STOCHASTIC needs a BF16 or FP16 field   */
  unsigned int seed;   /* seed of stochastic rounding                          */
  int pace_ns;         /* plane schedule of depth-1 launches, ns per z
                          iteration: 0 = the tuned one (tuned tile and
                          z-chunks only), -1 = none, 1..1000000 = this one.
                          A schedule only holds workgroups back; results are
                          the same with any schedule.  No counterpart in the
                          reference.                                         */
} d3h_run_config;

typedef struct {
  double kernel_ms;       /* event time around the launch sequence            */
  double wall_ms;         /* host wall clock of the whole call                */
  double avg_launch_ms;   /* kernel_ms / launches                             */
  int launches;           /* kernel launches issued                           */
  int steps;              /* steps advanced (== count)                        */
  int temporal_depth;     /* depth used by the fused launches                 */
  int tile;               /* kernel/tile index used                           */
  int zchunks;            /* z-chunks used                                    */
  int blocks;             /* workgroups per launch                            */
  int pace_ns;            /* plane schedule used (0: none)                    */
} d3h_timing;

void d3h_run_config_default(d3h_run_config *cfg);

int d3h_field_create(d3h_context *ctx, int nx, int ny, int nz, int dtype,
                     d3h_field **out);
int d3h_field_destroy(d3h_field *f);
int d3h_field_dims(const d3h_field *f, int *nx, int *ny, int *nz, int *dtype);
/* Copy a dense host array in; host_dtype may differ from the field's dtype
 * (converted on the device, round-to-nearest-even when narrowing). */
int d3h_field_upload(d3h_field *f, const void *host, int host_dtype);
int d3h_field_download(d3h_field *f, void *host, int host_dtype);
/* Device-to-device copy of src's current contents into dst (same dims, dtype
 * and row order; different fields of one context); dst also takes src's step
 * count, so stochastic rounding continues src's step numbers.  Synchronous.
 * Restores a field without a host round trip (no counterpart in the
 * reference, whose variants re-initialise on the host). */
int d3h_field_copy(d3h_field *dst, const d3h_field *src);
/* Raw field files: a 40-byte header ("D3HFIELD", version, dtype, nx, ny, nz,
 * flags, steps since upload) and the dense field in the field's dtype and the
 * host's row order -- bit-exact, unlike the reference's "%.7E" text dump
 * (diffusion3d/baseline.cc:62-70).  load needs a field of the same
 * dimensions (the stored dtype is converted like an upload) and restores the
 * step count, so stochastic rounding resumes where it stopped. */
int d3h_field_save(d3h_field *f, const char *path);
int d3h_field_load(d3h_field *f, const char *path);
/* The same file format for host arrays (e.g. the driver's --dump-raw).
 * d3h_raw_read with host == NULL reads the header only; otherwise it needs
 * capacity >= nx*ny*nz*size(dtype).  Outputs may be NULL.  flags bit 0
 * records reversed rows. */
int d3h_raw_write(const char *path, const void *host, int dtype, int nx, int ny, int nz,
                  uint64_t steps, int flags);
int d3h_raw_read(const char *path, void *host, size_t capacity, int *dtype, int *nx, int *ny,
                 int *nz, uint64_t *steps, int *flags);
/* Store the rows of every plane in reverse order on the device (uploads and
 * downloads flip them).  The kernels sum the y+1 term before the y-1 term as
 * diffusion3d does (diffusion3d/baseline.cc:39-40); the diffusion/ family
 * names y the other way round and sums y-1 first (diffusion/baseline.h:59-67),
 * so its results are reproduced bitwise on a reversed field.  The stencil is
 * symmetric in y, so only the rounding order changes.  The flag applies to
 * later uploads and downloads (set it before the first upload); device-side
 * comparisons and norms see the device order. */
int d3h_field_set_y_reversed(d3h_field *f, int reversed);
/* Advance the field by exactly `count` steps on the device (synchronous). */
int d3h_field_run(d3h_field *f, int count, const double coef[7],
                  const d3h_run_config *cfg, d3h_timing *timing);
/* FP64 error norms on the device: out[0] = ||a-b||_2 / ||b||_2,
 * out[1] = max |a-b|, out[2] = ||a-b||_2.  Same dims required. */
int d3h_field_compare(const d3h_field *a, const d3h_field *b, double out[3]);
/* RMS (FP64 accumulation) of the field against the analytical cosine product
 * at time t, evaluated in FP64 on the device:
 * 0.125 * prod_i (1 - exp(-kappa t k_i^2) cos(k_i x_i)), x_i = d_i (j + 0.5)
 * (diffusion3d/diffusion3d.cc:14-39).  out[0] = RMS, out[1] = max abs. */
int d3h_field_error_analytical(const d3h_field *f, double kappa, double time,
                               const double k[3], const double d[3], double out[2]);
/* The same norms, computed without allocating anything: no counterpart in
 * the reference; for callers that cannot allocate device or pinned memory
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  d3h_field_error_analytical allocates its per-axis table on
 * every call and its reduction scratch on a context's first call; this
 * variant keeps the table, the per-workgroup partial sums and the results in
 * a static device array of 557 KB per device (set up by d3h_context_create)
 * and reduces on the device, so it calls no hipMalloc/hipFree.  Limit:
 * nx + ny + nz <= 65536 (D3H_ERR_INVALID beyond), devices 0-63.  Every call
 * on one device shares that device's array: calls that overlap on one device
 * (other contexts or host threads) race and must be serialised by the
 * caller.  Results equal d3h_field_error_analytical's to rounding of the FP64
 * reductions (the factors are evaluated with the device's cos). */
int d3h_field_error_analytical_noalloc(const d3h_field *f, double kappa, double time,
                                       const double k[3], const double d[3], double out[2]);

/* Kernel/tile table (the tuning space of diffusion3d/build.py:6-34, searched
 * at run time).  Returns D3H_ERR_NOT_FOUND past the end. */
typedef struct {
  int dtype, temporal_depth, arith;
  int groups;          /* 16-byte vectors per lane along x                  */
  int rows_per_wave;   /* rows each wavefront owns                          */
  int waves;           /* wavefronts per workgroup                          */
  int tile_x, tile_y;  /* loaded tile extent in cells / rows               */
  int prefetch;        /* planes requested ahead of the one being computed;
                          0: the three-slot ring of fused depths (the next
                          plane requested after the first stage, every
                          level keeping 3 planes)                           */
  int row_stream;      /* 1: 2D kernel, streams the rows of nz == 1 fields in
                          blocks of rows_per_wave rows; such a
                          tile runs only fields with nz == 1 and ct == cb == 0,
                          and is chosen for them when no tile is given      */
} d3h_tile_desc;
int d3h_tile_count(void);
int d3h_tile_info(int index, d3h_tile_desc *out);

/* Autotuner: time every matching tile (and z-chunk count) on a scratch field of
 * the given shape, remember the fastest for that size class; the table can be
 * saved to / loaded from a text file.  best_tile / best_zchunks may be NULL.
 * temporal_depth 0 searches depths 1, 2 and 4 (and 8 for nz == 1 fields,
 * whose 2D kernels gain from it; 3D depth-8 tiles are far slower than depth
 * 4) and also remembers the fastest depth for the class, used by runs,
 * batches and servers given depth 0; best_tile then names the chosen depth's
 * tile (d3h_tile_info).  Candidates are timed in two interleaved passes (the
 * second only for those within 8 % of the first pass's fastest), each keeping
 * its faster time; at depth 1 the six fastest also get plane schedules
 * (d3h_run_config.pace_ns), and the tuned schedule is remembered with the
 * tile.  BF16/FP16 storage is timed with stochastic rounding. */
int d3h_autotune(d3h_context *ctx, int nx, int ny, int nz, int dtype,
                 int temporal_depth, int arith, int *best_tile, int *best_zchunks,
                 double *best_ms_per_step);
int d3h_autotune_save(d3h_context *ctx, const char *path);
int d3h_autotune_load(d3h_context *ctx, const char *path);

/* ---- bandwidth probe ---------------------------------------------------------
 * Streams f2 = cc * f1 over n elements of `dtype` `reps` times and reports the
 * average GB/s (read + write bytes), the measured copy roofline
 * (the reference's "copy" variant, diffusion/diffusion_cuda_copy.cu:8-24). */
int d3h_bandwidth_probe(d3h_context *ctx, size_t n, int dtype, int reps,
                        double *gbps, double *avg_ms);

/* ---- host-side analytical field ------------------------------------------------
 * The constructor's constants and the initial/analytical field of the
 * reference's benchmark class, computed on the host exactly as the reference
 * does (diffusion3d/diffusion3d.h:54-66, diffusion3d/diffusion3d.cc:14-39), so
 * inputs are byte-identical to its CPU path.  constants[15] = kappa, dx, dy,
 * dz, kx, ky, kz, dt, ce, cw, cn, cs, ct, cb, cc. */
int d3h_constants_f32(int nx, int ny, int nz, float constants[15]);
int d3h_constants_f64(int nx, int ny, int nz, double constants[15]);
int d3h_initialize_f32(float *buf, int nx, int ny, int nz,
                       const float constants[15], float time);
int d3h_initialize_f64(double *buf, int nx, int ny, int nz,
                       const double constants[15], double time);
/* 2D form of the diffusion/ family (nz = 1, no t/b terms): constants of
 * Diffusion(2, {nx, ny}) (diffusion/diffusion.h:93-108; ct = cb = 0 and
 * cc = 1 - (ce + cw + cn + cs)) and the 2D cosine product
 * (diffusion/diffusion.h:48-77 with ndim = 2).  Advance such a field with
 * d3h_field_run on an nx x ny x 1 field (y reversed) and these coefficients. */
int d3h_constants2d_f32(int nx, int ny, float constants[15]);
int d3h_initialize2d_f32(float *buf, int nx, int ny, const float constants[15], float time);
/* Cosine product plus counter-based noise U(-amp, amp) keyed by (seed, index). */
int d3h_add_noise_f32(float *buf, size_t n, uint64_t seed, float amp);

/* ---- batched uneven grids ------------------------------------------------------
 * Many independent problems advanced together: per size class one kernel
 * launch covers every tile of every member, T steps per launch, and the
 * whole step sequence is captured once in a hipGraph and replayed. */
typedef struct d3h_batch d3h_batch;
typedef struct {
  int nx, ny, nz;
  int dtype;        /* D3H_F32 / D3H_BF16 / D3H_F16 / D3H_F64 */
  int steps;        /* exact number of steps for this problem */
  double coef[7];   /* ce cw cn cs ct cb cc                    */
  int rounding;     /* D3H_ROUND_*; STOCHASTIC: BF16 or FP16   */
  unsigned int seed;
} d3h_problem;

/* Host-only checks (no device needed).  d3h_problem_check: dims, dtype,
 * steps >= 0, coefficients finite, >= 0 and summing to 1 (D3H_ERR_UNSTABLE
 * otherwise), rounding valid for the dtype.  d3h_problem_footprint: device
 * bytes a server request needs (two pitched buffers + conversion staging),
 * which admission compares with the free part of the server's pool. */
int d3h_problem_check(const d3h_problem *p);
int d3h_problem_footprint(const d3h_problem *p, int host_dtype, int out_dtype, size_t *bytes);
int d3h_batch_create(d3h_context *ctx, const d3h_problem *problems, int count,
                     d3h_batch **out);
int d3h_batch_destroy(d3h_batch *b);
/* Upload problem i's initial field (dense host array of host_dtype). */
int d3h_batch_upload(d3h_batch *b, int i, const void *host, int host_dtype);
int d3h_batch_download(d3h_batch *b, int i, void *host, int host_dtype);
/* Every problem at once: hosts[i] is problem i's dense host array (all of
 * host_dtype).  Copies and conversions are queued together and the call
 * synchronises once; with pinned buffers (d3h_host_alloc) the copies overlap.
 * Keeps a staging area of the batch's total dense size. */
int d3h_batch_upload_all(d3h_batch *b, const void *const *hosts, int host_dtype);
int d3h_batch_download_all(d3h_batch *b, void *const *hosts, int host_dtype);
/* Run every problem to its step count.  arith as above; temporal_depth 1, 2,
 * 4 or 8, or 0 = the depth d3h_autotune(..., 0, ...) chose for each problem's
 * size class, else 2 (FP32/FP64 storage) or 1 (BF16/FP16 storage);
 * use_graph 1 = capture + replay.  Stochastic rounding numbers the steps of
 * every call from 0: two consecutive runs of N steps reuse the same random
 * bits per step and differ from one field run of 2N steps (which continues
 * its step numbers); upload a problem again before running it again. */
int d3h_batch_run(d3h_batch *b, int arith, int temporal_depth, int use_graph,
                  d3h_timing *timing);
/* Batch autotuner (the reference's tile sweep, diffusion3d/build.py:6-44,
 * redone at run time for batched sweeps): for each size class in the batch
 * (dtype and ceil log2 of each dimension), times depths 1, 2 and 4 with every
 * batched tile, then z-chunk counts 1, 2, 4 and 8 for the fastest, each as an
 * 8-step sweep of the whole batch (the other classes keep their choice), and
 * keeps the fastest in the context's tune table as batch entries: depth-0
 * batch runs and servers on this context use them, d3h_autotune_save/load
 * keep them.  It runs the kernels on the batch's buffers, so it leaves their
 * contents undefined: call it before uploading.  classes, ms_before and
 * ms_after (may be NULL) get the number of classes tuned and the 8-step sweep
 * time before and after. */
int d3h_batch_autotune(d3h_batch *b, int arith, int *classes, double *ms_before,
                       double *ms_after);

/* ---- serving ----------------------------------------------------------------------
 * A request server over one device: requests (grid + steps + initial field)
 * are admitted as device memory and slots allow, kept resident, batched with
 * whatever else is running (continuous batching at sweep boundaries), and
 * copied back when done.  An admitted request joins the sweeps once its own
 * upload has completed, so running requests are not held up by it; when more
 * cells are still uploading than are ready, the next sweep waits for those
 * uploads instead.  Submission and polling are thread-safe. */
typedef struct d3h_server d3h_server;
typedef struct {
  size_t hbm_budget_bytes;  /* device bytes the server may hold; 0 = 90% of free */
  int max_active;           /* max concurrently resident requests; 0 = 4096     */
  int arith;                /* D3H_ARITH_EXACT / D3H_ARITH_FMA                   */
  int temporal_depth;       /* steps per sweep launch: 1, 2 or 4; 0 = the tuned
                               depth of each request's size class, else 2
                               (FP32/FP64) or 1 (BF16/FP16)                    */
  int sweep_steps;          /* steps per scheduling round; 0 = 8                 */
} d3h_server_config;

int d3h_server_create(d3h_context *ctx, const d3h_server_config *cfg,
                      d3h_server **out);
int d3h_server_destroy(d3h_server *s);
/* host_in: dense initial field of host_dtype; host_out: where the result is
 * written (dtype out_dtype) when the request completes.  Both must stay valid
 * until the request is done.  Returns the request id in *id. */
int d3h_server_submit(d3h_server *s, const d3h_problem *p, const void *host_in,
                      int host_dtype, void *host_out, int out_dtype, int64_t *id);
/* done = 1 when finished; latency_ms from submit to result-in-host-memory.
 * Polling never releases the request's record. */
int d3h_server_poll(d3h_server *s, int64_t id, int *done, double *latency_ms);
/* Blocks until the request is finished.  A successful wait releases the
 * request's record: later calls with its id return D3H_ERR_NOT_FOUND, so wait
 * exactly once per request to keep a long-running server's memory bounded.
 * A failed request keeps its record (and reports its error) until the server
 * is destroyed. */
int d3h_server_wait(d3h_server *s, int64_t id, double *latency_ms);
int d3h_server_stats(d3h_server *s, int64_t *completed, int64_t *rejected,
                     int64_t *sweeps, double *busy_ms);

#ifdef __cplusplus
}
#endif

#endif /* DIFFUSION3D_HIP_H_ */
