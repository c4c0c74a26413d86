// Batched execution of many independent grids of uneven sizes, and the sweep
// planner it shares with the request server (plan.hpp).
//
// A batch owns one device pool holding both buffers of every problem.  A run
// is one sweep over all problems (each by its own step count); its launches
// are captured once into a hipGraph (classes on separate streams become
// independent graph branches) and the graph is replayed on later runs.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

#include "plan.hpp"

using namespace d3h;

namespace d3h {

int check_problem(const d3h_problem &p) {
  if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0) return fail(D3H_ERR_INVALID, "dims %dx%dx%d", p.nx, p.ny, p.nz);
  if (!dtype_valid(p.dtype)) return fail(D3H_ERR_INVALID, "dtype %d", p.dtype);
  if (p.steps < 0) return fail(D3H_ERR_INVALID, "steps %d", p.steps);
  double sum = 0;
  for (int i = 0; i < 7; ++i) {
    if (!(p.coef[i] >= 0.0) || !std::isfinite(p.coef[i]))
      return fail(D3H_ERR_UNSTABLE, "coefficient %d = %g (all must be >= 0)", i, p.coef[i]);
    sum += p.coef[i];
  }
  if (std::fabs(sum - 1.0) > 1e-5) return fail(D3H_ERR_UNSTABLE, "coefficients sum to %.9g, not 1", sum);
  if (p.rounding != D3H_ROUND_NEAREST && p.rounding != D3H_ROUND_STOCHASTIC)
    return fail(D3H_ERR_INVALID, "rounding %d", p.rounding);
  if (p.rounding == D3H_ROUND_STOCHASTIC && p.dtype != D3H_BF16 && p.dtype != D3H_F16)
    return fail(D3H_ERR_UNSUPPORTED, "stochastic rounding is for 16-bit storage (dtype %d)", p.dtype);
  return D3H_OK;
}

const KernelEntry *pick_batch_entry(d3h_context *ctx, int nx, int ny, int nz, int dtype, int T, int fma,
                                    int *zchunks) {
  if (zchunks) *zchunks = 0;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (int batch : {1, 0}) {
      auto it = ctx->tuned.find(tune_key(dtype, T, fma, nx, ny, nz, batch));
      if (it == ctx->tuned.end()) continue;
      const TuneVal &v = it->second;
      const KernelEntry *e = find_entry(dtype, T, fma, v.vec, v.G, v.RPW, v.NW, v.pf, 0);
      if (e && e->launch_batch) {
        if (zchunks && batch) *zchunks = v.zchunks;
        return e;
      }
    }
  }
  const KernelEntry *best = nullptr;
  for (const auto &e : registry()) {
    if (!e.launch_batch || e.dtype != dtype || e.T != T || e.fma != fma) continue;
    if (occupancy_of(e) <= 0) continue;
    if (!best) {
      best = &e;
      continue;
    }
    const bool fits = e.tile_x() >= nx, bfits = best->tile_x() >= nx;
    if (fits && (!bfits || e.tile_x() < best->tile_x() ||
                 (e.tile_x() == best->tile_x() && e.tile_y() > best->tile_y())))
      best = &e;
    else if (!fits && !bfits && e.tile_x() > best->tile_x())
      best = &e;
  }
  return best;
}

int batch_depth(d3h_context *ctx, int dtype, int fma, int nx, int ny, int nz, int dflt) {
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tuned.find(tune_key(dtype, 0, fma, nx, ny, nz, 1));
    if (it != ctx->tuned.end()) return it->second.T;
  }
  return tuned_depth(ctx, dtype, fma, nx, ny, nz, dflt);
}

// Chunk depth for batched members: classes fill the device together, so a
// member only needs enough z-chunks to keep its halo planes a small overhead.
static int batch_zchunks(int nz, int T) { return std::max(1, nz / std::max(32, 8 * T)); }

int plan_sweep(d3h_context *ctx, const std::vector<PlanMember> &ms, int T, int fma, SweepPlan *plan) {
  const int n = (int)ms.size();
  std::vector<const KernelEntry *> ea(n, nullptr), eb(n, nullptr);
  std::vector<int> qa(n, 0), rb(n, 0), ta(n, 1), za(n, 0), zb(n, 0);
  for (int i = 0; i < n; ++i) {
    const PlanMember &m = ms[i];
    // depth 0: the depth the autotuners chose for the member's size class,
    // else 2 for FP32/FP64 storage and 1 for 16-bit storage (whose depth-1
    // kernels already keep the VALU pipe 75 % busy, so fused steps become
    // VALU-bound and their halo work costs more than the traffic saved)
    const int dflt = dtype_bytes(m.dtype) >= 4 ? 2 : 1;
    int Tm = T != 0 ? T : batch_depth(ctx, m.dtype, fma, m.nx, m.ny, m.nz, dflt);
    const KernelEntry *e = pick_batch_entry(ctx, m.nx, m.ny, m.nz, m.dtype, Tm, fma, &za[i]);
    if (!e && Tm != 1) {
      Tm = 1;
      e = pick_batch_entry(ctx, m.nx, m.ny, m.nz, m.dtype, 1, fma, &za[i]);
    }
    if (!e) return fail(D3H_ERR_UNSUPPORTED, "no batched kernel for dtype %d", m.dtype);
    ea[i] = e;
    ta[i] = Tm;
    qa[i] = m.steps / Tm;
    rb[i] = m.steps % Tm;
    if (rb[i]) {
      eb[i] = pick_batch_entry(ctx, m.nx, m.ny, m.nz, m.dtype, 1, fma, &zb[i]);
      if (!eb[i]) return fail(D3H_ERR_UNSUPPORTED, "no batched plain kernel for dtype %d", m.dtype);
    }
  }
  auto make = [&](const std::vector<const KernelEntry *> &ent, const std::vector<int> &cnt,
                  const std::vector<int> &zc, std::vector<ClassPlan> &out, bool tail) -> int {
    std::vector<const KernelEntry *> uniq;
    for (int i = 0; i < n; ++i)
      if (ent[i] && cnt[i] > 0 && std::find(uniq.begin(), uniq.end(), ent[i]) == uniq.end())
        uniq.push_back(ent[i]);
    for (const KernelEntry *e : uniq) {
      ClassPlan c;
      c.e = e;
      for (int i = 0; i < n; ++i)
        if (ent[i] == e && cnt[i] > 0) c.members.push_back(i);
      std::stable_sort(c.members.begin(), c.members.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
      int blocks = 0;
      c.prefix_blocks.push_back(0);
      for (int i : c.members) {
        const PlanMember &pm = ms[i];
        BatchMember m{};
        m.buf[0] = pm.buf[0];
        m.buf[1] = pm.buf[1];
        unsigned nb = 0;
        const int nzc = zc[i] > 0 ? zc[i] : batch_zchunks(pm.nz, e->T);
        D3H_TRY(make_geometry(*e, pm.nx, pm.ny, pm.nz, pm.pitch, nzc, ctx->num_cus, occupancy_of(*e),
                              &m.geo, &nb));
        for (int k = 0; k < 7; ++k) m.coef[k] = pm.coef[k];
        m.block_start = blocks;
        m.parity0 = tail ? (pm.parity0 + qa[i]) & 1 : pm.parity0;
        // the kernel numbers the steps of launch k from step0 + k * T
        m.geo.sr = pm.sr;
        m.geo.seed = pm.seed;
        m.geo.step0 = pm.step_base + (tail ? qa[i] * ta[i] : 0);
        c.table.push_back(m);
        c.launches.push_back(cnt[i]);
        for (unsigned k = 0; k < nb; ++k) c.block_member.push_back((int)c.table.size() - 1);
        blocks += (int)nb;
        c.prefix_blocks.push_back(blocks);
      }
      out.push_back(std::move(c));
    }
    return D3H_OK;
  };
  plan->A.clear();
  plan->B.clear();
  D3H_TRY(make(ea, qa, za, plan->A, false));
  D3H_TRY(make(eb, rb, zb, plan->B, true));
  size_t total = 0;
  int launches = 0;
  for (auto *cs : {&plan->A, &plan->B}) {
    for (auto &c : *cs) {
      total = (total + 255) / 256 * 256;
      c.table_off = total;
      total += c.table.size() * sizeof(BatchMember);
      total = (total + 255) / 256 * 256;
      c.block_off = total;
      total += c.block_member.size() * sizeof(int);
      launches += c.launches.empty() ? 0 : c.launches[0];
    }
  }
  plan->table_bytes = total;
  plan->launches = launches;
  plan->parity_after.assign(n, 0);
  for (int i = 0; i < n; ++i) plan->parity_after[i] = (ms[i].parity0 + qa[i] + rb[i]) & 1;
  return D3H_OK;
}

void plan_fill(const SweepPlan &plan, char *host) {
  for (const auto *cs : {&plan.A, &plan.B})
    for (const auto &c : *cs) {
      memcpy(host + c.table_off, c.table.data(), c.table.size() * sizeof(BatchMember));
      memcpy(host + c.block_off, c.block_member.data(), c.block_member.size() * sizeof(int));
    }
}

int plan_enqueue(std::vector<ClassPlan> &cs, void *dev_tables, hipStream_t main,
                 const std::vector<hipStream_t> &streams, const std::vector<hipEvent_t> &events,
                 int *launches) {
  if (cs.empty()) return D3H_OK;
  const int ns = (int)std::min<size_t>(cs.size(), streams.size() + 1);
  if (ns > 1) D3H_HIP(hipEventRecord(events[0], main));
  for (int s = 1; s < ns; ++s) D3H_HIP(hipStreamWaitEvent(streams[s - 1], events[0], 0));
  for (size_t ci = 0; ci < cs.size(); ++ci) {
    ClassPlan &c = cs[ci];
    hipStream_t st = (ci % ns == 0) ? main : streams[ci % ns - 1];
    const BatchMember *tab =
        reinterpret_cast<const BatchMember *>(static_cast<char *>(dev_tables) + c.table_off);
    const int *blk = reinterpret_cast<const int *>(static_cast<char *>(dev_tables) + c.block_off);
    const int maxl = c.launches.empty() ? 0 : c.launches[0];
    int active = (int)c.members.size();
    for (int k = 0; k < maxl; ++k) {
      while (active > 0 && c.launches[active - 1] <= k) --active;
      c.e->launch_batch(tab, blk, k, (unsigned)c.prefix_blocks[active], st);
      ++*launches;
    }
  }
  for (int s = 1; s < ns; ++s) {
    D3H_HIP(hipEventRecord(events[s], streams[s - 1]));
    D3H_HIP(hipStreamWaitEvent(main, events[s], 0));
  }
  D3H_HIP(hipGetLastError());
  return D3H_OK;
}

}  // namespace d3h

struct d3h_batch {
  d3h_context *ctx = nullptr;
  std::vector<d3h_problem> probs;
  std::vector<int> pitch;
  std::vector<size_t> off;    // byte offset of buffer 0 of each problem in the pool
  std::vector<size_t> bytes;  // bytes of one buffer (256-aligned)
  std::vector<int> cur;       // buffer holding each problem's current field
  void *pool = nullptr;
  void *staging = nullptr;
  size_t staging_bytes = 0;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  // cached plan + graph
  std::vector<long long> key;
  hipGraphExec_t gexec = nullptr;
  void *tables = nullptr;
  int plan_launches = 0;
  std::vector<int> cur_after;
};

namespace {

constexpr int kBatchStreams = 3;

void release_plan(d3h_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  if (b->tables) (void)hipFree(b->tables);
  b->tables = nullptr;
  b->key.clear();
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int batch_staging(d3h_batch *b, size_t bytes) {
  if (b->staging_bytes >= bytes) return D3H_OK;
  if (b->staging) (void)hipFree(b->staging);
  b->staging = nullptr;
  b->staging_bytes = 0;
  if (hipMalloc(&b->staging, bytes) != hipSuccess) return fail(D3H_ERR_NOMEM, "batch staging");
  b->staging_bytes = bytes;
  return D3H_OK;
}

}  // namespace

extern "C" {

int d3h_batch_create(d3h_context *ctx, const d3h_problem *problems, int count, d3h_batch **out) {
  if (!ctx || !out || (!problems && count > 0) || count < 0) return fail(D3H_ERR_INVALID, "bad batch arguments");
  *out = nullptr;
  for (int i = 0; i < count; ++i) D3H_TRY(check_problem(problems[i]));
  D3H_HIP(hipSetDevice(ctx->device));
  auto *b = new d3h_batch();
  b->ctx = ctx;
  b->probs.assign(problems, problems + count);
  size_t total = 0;
  for (const auto &p : b->probs) {
    const int pitch = pitch_elems(p.nx, p.dtype);
    const size_t bytes = ((size_t)pitch * p.ny * p.nz * dtype_bytes(p.dtype) + 255) / 256 * 256;
    b->pitch.push_back(pitch);
    b->bytes.push_back(bytes);
    b->off.push_back(total);
    total += 2 * bytes;
  }
  b->cur.assign(count, 0);
  if (total && hipMalloc(&b->pool, total) != hipSuccess) {
    delete b;
    return fail(D3H_ERR_NOMEM, "batch pool of %zu bytes", total);
  }
  if (total && hipMemset(b->pool, 0, total) != hipSuccess) {
    d3h_batch_destroy(b);
    return fail(D3H_ERR_HIP, "batch pool memset");
  }
  for (int s = 0; s < kBatchStreams; ++s) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      d3h_batch_destroy(b);
      return fail(D3H_ERR_HIP, "batch stream");
    }
    b->streams.push_back(st);
  }
  for (int s = 0; s <= kBatchStreams; ++s) {
    hipEvent_t ev;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      d3h_batch_destroy(b);
      return fail(D3H_ERR_HIP, "batch event");
    }
    b->events.push_back(ev);
  }
  *out = b;
  return D3H_OK;
}

int d3h_batch_destroy(d3h_batch *b) {
  if (!b) return D3H_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  release_plan(b);
  for (auto s : b->streams) (void)hipStreamDestroy(s);
  for (auto e : b->events) (void)hipEventDestroy(e);
  if (b->pool) (void)hipFree(b->pool);
  if (b->staging) (void)hipFree(b->staging);
  delete b;
  return D3H_OK;
}

int d3h_batch_upload(d3h_batch *b, int i, const void *host, int host_dtype) {
  if (!b || !host || i < 0 || i >= (int)b->probs.size() || !dtype_valid(host_dtype))
    return fail(D3H_ERR_INVALID, "bad upload arguments");
  const d3h_problem &p = b->probs[i];
  hipStream_t st = b->ctx->stream;
  D3H_HIP(hipSetDevice(b->ctx->device));
  char *dst = static_cast<char *>(b->pool) + b->off[i] + b->cur[i] * b->bytes[i];
  if (needs_staging(host_dtype, p.dtype, b->pitch[i], p.nx))
    D3H_TRY(batch_staging(b, (size_t)p.nx * p.ny * p.nz * dtype_bytes(host_dtype)));
  D3H_TRY(copy_in(host, host_dtype, dst, p.dtype, b->pitch[i], p.nx, p.ny, p.nz, b->staging, st));
  D3H_HIP(hipStreamSynchronize(st));
  return D3H_OK;
}

int d3h_batch_download(d3h_batch *b, int i, void *host, int host_dtype) {
  if (!b || !host || i < 0 || i >= (int)b->probs.size() || !dtype_valid(host_dtype))
    return fail(D3H_ERR_INVALID, "bad download arguments");
  const d3h_problem &p = b->probs[i];
  hipStream_t st = b->ctx->stream;
  D3H_HIP(hipSetDevice(b->ctx->device));
  const char *src = static_cast<const char *>(b->pool) + b->off[i] + b->cur[i] * b->bytes[i];
  if (needs_staging(host_dtype, p.dtype, b->pitch[i], p.nx))
    D3H_TRY(batch_staging(b, (size_t)p.nx * p.ny * p.nz * dtype_bytes(host_dtype)));
  D3H_TRY(copy_out(src, p.dtype, b->pitch[i], host, host_dtype, p.nx, p.ny, p.nz, b->staging, st));
  D3H_HIP(hipStreamSynchronize(st));
  return D3H_OK;
}

// Whole-batch transfers: every problem gets its own dense staging slot, all
// copies and conversions are queued on the stream, and the call synchronises
// once (pinned host buffers make the copies asynchronous DMA).
static int staging_slots(d3h_batch *b, int host_dtype, std::vector<size_t> *offs) {
  size_t total = 0;
  offs->clear();
  for (const auto &p : b->probs) {
    offs->push_back(total);
    total += ((size_t)p.nx * p.ny * p.nz * dtype_bytes(host_dtype) + 255) / 256 * 256;
  }
  return batch_staging(b, std::max<size_t>(total, 256));
}

int d3h_batch_upload_all(d3h_batch *b, const void *const *hosts, int host_dtype) {
  if (!b || !hosts || !dtype_valid(host_dtype)) return fail(D3H_ERR_INVALID, "bad upload arguments");
  for (size_t i = 0; i < b->probs.size(); ++i)
    if (!hosts[i]) return fail(D3H_ERR_INVALID, "host buffer %zu is NULL", i);
  D3H_HIP(hipSetDevice(b->ctx->device));
  std::vector<size_t> offs;
  D3H_TRY(staging_slots(b, host_dtype, &offs));
  hipStream_t st = b->ctx->stream;
  for (size_t i = 0; i < b->probs.size(); ++i) {
    const d3h_problem &p = b->probs[i];
    char *dst = static_cast<char *>(b->pool) + b->off[i] + b->cur[i] * b->bytes[i];
    D3H_TRY(copy_in(hosts[i], host_dtype, dst, p.dtype, b->pitch[i], p.nx, p.ny, p.nz,
                    static_cast<char *>(b->staging) + offs[i], st));
  }
  D3H_HIP(hipStreamSynchronize(st));
  return D3H_OK;
}

int d3h_batch_download_all(d3h_batch *b, void *const *hosts, int host_dtype) {
  if (!b || !hosts || !dtype_valid(host_dtype)) return fail(D3H_ERR_INVALID, "bad download arguments");
  for (size_t i = 0; i < b->probs.size(); ++i)
    if (!hosts[i]) return fail(D3H_ERR_INVALID, "host buffer %zu is NULL", i);
  D3H_HIP(hipSetDevice(b->ctx->device));
  std::vector<size_t> offs;
  D3H_TRY(staging_slots(b, host_dtype, &offs));
  hipStream_t st = b->ctx->stream;
  for (size_t i = 0; i < b->probs.size(); ++i) {
    const d3h_problem &p = b->probs[i];
    const char *src = static_cast<const char *>(b->pool) + b->off[i] + b->cur[i] * b->bytes[i];
    D3H_TRY(copy_out(src, p.dtype, b->pitch[i], hosts[i], host_dtype, p.nx, p.ny, p.nz,
                     static_cast<char *>(b->staging) + offs[i], st));
  }
  D3H_HIP(hipStreamSynchronize(st));
  return D3H_OK;
}

// Batch autotuner.  The batch's problems are grouped by size class (dtype and
// ceil log2 of each dimension).  Class by class, every candidate (depth 1, 2
// or 4 with each batched tile of that depth, then z-chunk counts 1, 2, 4 and
// 8 for the best of them) is timed as a sweep of kTuneSteps steps over the
// WHOLE batch, the other classes keeping their current choice, because
// classes share the device in a real run; the fastest stays.  The choices go
// to the context's tune table as batch entries (depth 0 and the chosen depth),
// which depth-0 batch runs and servers then use.
static constexpr int kTuneSteps = 8;  // a multiple of every batched depth
static constexpr int kTuneReps = 5;

static int time_batch_sweep(d3h_batch *b, const std::vector<PlanMember> &ms, int fma, void **tables,
                            size_t *cap, std::vector<char> *host, double *best_ms) {
  d3h_context *ctx = b->ctx;
  SweepPlan plan;
  D3H_TRY(plan_sweep(ctx, ms, 0, fma, &plan));
  if (plan.table_bytes > *cap) {
    if (*tables) (void)hipFree(*tables);
    *tables = nullptr;
    *cap = 0;
    if (hipMalloc(tables, plan.table_bytes) != hipSuccess) return fail(D3H_ERR_NOMEM, "tune tables");
    *cap = plan.table_bytes;
  }
  host->resize(plan.table_bytes);
  plan_fill(plan, host->data());
  if (plan.table_bytes)
    D3H_HIP(hipMemcpy(*tables, host->data(), plan.table_bytes, hipMemcpyHostToDevice));
  int launches = 0;
  *best_ms = 1e300;
  for (int r = 0; r <= kTuneReps; ++r) {  // r == 0 warms up
    D3H_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    D3H_TRY(plan_enqueue(plan.A, *tables, ctx->stream, b->streams, b->events, &launches));
    D3H_TRY(plan_enqueue(plan.B, *tables, ctx->stream, b->streams, b->events, &launches));
    D3H_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    D3H_HIP(hipEventSynchronize(ctx->ev1));
    float t = 0;
    D3H_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (r > 0) *best_ms = std::min(*best_ms, (double)t);
  }
  return D3H_OK;
}

int d3h_batch_autotune(d3h_batch *b, int arith, int *classes, double *ms_before, double *ms_after) {
  if (!b) return fail(D3H_ERR_INVALID, "batch is NULL");
  if (arith != D3H_ARITH_EXACT && arith != D3H_ARITH_FMA) return fail(D3H_ERR_INVALID, "arith %d", arith);
  const int fma = arith == D3H_ARITH_FMA;
  d3h_context *ctx = b->ctx;
  D3H_HIP(hipSetDevice(ctx->device));
  release_plan(b);  // the cached sequence may change with the choices
  std::vector<PlanMember> ms;
  std::map<TuneKey, std::vector<int>> groups;  // depth-0 batch key -> members
  for (size_t i = 0; i < b->probs.size(); ++i) {
    const d3h_problem &p = b->probs[i];
    PlanMember m{};
    m.nx = p.nx; m.ny = p.ny; m.nz = p.nz; m.dtype = p.dtype; m.pitch = b->pitch[i];
    char *base = static_cast<char *>(b->pool) + b->off[i];
    m.buf[0] = base;
    m.buf[1] = base + b->bytes[i];
    m.parity0 = b->cur[i];
    m.steps = kTuneSteps;
    for (int k = 0; k < 7; ++k) m.coef[k] = p.coef[k];
    m.sr = p.rounding == D3H_ROUND_STOCHASTIC;
    m.seed = p.seed;
    ms.push_back(m);
    groups[tune_key(p.dtype, 0, fma, p.nx, p.ny, p.nz, 1)].push_back((int)i);
  }
  void *tables = nullptr;
  size_t cap = 0;
  std::vector<char> host;
  auto set_entry = [&](TuneKey k, const KernelEntry &e, int zc, double t) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tuned[k] = TuneVal{e.vec, e.G, e.RPW, e.NW, e.pf, zc, t, e.T, 0};
  };
  auto with_depth = [](TuneKey k, int T) {
    k.T = T;
    return k;
  };
  int rc = D3H_OK;
  double t0 = 0, cur = 0;
  int tuned = 0;
  rc = time_batch_sweep(b, ms, fma, &tables, &cap, &host, &t0);
  cur = t0;
  for (auto it = groups.begin(); rc == D3H_OK && it != groups.end(); ++it) {
    const TuneKey &k0 = it->first;
    const PlanMember &rep = ms[it->second[0]];
    // the class's entries before this search, restored if it fails
    std::map<TuneKey, TuneVal> saved;
    std::vector<TuneKey> keys = {k0, with_depth(k0, 1), with_depth(k0, 2), with_depth(k0, 4)};
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      for (const auto &k : keys) {
        auto f = ctx->tuned.find(k);
        if (f != ctx->tuned.end()) saved[k] = f->second;
      }
    }
    auto restore = [&] {
      std::lock_guard<std::mutex> lk(ctx->mu);
      for (const auto &k : keys) {
        auto f = saved.find(k);
        if (f != saved.end()) ctx->tuned[k] = f->second;
        else ctx->tuned.erase(k);
      }
    };
    std::map<int, std::pair<const KernelEntry *, double>> best_of;  // depth -> fastest tile
    const KernelEntry *best = nullptr;
    int best_zc = 0;
    double best_t = 1e300;
    // candidates: depth and tile (the depth-0 entry selects the depth)
    for (int T : {1, 2, 4}) {
      for (const auto &e : registry()) {
        if (!e.launch_batch || e.dtype != k0.dtype || e.T != T || e.fma != fma || occupancy_of(e) <= 0) continue;
        set_entry(k0, e, 0, 0.0);
        set_entry(with_depth(k0, T), e, 0, 0.0);
        double t = 0;
        rc = time_batch_sweep(b, ms, fma, &tables, &cap, &host, &t);
        if (rc != D3H_OK) break;
        auto bo = best_of.find(T);
        if (bo == best_of.end() || t < bo->second.second) best_of[T] = {&e, t};
        if (t < best_t) {
          best_t = t;
          best = &e;
        }
      }
      if (rc != D3H_OK) break;
    }
    if (rc != D3H_OK || !best) {  // an error, or no batched kernel for this dtype
      restore();
      if (rc != D3H_OK) break;
      continue;
    }
    // every tried depth keeps its fastest tile (explicit-depth batch runs)
    for (const auto &kv : best_of)
      set_entry(with_depth(k0, kv.first), *kv.second.first, 0, kv.second.second / kTuneSteps);
    // then the z-chunk count of the best
    set_entry(k0, *best, 0, 0.0);
    for (int zc : {1, 2, 4, 8}) {
      if (zc > rep.nz) continue;
      set_entry(k0, *best, zc, 0.0);
      set_entry(with_depth(k0, best->T), *best, zc, 0.0);
      double t = 0;
      rc = time_batch_sweep(b, ms, fma, &tables, &cap, &host, &t);
      if (rc != D3H_OK) break;
      if (t < best_t) {
        best_t = t;
        best_zc = zc;
      }
    }
    if (rc != D3H_OK) {
      restore();
      break;
    }
    set_entry(k0, *best, best_zc, best_t / kTuneSteps);
    set_entry(with_depth(k0, best->T), *best, best_zc, best_t / kTuneSteps);
    cur = best_t;
    ++tuned;
  }
  if (tables) (void)hipFree(tables);
  if (rc != D3H_OK) return rc;
  if (classes) *classes = tuned;
  if (ms_before) *ms_before = t0;
  if (ms_after) *ms_after = cur;
  return D3H_OK;
}

int d3h_batch_run(d3h_batch *b, int arith, int temporal_depth, int use_graph, d3h_timing *timing) {
  const double t0 = now_ms();
  if (!b) return fail(D3H_ERR_INVALID, "batch is NULL");
  if (arith != D3H_ARITH_EXACT && arith != D3H_ARITH_FMA) return fail(D3H_ERR_INVALID, "arith %d", arith);
  const int T = temporal_depth;
  if (T != 0 && T != 1 && T != 2 && T != 4 && T != 8) return fail(D3H_ERR_INVALID, "depth %d", T);
  const int fma = arith == D3H_ARITH_FMA;
  d3h_context *ctx = b->ctx;
  D3H_HIP(hipSetDevice(ctx->device));
  std::vector<long long> key = {T, fma};
  for (int c : b->cur) key.push_back(c);
  int launches = 0;
  if (!use_graph || key != b->key) {
    release_plan(b);
    std::vector<PlanMember> ms;
    for (size_t i = 0; i < b->probs.size(); ++i) {
      const d3h_problem &p = b->probs[i];
      PlanMember m{};
      m.nx = p.nx; m.ny = p.ny; m.nz = p.nz; m.dtype = p.dtype; m.pitch = b->pitch[i];
      char *base = static_cast<char *>(b->pool) + b->off[i];
      m.buf[0] = base;
      m.buf[1] = base + b->bytes[i];
      m.parity0 = b->cur[i];
      m.steps = p.steps;
      for (int k = 0; k < 7; ++k) m.coef[k] = p.coef[k];
      m.sr = p.rounding == D3H_ROUND_STOCHASTIC;
      m.seed = p.seed;
      m.step_base = 0;  // step numbers start at 0 in every run
      ms.push_back(m);
    }
    SweepPlan plan;
    D3H_TRY(plan_sweep(ctx, ms, T, fma, &plan));
    if (plan.table_bytes) {
      if (hipMalloc(&b->tables, plan.table_bytes) != hipSuccess) return fail(D3H_ERR_NOMEM, "batch tables");
      std::vector<char> host(plan.table_bytes);
      plan_fill(plan, host.data());
      D3H_HIP(hipMemcpy(b->tables, host.data(), plan.table_bytes, hipMemcpyHostToDevice));
    }
    b->cur_after = plan.parity_after;
    if (use_graph) {
      hipGraph_t graph = nullptr;
      D3H_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int rc = plan_enqueue(plan.A, b->tables, ctx->stream, b->streams, b->events, &launches);
      if (rc == D3H_OK) rc = plan_enqueue(plan.B, b->tables, ctx->stream, b->streams, b->events, &launches);
      hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
      if (rc != D3H_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (ce != hipSuccess) return fail(D3H_ERR_HIP, "batch capture: %s", hipGetErrorString(ce));
      hipError_t ie = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) return fail(D3H_ERR_HIP, "batch instantiate: %s", hipGetErrorString(ie));
      b->key = key;
      b->plan_launches = launches;
    } else {
      D3H_HIP(hipEventRecord(ctx->ev0, ctx->stream));
      D3H_TRY(plan_enqueue(plan.A, b->tables, ctx->stream, b->streams, b->events, &launches));
      D3H_TRY(plan_enqueue(plan.B, b->tables, ctx->stream, b->streams, b->events, &launches));
      D3H_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    }
  }
  if (use_graph) {
    launches = b->plan_launches;
    D3H_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    D3H_HIP(hipGraphLaunch(b->gexec, ctx->stream));
    D3H_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  }
  D3H_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0;
  D3H_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  b->cur = b->cur_after;
  if (!use_graph) release_plan(b);
  if (timing) {
    memset(timing, 0, sizeof *timing);
    timing->kernel_ms = ms;
    timing->launches = launches;
    timing->avg_launch_ms = launches ? ms / launches : 0;
    timing->temporal_depth = T;
    timing->tile = -1;
    timing->wall_ms = now_ms() - t0;
  }
  return D3H_OK;
}

}  // extern "C"
