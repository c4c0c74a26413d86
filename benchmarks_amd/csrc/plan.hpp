// Sweep planner shared by batched execution and the request server.
//
// A sweep advances a set of members (grids with their own buffers, dims,
// dtype, coefficients and step count) by their counts.  Members sharing a
// kernel form a class; one launch of a class covers every tile of every
// still-active member (members sorted by launch count, so the active set of
// launch k is a prefix and the block table of launch k is a prefix of the
// class table).  Phase A runs the T-fused launches, phase B the count % T
// plain tail launches.
#pragma once

#include <vector>

#include "runtime.hpp"

namespace d3h {

struct PlanMember {
  int nx, ny, nz, dtype, pitch;
  void *buf[2];
  int parity0;   // buffer holding the input
  int steps;     // steps to advance in this sweep
  double coef[7];
  int sr;        // stochastic rounding of BF16 storage
  unsigned seed;
  int step_base; // steps the member has taken before this sweep
};

struct ClassPlan {
  const KernelEntry *e = nullptr;
  std::vector<int> members;        // member indices, sorted by launches desc
  std::vector<int> launches;       // per sorted member
  std::vector<BatchMember> table;
  std::vector<int> block_member;
  std::vector<int> prefix_blocks;  // blocks of the first n members
  size_t table_off = 0, block_off = 0;
};

struct SweepPlan {
  std::vector<ClassPlan> A, B;
  std::vector<int> parity_after;   // per member
  size_t table_bytes = 0;
  int launches = 0;
};

// Kernel for one member at depth T: the batch autotuner's entry for its size
// class, else the single-grid autotuner's entry if it has a batched kernel,
// else the narrowest batched tile covering nx.  *zchunks (if given) gets the
// batch autotuner's z-chunk count, or 0 for the default.  nullptr if none.
const KernelEntry *pick_batch_entry(d3h_context *ctx, int nx, int ny, int nz, int dtype, int T, int fma,
                                    int *zchunks = nullptr);
// Depth for a member run at depth 0: the batch autotuner's choice for its size
// class, else the single-grid autotuner's, else `dflt`.
int batch_depth(d3h_context *ctx, int dtype, int fma, int nx, int ny, int nz, int dflt);

// Build the class plans (host side) and the table layout.
int plan_sweep(d3h_context *ctx, const std::vector<PlanMember> &members, int T, int fma,
               SweepPlan *plan);
// Write the tables of `plan` into a host buffer of plan->table_bytes.
void plan_fill(const SweepPlan &plan, char *host);
// Enqueue the launches of a phase (class c on streams[c % n], forked from and
// joined back to `main` with the given events; events.size() >= streams.size()+1).
int plan_enqueue(std::vector<ClassPlan> &cs, void *dev_tables, hipStream_t main,
                 const std::vector<hipStream_t> &streams, const std::vector<hipEvent_t> &events,
                 int *launches);
// Validate a problem: dims, dtype, steps >= 0, coefficients >= 0 summing to 1,
// rounding (stochastic only with BF16 storage).
int check_problem(const d3h_problem &p);

}  // namespace d3h
