// Registry of compiled stencil kernels (one entry per dtype / lane vector /
// temporal depth / tile shape / arithmetic mode).
#pragma once

#include <hip/hip_runtime.h>

#include "stencil.hpp"

namespace d3h {

typedef void (*LaunchFn)(const void *src, void *dst, const Geometry &g, const double *coef,
                         unsigned blocks, hipStream_t stream);
typedef void (*LaunchBatchFn)(const BatchMember *members, const int *block_member, int launch,
                              unsigned blocks, hipStream_t stream);

struct KernelEntry {
  int dtype;  // StoreKind
  int vec;    // cells per lane vector
  int T;      // fused steps per launch
  int G;      // lane vectors per row (groups)
  int RPW;    // rows per wave
  int NW;     // waves per workgroup
  int fma;    // 0 = exact, 1 = fma chain
  int pf;     // prefetch distance in planes
  const void *func;
  LaunchFn launch;
  const void *batch_func;      // nullptr when no batched instantiation exists
  LaunchBatchFn launch_batch;
  // 1: 2D entry (stencil2d_kernel: one wave, RPW rows per streamed "plane",
  // the 5-point update); it runs nz == 1 fields with ct == cb == 0 as a
  // stream over their rows and nothing else.
  int rows = 0;
  int tile_x() const { return 64 * G * vec; }
  int tile_y() const { return NW * RPW; }
};

template <int S, int VEC, int T, int G, int RPW, int NW, bool FMA, int PF>
void launch_stencil(const void *src, void *dst, const Geometry &g, const double *c, unsigned blocks,
                    hipStream_t stream) {
  using C = typename Traits<S>::C;
  using ST = typename Traits<S>::T;
  Coef<C> k{(C)c[0], (C)c[1], (C)c[2], (C)c[3], (C)c[4], (C)c[5], (C)c[6]};
  hipLaunchKernelGGL((stencil_kernel<S, VEC, T, G, RPW, NW, FMA, PF>), dim3(blocks), dim3(64 * NW), 0,
                     stream, (const ST *)src, (ST *)dst, g, k);
}

template <int S, int VEC, int T, int G, int RPW, int NW, bool FMA, int PF>
void launch_stencil_batch(const BatchMember *members, const int *block_member, int launch,
                          unsigned blocks, hipStream_t stream) {
  hipLaunchKernelGGL((stencil_batch_kernel<S, VEC, T, G, RPW, NW, FMA, PF>), dim3(blocks), dim3(64 * NW),
                     0, stream, members, block_member, launch);
}

template <int S, int VEC, int T, int G, int RPW, bool FMA, int PF>
void launch_stencil2d(const void *src, void *dst, const Geometry &g, const double *c, unsigned blocks,
                      hipStream_t stream) {
  using C = typename Traits<S>::C;
  using ST = typename Traits<S>::T;
  Coef<C> k{(C)c[0], (C)c[1], (C)c[2], (C)c[3], (C)c[4], (C)c[5], (C)c[6]};
  hipLaunchKernelGGL((stencil2d_kernel<S, VEC, T, G, RPW, FMA, PF>), dim3(blocks), dim3(64), 0, stream,
                     (const ST *)src, (ST *)dst, g, k);
}

// Single-grid kernel only (prefetch distance PF).
#define D3H_ENTRY_PF(S, VEC, T, G, RPW, NW, FMA, PF)                                          \
  KernelEntry {                                                                               \
    S, VEC, T, G, RPW, NW, FMA, PF, (const void *)&stencil_kernel<S, VEC, T, G, RPW, NW, FMA, PF>, \
        &launch_stencil<S, VEC, T, G, RPW, NW, FMA, PF>, nullptr, nullptr                     \
  }
#define D3H_ENTRY(S, VEC, T, G, RPW, NW, FMA) D3H_ENTRY_PF(S, VEC, T, G, RPW, NW, FMA, 1)
// Single-grid and batched kernels.
#define D3H_ENTRY_B(S, VEC, T, G, RPW, NW, FMA)                                               \
  KernelEntry {                                                                               \
    S, VEC, T, G, RPW, NW, FMA, 1, (const void *)&stencil_kernel<S, VEC, T, G, RPW, NW, FMA, 1>, \
        &launch_stencil<S, VEC, T, G, RPW, NW, FMA, 1>,                                       \
        (const void *)&stencil_batch_kernel<S, VEC, T, G, RPW, NW, FMA, 1>,                   \
        &launch_stencil_batch<S, VEC, T, G, RPW, NW, FMA, 1>                                  \
  }

// 2D kernel (one wave, RPW rows per streamed plane, 5-point update).
#define D3H_ENTRY_2D(S, VEC, T, G, RPW, FMA, PF)                                              \
  KernelEntry {                                                                               \
    S, VEC, T, G, RPW, 1, FMA, PF, (const void *)&stencil2d_kernel<S, VEC, T, G, RPW, FMA, PF>, \
        &launch_stencil2d<S, VEC, T, G, RPW, FMA, PF>, nullptr, nullptr, 1                    \
  }

// Each kernels_*.hip file defines one table.
const KernelEntry *kernels_f32_plain(int *n);
const KernelEntry *kernels_f32_tb(int *n);
const KernelEntry *kernels_f64(int *n);
const KernelEntry *kernels_bf16(int *n);
const KernelEntry *kernels_f16(int *n);
const KernelEntry *kernels_f32_rows(int *n);

}  // namespace d3h
