// FP32 plain-step kernels (one step per launch).  The first entry is the
// default when no tuned choice exists: 512-wide rows (G = 2), so a 512^3 grid
// needs no x halo at all.
#include "kernels.hpp"

namespace d3h {

#define F32_PLAIN(A)                                                              \
  D3H_ENTRY_B(kF32, 4, 1, 2, 2, 4, A), D3H_ENTRY(kF32, 4, 1, 2, 4, 2, A),           \
  D3H_ENTRY(kF32, 4, 1, 2, 2, 2, A), D3H_ENTRY(kF32, 4, 1, 2, 4, 1, A),           \
  D3H_ENTRY(kF32, 4, 1, 2, 1, 8, A),           \
  D3H_ENTRY(kF32, 4, 1, 2, 1, 4, A), D3H_ENTRY(kF32, 4, 1, 1, 4, 4, A),           \
  D3H_ENTRY_B(kF32, 4, 1, 1, 2, 4, A), D3H_ENTRY(kF32, 4, 1, 1, 4, 2, A),           \
  D3H_ENTRY(kF32, 4, 1, 1, 2, 8, A), D3H_ENTRY(kF32, 4, 1, 1, 8, 1, A),           \
  D3H_ENTRY(kF32, 4, 1, 1, 4, 1, A), D3H_ENTRY_B(kF32, 2, 1, 1, 4, 2, A),           \
  D3H_ENTRY_B(kF32, 2, 1, 1, 2, 4, A), D3H_ENTRY(kF32, 4, 1, 3, 2, 4, A),          \
  D3H_ENTRY(kF32, 4, 1, 3, 1, 8, A), D3H_ENTRY(kF32, 4, 1, 3, 2, 2, A),            \
  D3H_ENTRY(kF32, 4, 1, 3, 1, 4, A),                                               \
  /* deeper prefetch (2-3 planes in flight per wave) */                          \
  D3H_ENTRY_PF(kF32, 4, 1, 2, 1, 4, A, 2), D3H_ENTRY_PF(kF32, 4, 1, 2, 1, 4, A, 3), \
  D3H_ENTRY_PF(kF32, 4, 1, 2, 2, 4, A, 2), D3H_ENTRY_PF(kF32, 4, 1, 2, 1, 8, A, 2), \
  D3H_ENTRY_PF(kF32, 4, 1, 2, 2, 2, A, 2), D3H_ENTRY_PF(kF32, 4, 1, 3, 1, 4, A, 2), \
  D3H_ENTRY_PF(kF32, 4, 1, 2, 4, 1, A, 2),                                         \
  /* one cell per lane: 64-wide tiles for narrow (batched) grids */              \
  D3H_ENTRY_B(kF32, 1, 1, 1, 4, 4, A), D3H_ENTRY_B(kF32, 1, 1, 1, 8, 2, A)

static const KernelEntry kTable[] = {F32_PLAIN(false), F32_PLAIN(true)};

const KernelEntry *kernels_f32_plain(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace d3h
