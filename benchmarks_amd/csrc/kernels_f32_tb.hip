// FP32 temporally blocked kernels (T = 2, 4, 8 fused steps per launch).  Tile
// shapes are limited by the register file: every level keeps three planes of
// the lane's cells live (plus the planes in flight), so deeper T uses fewer
// cells per lane (8-byte lane vectors at T >= 4) and taller workgroups to keep
// the redundant halo rows a small fraction of the tile.  The three-slot ring
// (prefetch 0) trades some load-latency hiding for registers: at 768^3 it
// lifts depth 8 from 532 to 809 Gcell/s effective and runs depth 4 with 32-row
// tiles as fast as the 24-row prefetch-1 tile.
#include "kernels.hpp"

namespace d3h {

#define F32_TB(A)                                                                 \
  D3H_ENTRY(kF32, 4, 2, 1, 4, 8, A), D3H_ENTRY_B(kF32, 4, 2, 1, 2, 8, A),           \
  D3H_ENTRY(kF32, 4, 2, 1, 2, 16, A), D3H_ENTRY(kF32, 4, 2, 1, 4, 4, A),          \
  D3H_ENTRY_B(kF32, 2, 2, 1, 4, 8, A), D3H_ENTRY(kF32, 4, 2, 2, 2, 4, A),           \
  D3H_ENTRY(kF32, 2, 4, 1, 3, 8, A), D3H_ENTRY_B(kF32, 2, 4, 1, 2, 16, A), D3H_ENTRY(kF32, 4, 4, 1, 2, 8, A),          \
  D3H_ENTRY(kF32, 2, 4, 1, 4, 8, A), D3H_ENTRY(kF32, 4, 2, 3, 2, 8, A),            \
  D3H_ENTRY(kF32, 2, 8, 1, 3, 8, A),                                               \
  /* one cell per lane: 64-wide tiles for narrow (batched) grids */              \
  D3H_ENTRY_B(kF32, 1, 2, 1, 4, 8, A), D3H_ENTRY_B(kF32, 1, 4, 1, 4, 8, A),       \
  /* three-slot register ring (prefetch 0): taller tiles at the same occupancy */ \
  D3H_ENTRY_PF(kF32, 2, 4, 1, 4, 8, A, 0), D3H_ENTRY_PF(kF32, 2, 8, 1, 4, 8, A, 0)

static const KernelEntry kTable[] = {F32_TB(false), F32_TB(true)};

const KernelEntry *kernels_f32_tb(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace d3h
