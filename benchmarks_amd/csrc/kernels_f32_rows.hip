// FP32 row-streaming kernels for nz == 1 fields (the 2D 5-point operator).
// One wave per workgroup streams a strip of 256 * G columns down a chunk of
// rows; the tile runner's plane pipeline holds rows instead of planes, so T
// fused steps cost one read and one write of the strip plus T halo rows per
// chunk and 8-cell-aligned x halos.  The first entry of each depth is the
// default (best at 8192^2 and 16384^2 with the default chunking).
#include "kernels.hpp"

namespace d3h {

#define F32_ROWS(A)                                                               \
  D3H_ENTRY_ROW(kF32, 4, 1, 2, A, 2), D3H_ENTRY_ROW(kF32, 4, 1, 1, A, 1),           \
  D3H_ENTRY_ROW(kF32, 4, 1, 4, A, 1), D3H_ENTRY_ROW(kF32, 4, 1, 2, A, 1),           \
  D3H_ENTRY_ROW(kF32, 4, 1, 2, A, 3), D3H_ENTRY_ROW(kF32, 4, 1, 4, A, 2),           \
  D3H_ENTRY_ROW(kF32, 4, 1, 1, A, 3),                                              \
  D3H_ENTRY_ROW(kF32, 4, 2, 1, A, 1), D3H_ENTRY_ROW(kF32, 4, 2, 2, A, 1),           \
  D3H_ENTRY_ROW(kF32, 4, 2, 4, A, 1), D3H_ENTRY_ROW(kF32, 4, 2, 2, A, 2),           \
  D3H_ENTRY_ROW(kF32, 4, 4, 1, A, 1), D3H_ENTRY_ROW(kF32, 4, 4, 2, A, 1),           \
  D3H_ENTRY_ROW(kF32, 4, 4, 2, A, 2),                                              \
  D3H_ENTRY_ROW(kF32, 4, 8, 1, A, 1)

static const KernelEntry kTable[] = {F32_ROWS(false), F32_ROWS(true)};

const KernelEntry *kernels_f32_rows(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace d3h
