// FP32 2D kernels for nz == 1 fields (the diffusion/ family's 5-point
// operator).  One wave per workgroup streams a strip of 256 * G columns down a
// chunk of the field's rows, RPW rows per streamed "plane": a row's y
// neighbours are the adjacent rows of its block, or across a block boundary
// the adjacent block's edge row, already in registers.  T fused steps cost one
// read and one write of the strip plus T halo blocks per chunk and the
// 8-cell-aligned x halos.  The first entry of each depth is the default (the
// fastest at 16384^2).
#include "kernels.hpp"

namespace d3h {

#define F32_2D(A)                                                                 \
  D3H_ENTRY_2D(kF32, 4, 1, 2, 4, A, 2), D3H_ENTRY_2D(kF32, 4, 1, 2, 4, A, 1),       \
  D3H_ENTRY_2D(kF32, 4, 1, 2, 2, A, 2), D3H_ENTRY_2D(kF32, 4, 1, 2, 2, A, 1),       \
  D3H_ENTRY_2D(kF32, 4, 1, 2, 1, A, 2), D3H_ENTRY_2D(kF32, 4, 1, 1, 8, A, 1),       \
  D3H_ENTRY_2D(kF32, 4, 2, 2, 2, A, 1), D3H_ENTRY_2D(kF32, 4, 2, 1, 4, A, 1),       \
  D3H_ENTRY_2D(kF32, 4, 2, 2, 1, A, 1),                                            \
  D3H_ENTRY_2D(kF32, 4, 4, 1, 2, A, 1), D3H_ENTRY_2D(kF32, 4, 4, 2, 1, A, 1),       \
  D3H_ENTRY_2D(kF32, 4, 4, 1, 1, A, 1),                                            \
  D3H_ENTRY_2D(kF32, 4, 8, 1, 1, A, 1)

static const KernelEntry kTable[] = {F32_2D(false), F32_2D(true)};

const KernelEntry *kernels_f32_rows(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace d3h
