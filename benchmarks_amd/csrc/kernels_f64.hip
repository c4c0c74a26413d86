// FP64 kernels.  16-byte lanes hold 2 cells; G = 4 gives 512-wide rows.
#include "kernels.hpp"

namespace d3h {

#define F64_TILES(A)                                                              \
  D3H_ENTRY(kF64, 2, 1, 4, 2, 4, A), D3H_ENTRY(kF64, 2, 1, 4, 1, 4, A),           \
  D3H_ENTRY(kF64, 2, 1, 4, 2, 2, A), D3H_ENTRY(kF64, 2, 1, 4, 1, 8, A),           \
  D3H_ENTRY_B(kF64, 2, 1, 2, 2, 4, A), D3H_ENTRY(kF64, 2, 1, 2, 4, 2, A),           \
  D3H_ENTRY(kF64, 2, 1, 1, 4, 2, A), D3H_ENTRY_B(kF64, 2, 1, 1, 2, 4, A),           \
  D3H_ENTRY(kF64, 2, 2, 1, 4, 8, A), D3H_ENTRY_B(kF64, 2, 2, 1, 2, 16, A),          \
  D3H_ENTRY(kF64, 1, 4, 1, 2, 16, A), D3H_ENTRY(kF64, 2, 4, 1, 2, 8, A),          \
  D3H_ENTRY(kF64, 1, 8, 1, 3, 8, A)

// 2D kernels (nz == 1 fields, 5-point update): 512-wide strips, 2 or 4 rows
// per streamed block.
#define F64_2D(A)                                                                 \
  D3H_ENTRY_2D(kF64, 2, 1, 4, 2, A, 1), D3H_ENTRY_2D(kF64, 2, 1, 4, 1, A, 2),       \
  D3H_ENTRY_2D(kF64, 2, 2, 2, 2, A, 1), D3H_ENTRY_2D(kF64, 2, 4, 1, 2, A, 1)

static const KernelEntry kTable[] = {F64_TILES(false), F64_TILES(true), F64_2D(false), F64_2D(true)};

const KernelEntry *kernels_f64(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace d3h
