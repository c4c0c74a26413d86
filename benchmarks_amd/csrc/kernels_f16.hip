// FP16-storage kernels (FP32 arithmetic, round-to-nearest-even on store).
#include "kernels.hpp"
#include "kernels_lp.inc"

namespace d3h {

static const KernelEntry kTable[] = {LP_TILES(kF16, false), LP_TILES(kF16, true)};

const KernelEntry *kernels_f16(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace d3h
