// Instantiations of the stage kernel with the dust fluid
// (components/dust_fluid/dust.cc): FP64 and FP32-split storage, orders 4 and 6.
#include "bssn_stage_kernel.h"

namespace cg {
#define CG_DUST(R, TI, TJ, TK, WPE)                                   \
  CG_REGISTER_DUST(double, double, 0, R, R, TI, TJ, TK, WPE, false)  \
  CG_REGISTER_DUST(double, double, 0, R, R, TI, TJ, TK, WPE, true)   \
  CG_REGISTER_DUST(float, float, 1, R, R, TI, TJ, TK, WPE, false)    \
  CG_REGISTER_DUST(float, float, 1, R, R, TI, TJ, TK, WPE, true)
// with the LDS-DMA twin (stages without metric normalisation, nz % 4 == 0)
#define CG_DUST32(R, TI, TJ, TK, WPE)                                  \
  CG_REGISTER_DUST_D(float, float, 1, R, R, TI, TJ, TK, WPE, false)    \
  CG_REGISTER_DUST_D(float, float, 1, R, R, TI, TJ, TK, WPE, true)
#define CG_DUST64(R, TI, TJ, TK, WPE)                                   \
  CG_REGISTER_DUST_D(double, double, 0, R, R, TI, TJ, TK, WPE, false)  \
  CG_REGISTER_DUST_D(double, double, 0, R, R, TI, TJ, TK, WPE, true)
CG_DUST(2, 4, 8, 8, 1)
CG_DUST(3, 4, 8, 8, 1)
// register budgets for 3 (FP32) and 2 (FP64) waves per SIMD, and a tile
// long along k; the autotuner keeps whichever is faster
CG_DUST32(3, 4, 8, 8, 3)
CG_DUST32(3, 4, 4, 16, 3)
CG_DUST64(3, 4, 8, 8, 2)
CG_DUST32(2, 4, 8, 8, 3)
}  // namespace cg
