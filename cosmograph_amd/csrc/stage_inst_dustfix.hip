// Instantiations of the stage kernel with the dust fluid and the
// fixed-point metric (every other field FP32), orders 4 and 6: tiles with
// the register budget of 3 waves per SIMD and, at order 6, tiles without a
// budget (the frozen-lapse kernels fit 3 waves in 166 VGPRs without
// spilling, where the budgeted ones spill 24-36 B), each with its LDS-DMA
// twin; the autotuner keeps the fastest.
#include "bssn_stage_kernel.h"

namespace cg {
constexpr int MODE_DUST_MQ = MODE_DUST | MODE_MQ_STAGE;
#define CG_REGISTER_DMQ(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA) \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_DUST_MQ, nullptr, nullptr, nullptr, nullptr)
#define CG_REGISTER_DMQ_D(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA)                                     \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_DUST_MQ, nullptr,                 \
                      CG_DMA_L(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_DUST_MQ),                      \
                      CG_DMA_I(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_DUST_MQ), nullptr)
#define CG_DQ(REG, R, TI, TJ, TK, WPE)               \
  REG(float, float, 3, R, R, TI, TJ, TK, WPE, false) \
  REG(float, float, 3, R, R, TI, TJ, TK, WPE, true)
CG_DQ(CG_REGISTER_DMQ, 2, 4, 8, 8, 1)
CG_DQ(CG_REGISTER_DMQ_D, 3, 4, 8, 8, 1)
CG_DQ(CG_REGISTER_DMQ_D, 2, 4, 8, 8, 3)
CG_DQ(CG_REGISTER_DMQ_D, 3, 4, 8, 8, 3)
CG_DQ(CG_REGISTER_DMQ_D, 3, 4, 4, 16, 3)
CG_DQ(CG_REGISTER_DMQ_D, 3, 4, 4, 16, 1)
}  // namespace cg
