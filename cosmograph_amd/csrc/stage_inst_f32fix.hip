// Instantiations of the stage kernel with the fixed-point metric (the
// conformal metric in 32-bit fixed point, every other field in FP32),
// orders 4 and 6: the default tile at one wave per SIMD and the
// 3-waves-per-SIMD tiles with their LDS-DMA twins.
#include "bssn_stage_kernel.h"

namespace cg {
#define CG_REGISTER_MQ_D(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA)                                      \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_MQ_STAGE, nullptr,                \
                      CG_DMA_L(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_MQ_STAGE),                     \
                      CG_DMA_I(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_MQ_STAGE), nullptr)
#define CG_REGISTER_MQ(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA) \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_MQ_STAGE, nullptr, nullptr, nullptr, nullptr)
#define CG_F32Q(REG, R, TI, TJ, TK, WPE)                \
  REG(float, float, 3, R, R, TI, TJ, TK, WPE, false)    \
  REG(float, float, 3, R, R, TI, TJ, TK, WPE, true)
CG_F32Q(CG_REGISTER_MQ, 2, 4, 8, 8, 1)
CG_F32Q(CG_REGISTER_MQ_D, 2, 4, 8, 8, 3)
CG_F32Q(CG_REGISTER_MQ_D, 2, 4, 4, 16, 3)
CG_F32Q(CG_REGISTER_MQ, 3, 4, 8, 8, 1)
CG_F32Q(CG_REGISTER_MQ_D, 3, 4, 8, 8, 3)
CG_F32Q(CG_REGISTER_MQ_D, 3, 4, 4, 16, 3)
}  // namespace cg
