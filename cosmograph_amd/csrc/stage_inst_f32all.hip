// Instantiations of the stage kernel with all-FP32 storage (the conformal
// metric in FP32 too), orders 4 and 6: the default tile at one wave per SIMD
// and the 3-waves-per-SIMD tiles with their LDS-DMA twins.
#include "bssn_stage_kernel.h"

namespace cg {
#define CG_REGISTER_M32_D(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA)                                     \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_M32, nullptr,                     \
                      CG_DMA_L(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_M32),                          \
                      CG_DMA_I(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_M32), nullptr)
#define CG_REGISTER_M32(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA) \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_M32, nullptr, nullptr, nullptr, nullptr)
#define CG_F32A(REG, R, TI, TJ, TK, WPE)                \
  REG(float, float, 2, R, R, TI, TJ, TK, WPE, false)    \
  REG(float, float, 2, R, R, TI, TJ, TK, WPE, true)
CG_F32A(CG_REGISTER_M32, 2, 4, 8, 8, 1)
CG_F32A(CG_REGISTER_M32_D, 2, 4, 8, 8, 3)
CG_F32A(CG_REGISTER_M32_D, 2, 4, 4, 16, 3)
CG_F32A(CG_REGISTER_M32, 3, 4, 8, 8, 1)
CG_F32A(CG_REGISTER_M32_D, 3, 4, 8, 8, 3)
CG_F32A(CG_REGISTER_M32_D, 3, 4, 4, 16, 3)
}  // namespace cg
