// Instantiations of the stage kernel: FP32 split, orders 4 and 6.  Order 4
// (the batched-ensemble order) and the default order-6 tile also get the
// batched kernel; the tiles with 3 waves per SIMD get LDS-DMA twins.
#include "bssn_stage_kernel.h"

namespace cg {
#define CG_F32(REG, R, TI, TJ, TK, WPE)                 \
  REG(float, float, 1, R, R, TI, TJ, TK, WPE, false)    \
  REG(float, float, 1, R, R, TI, TJ, TK, WPE, true)
CG_F32(CG_REGISTER_STAGE_B, 2, 4, 8, 8, 1)
CG_F32(CG_REGISTER_STAGE_BD, 2, 4, 8, 8, 3)
CG_F32(CG_REGISTER_STAGE_BD, 2, 2, 8, 16, 3)
CG_F32(CG_REGISTER_STAGE_BD, 2, 4, 4, 16, 3)
CG_F32(CG_REGISTER_STAGE_B, 2, 8, 8, 8, 2)
CG_F32(CG_REGISTER_STAGE, 3, 4, 8, 8, 1)
CG_F32(CG_REGISTER_STAGE_BD, 3, 4, 8, 8, 3)
CG_F32(CG_REGISTER_STAGE_D, 3, 2, 8, 16, 3)
CG_F32(CG_REGISTER_STAGE_D, 3, 4, 4, 16, 3)
CG_F32(CG_REGISTER_STAGE, 3, 8, 8, 8, 2)
}  // namespace cg
