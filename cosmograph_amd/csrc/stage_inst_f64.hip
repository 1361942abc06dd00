// Instantiations of the stage kernel: FP64 drop-in, orders 4 and 6.  The
// order-4 tiles also get the persistent batched kernel.
#include "bssn_stage_kernel.h"

namespace cg {
#define CG_F64(REG, R, TI, TJ, TK, WPE)                  \
  REG(double, double, 0, R, R, TI, TJ, TK, WPE, false)   \
  REG(double, double, 0, R, R, TI, TJ, TK, WPE, true)
CG_F64(CG_REGISTER_STAGE_B, 2, 4, 8, 8, 1)
CG_F64(CG_REGISTER_STAGE_BD, 2, 4, 8, 8, 2)
CG_F64(CG_REGISTER_STAGE_BD, 2, 2, 8, 16, 2)
CG_F64(CG_REGISTER_STAGE_BD, 2, 4, 4, 16, 2)
CG_F64(CG_REGISTER_STAGE, 3, 4, 8, 8, 1)
CG_F64(CG_REGISTER_STAGE_D, 3, 4, 8, 8, 2)
CG_F64(CG_REGISTER_STAGE_D, 3, 2, 8, 16, 2)
CG_F64(CG_REGISTER_STAGE_D, 3, 4, 4, 16, 2)
}  // namespace cg
