// Instantiations of the stage kernel: FP32 split, orders 4 and 6.
#include "bssn_stage_kernel.h"

namespace cg {
#define CG_F32(R, TI, TJ, TK, WPE)                                   \
  CG_REGISTER_STAGE(float, float, 1, R, R, TI, TJ, TK, WPE, false)   \
  CG_REGISTER_STAGE(float, float, 1, R, R, TI, TJ, TK, WPE, true)
CG_F32(2, 4, 8, 8, 1)
CG_F32(2, 4, 8, 8, 3)
CG_F32(2, 2, 8, 16, 3)
CG_F32(2, 4, 4, 16, 3)
CG_F32(2, 8, 8, 8, 2)
CG_F32(3, 4, 8, 8, 1)
CG_F32(3, 4, 8, 8, 3)
CG_F32(3, 2, 8, 16, 3)
CG_F32(3, 4, 4, 16, 3)
CG_F32(3, 8, 8, 8, 2)
}  // namespace cg
