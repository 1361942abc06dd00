// Instantiations of the stage kernel with the dust fluid
// (components/dust_fluid/dust.cc): FP64 and FP32-split storage, orders 4 and 6.
#include "bssn_stage_kernel.h"

namespace cg {
#define CG_DUST(R, TI, TJ, TK, WPE)                                   \
  CG_REGISTER_DUST(double, double, 0, R, R, TI, TJ, TK, WPE, false)  \
  CG_REGISTER_DUST(double, double, 0, R, R, TI, TJ, TK, WPE, true)   \
  CG_REGISTER_DUST(float, float, 1, R, R, TI, TJ, TK, WPE, false)    \
  CG_REGISTER_DUST(float, float, 1, R, R, TI, TJ, TK, WPE, true)
CG_DUST(2, 4, 8, 8, 1)
CG_DUST(3, 4, 8, 8, 1)
}  // namespace cg
