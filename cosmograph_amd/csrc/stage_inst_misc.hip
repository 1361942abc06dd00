// Instantiations of the stage kernel: orders 2 and 8 (with and without the
// wider Kreiss-Oliger halo), both precisions, default tile.
#include "bssn_stage_kernel.h"

namespace cg {
std::vector<StageVariant>& stage_variants() {
  static std::vector<StageVariant> v;
  return v;
}

#define CG_MISC(R, H)                                                     \
  CG_REGISTER_STAGE(double, double, 0, R, H, 4, 8, 8, 1, true)             \
  CG_REGISTER_STAGE(double, double, 0, R, H, 4, 8, 8, 1, false)            \
  CG_REGISTER_STAGE(float, float, 1, R, H, 4, 8, 8, 1, true)               \
  CG_REGISTER_STAGE(float, float, 1, R, H, 4, 8, 8, 1, false)
CG_MISC(1, 1)
CG_MISC(1, 2)
CG_MISC(4, 4)
CG_MISC(4, 5)
}  // namespace cg
