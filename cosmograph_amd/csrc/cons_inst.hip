// Instantiations of the constraint-diagnostics mode of the stage kernel
// (BSSN::setConstraintCalcs), FP64 arithmetic from either storage type.
#include "bssn_stage_kernel.h"

namespace cg {

template <typename TS, bool LA>
static int cons_by_order(int order, const StageArgs& a, hipStream_t s) {
  switch (order) {
  case 2: return cons_launch_t<TS, 1, LA>(a, s);
  case 4: return cons_launch_t<TS, 2, LA>(a, s);
  case 6: return cons_launch_t<TS, 3, LA>(a, s);
  case 8: return cons_launch_t<TS, 4, LA>(a, s);
  }
  return (int)hipErrorInvalidValue;
}

int launch_constraints(int prec, int order, bool live_alpha, const StageArgs& a, hipStream_t s) {
  if (prec == 0) return live_alpha ? cons_by_order<double, true>(order, a, s) : cons_by_order<double, false>(order, a, s);
  return live_alpha ? cons_by_order<float, true>(order, a, s) : cons_by_order<float, false>(order, a, s);
}

}  // namespace cg
