// Instantiations of the constraint-diagnostics mode of the stage kernel
// (BSSN::setConstraintCalcs), FP64 arithmetic from either storage type;
// with the dust fluid's matter sources for dust contexts.
#include "bssn_stage_kernel.h"

namespace cg {

template <typename TS, bool LA, int MODE>
static int cons_by_order(int order, const StageArgs& a, hipStream_t s) {
  switch (order) {
  case 2: return cons_launch_t<TS, 1, LA, MODE>(a, s);
  case 4: return cons_launch_t<TS, 2, LA, MODE>(a, s);
  case 6: return cons_launch_t<TS, 3, LA, MODE>(a, s);
  case 8: return cons_launch_t<TS, 4, LA, MODE>(a, s);
  }
  return (int)hipErrorInvalidValue;
}

int launch_constraints(int prec, int order, bool live_alpha, bool dust, const StageArgs& a, hipStream_t s) {
  if (prec == 3) {  // fixed-point metric, other fields FP32
    if (dust)
      return live_alpha ? cons_by_order<float, true, MODE_MQ_CONS | MODE_DUST>(order, a, s)
                        : cons_by_order<float, false, MODE_MQ_CONS | MODE_DUST>(order, a, s);
    return live_alpha ? cons_by_order<float, true, MODE_MQ_CONS>(order, a, s)
                      : cons_by_order<float, false, MODE_MQ_CONS>(order, a, s);
  }
  if (prec == 2)  // all-FP32 storage (no dust fluid)
    return live_alpha ? cons_by_order<float, true, MODE_M32_CONS>(order, a, s)
                      : cons_by_order<float, false, MODE_M32_CONS>(order, a, s);
  if (dust) {
    if (prec == 0)
      return live_alpha ? cons_by_order<double, true, MODE_DUST_CONS>(order, a, s)
                        : cons_by_order<double, false, MODE_DUST_CONS>(order, a, s);
    return live_alpha ? cons_by_order<float, true, MODE_DUST_CONS>(order, a, s)
                      : cons_by_order<float, false, MODE_DUST_CONS>(order, a, s);
  }
  if (prec == 0)
    return live_alpha ? cons_by_order<double, true, MODE_CONS>(order, a, s)
                      : cons_by_order<double, false, MODE_CONS>(order, a, s);
  return live_alpha ? cons_by_order<float, true, MODE_CONS>(order, a, s)
                    : cons_by_order<float, false, MODE_CONS>(order, a, s);
}

}  // namespace cg
