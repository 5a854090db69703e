// Passes whose HIP kernels are not built yet report TPT_ESTATE instead of
// silently doing nothing.

#include "tpt_internal.h"

int tpt_thermal_step(tpt_ctx *, uint32_t, int) { return tpt_set_error(TPT_ESTATE, "thermal pass not built"); }
int tpt_movement_step(tpt_ctx *, uint32_t) { return tpt_set_error(TPT_ESTATE, "movement pass not built"); }
