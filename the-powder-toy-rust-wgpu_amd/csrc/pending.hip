// Passes whose HIP kernels are not built yet report TPT_ESTATE instead of
// silently doing nothing.

#include "tpt_internal.h"

int tpt_gravity_step(tpt_ctx *, uint32_t) { return tpt_set_error(TPT_ESTATE, "gravity pass not built"); }
int tpt_gravity_solve_only(tpt_ctx *) { return tpt_set_error(TPT_ESTATE, "gravity pass not built"); }
int tpt_gravity_free(tpt_ctx *) { return TPT_OK; }
int tpt_recalc_step(tpt_ctx *, uint32_t, int) { return tpt_set_error(TPT_ESTATE, "recalc pass not built"); }
int tpt_thermal_step(tpt_ctx *, uint32_t, int) { return tpt_set_error(TPT_ESTATE, "thermal pass not built"); }
int tpt_movement_step(tpt_ctx *, uint32_t) { return tpt_set_error(TPT_ESTATE, "movement pass not built"); }
