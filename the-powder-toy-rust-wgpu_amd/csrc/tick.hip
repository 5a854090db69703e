// Whole-tick driver and the per-pass C entry points.
//
// tpt_step follows SimulationEngine::step (crates/tpt-sim/src/lib.rs:303-459)
// for the passes this library implements: air -> Newtonian gravity (or the
// force-grid clear) -> RecalcFreeParticles (+ CheckStacking every 10th
// frame) -> conduction/transitions -> movement.  The element reactions
// (reaction_kernel::step_reactions_gpu, lib.rs:379) are not part of this
// library.

#include "tpt_internal.h"

static int require(tpt_ctx *ctx) {
    if (!ctx) return tpt_set_error(TPT_EINVAL, "null context");
    if (!ctx->tables_set) return tpt_set_error(TPT_ESTATE, "element table not set");
    return TPT_OK;
}

extern "C" int tpt_step_gravity(tpt_ctx *ctx, uint32_t seed) {
    int rc = require(ctx);
    return rc ? rc : tpt_gravity_step(ctx, seed);
}

extern "C" int tpt_gravity_solve(tpt_ctx *ctx) {
    int rc = require(ctx);
    return rc ? rc : tpt_gravity_solve_only(ctx);
}

extern "C" int tpt_recalc_free_particles(tpt_ctx *ctx, uint32_t seed, int stacking) {
    int rc = require(ctx);
    return rc ? rc : tpt_recalc_step(ctx, seed, stacking);
}

extern "C" int tpt_step_thermal(tpt_ctx *ctx, uint32_t seed, int legacy) {
    int rc = require(ctx);
    return rc ? rc : tpt_thermal_step(ctx, seed, legacy);
}

extern "C" int tpt_step_movement(tpt_ctx *ctx, uint32_t seed) {
    int rc = require(ctx);
    return rc ? rc : tpt_movement_step(ctx, seed);
}

extern "C" int tpt_step(tpt_ctx *ctx, uint64_t frame_count) {
    int rc = require(ctx);
    if (rc) return rc;
    const uint32_t seed = tpt_frame_seed(frame_count);
    if ((rc = tpt_air_step(ctx))) return rc;
    if ((rc = tpt_gravity_step(ctx, seed))) return rc;
    if ((rc = tpt_recalc_step(ctx, seed, (frame_count % 10u) == 0u))) return rc;
    if ((rc = tpt_thermal_step(ctx, seed, ctx->cfg.heat_enable ? 0 : 1))) return rc;
    return tpt_movement_step(ctx, seed);
}
