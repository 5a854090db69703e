/*
 * CPU oracle for the tpt-sim tick -- TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded C restatement of the reference kernels, written from the
 * Rust/CubeCL source text of pppopipupu/The-Powder-Toy-Rust-wgpu.  Each
 * function cites the reference kernel it follows.  Only the tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load it; the product
 * library (libtpt_hip) never does.
 *
 * Numerics: built with -ffp-contract=off and no fast-math; expressions keep
 * the reference's order; float->int conversions saturate toward zero with
 * NaN -> 0 (the same convention as the HIP kernels).
 *
 * Races in the reference (non-atomic read-modify-writes) are replaced by one
 * canonical serialisation shared with the HIP kernels; each affected function
 * names the rule it implements.
 *
 * Pinning: every function is checked by tests/test_oracle_*.py against the
 * reference's own test values, closed forms, invariants or brute force on
 * small inputs.  Parity unpinned: the outcome of the reference's racy
 * kernels (which these canonical rules replace), its CubeCL lowering
 * details (contraction, division ULP), and the f32 twiddles of its FFT
 * (the solve here is FP64 and the HIP FFT is held to it within 1e-4).
 */
#ifndef TPT_ORACLE_H
#define TPT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Simulation state: every pointer is a host array in the reference layout
 * (slot arrays of `capacity` words, pixel maps of xres*yres words, cell grids
 * of xcells*ycells words).  Field names follow SimBuffers. */
typedef struct oracle_state {
    /* config (SimulationConfig, crates/tpt-core/src/particle.rs:166-346) */
    uint32_t xres, yres, capacity, slot_bound;
    uint32_t aheat_enable, heat_enable, newton_grav;
    uint32_t edge_mode, gravity_mode, air_mode, convection_mode, grav_mass_mode, water_equal_test;
    float ambient_air_temp, edge_pressure, edge_velocity_x, edge_velocity_y;
    float custom_gravity_x, custom_gravity_y, vorticity_coeff;
    /* slot arrays */
    uint32_t *part_ty; int32_t *part_life; uint32_t *part_ctype;
    float *part_x, *part_y, *part_vx, *part_vy, *part_temp, *part_equil;
    uint32_t *part_flags; int32_t *part_tmp, *part_tmp2, *part_tmp3, *part_tmp4;
    uint32_t *part_dcolour, *intent_target;
    /* pixel maps */
    uint32_t *pmap, *photons, *pmap_count, *cell_winner;
    /* cell grids */
    float *air_vx, *air_vy, *air_pv, *air_hv, *air_ovx, *air_ovy, *air_opv, *air_ohv;
    float *fan_vx, *fan_vy;
    uint32_t *wall_bmap, *wall_emap, *block_air, *block_airh, *elem_block_air, *elem_block_airh;
    uint32_t *grav_mask; int32_t *grav_mass, *grav_mass_prev;
    float *grav_fx, *grav_fy;
    /* tables */
    const uint32_t *elem_props;   /* 256 x 32 */
    const uint32_t *move_lut;     /* 256 x 12 + 256 x 256 */
    const uint32_t *photon_mask;  /* 256 */
    const uint32_t *spawn_defaults; /* 256 x 8 (to_gpu_spawn_defaults_u32) */
    /* gravity tool wells and element mass overrides (cell grids,
     * buffers/mod.rs:316-347) */
    float *grav_tool_mass; int32_t *grav_mass_override; uint32_t *grav_mass_override_set;
} oracle_state;

/* air_kernel::step_air_gpu (air_kernel/mod.rs:120-299). */
void oracle_step_air(oracle_state *s);

/* The individual air kernels, in launch order. */
void oracle_air_pressure(oracle_state *s);      /* air_kernel/pressure.rs:174-238 */
void oracle_air_velocity(oracle_state *s);      /* air_kernel/pressure.rs:248-299 */
void oracle_air_smooth_advect(oracle_state *s); /* air_kernel/advect.rs:176-360   */
void oracle_air_heat_convect(oracle_state *s);  /* air_kernel/heat.rs:31-253      */
void oracle_air_copy_back(oracle_state *s);     /* air_kernel/settle.rs:22-28     */
void oracle_air_block_masks(oracle_state *s);   /* air_kernel/settle.rs:51-93     */
void oracle_hot_air_pressure(oracle_state *s);  /* air_kernel/settle.rs:119-160   */

/* far_advect_target (air_kernel/advect.rs:66-152); exposed for unit tests. */
void oracle_far_advect_target(uint32_t x, uint32_t y, float dx, float dy, const uint32_t *block,
                              uint32_t xcells, uint32_t ycells, uint32_t heat, float *tx, float *ty);

/* Gravity (gravity_gpu/{mod,fft}.rs, gravity_kernel.rs). */
void oracle_fft_twiddle_f32(uint32_t max_len, float *re, float *im);
void oracle_fft_transform_f32(float *data, uint32_t rows, uint32_t len, uint32_t inner, uint32_t stride,
                              const float *tw_re, const float *tw_im, uint32_t tw_bits, float sign);
void oracle_gravity_solve(oracle_state *s);
void oracle_gravity_solve_f64(const oracle_state *s, double *fx, double *fy);
void oracle_build_gravity_mass(oracle_state *s);
void oracle_apply_gravity(oracle_state *s);
void oracle_step_gravity(oracle_state *s);
/* The element pass's clear of the mass override channel
 * (reaction_kernel/modext/mod.rs:53-99), the only part of
 * step_reactions_gpu the tick here performs. */
void oracle_clear_grav_override(oracle_state *s);
void oracle_gravity_mask(const uint32_t *wall, uint32_t xc, uint32_t yc, uint32_t *mask);

/* RecalcFreeParticles + CheckStacking (buffers/spatial.rs:81-472). */
uint32_t oracle_lowbias_hash(uint32_t seed);
void oracle_recalc_free_particles(oracle_state *s, uint32_t seed, int stacking);

/* Conduction, apply, phase transitions, pmap refresh (thermal_kernel.rs). */
uint32_t oracle_gpu_hash(uint32_t seed);
void oracle_thermal_conduction(oracle_state *s, uint32_t seed);
void oracle_thermal_apply(oracle_state *s);
void oracle_phase_transitions(oracle_state *s, uint32_t seed, int legacy);
void oracle_refresh_pmap_types(oracle_state *s);
void oracle_step_thermal(oracle_state *s, uint32_t seed, int legacy);

/* Movement (move_kernel/{mod,velocity,liquid,commit}.rs). */
uint32_t oracle_pcg_hash(uint32_t seed);
void oracle_air_deposit(oracle_state *s);
void oracle_detect_charge(oracle_state *s);
void oracle_emap_closure(oracle_state *s);
void oracle_velocity_intent(oracle_state *s, uint32_t seed);
void oracle_commit(oracle_state *s);
void oracle_rebuild_maps(oracle_state *s);
void oracle_step_movement(oracle_state *s, uint32_t seed);
/* water_equal_gpu::step_water_equal_gpu (water_equal_gpu/mod.rs:67-166). */
void oracle_water_equal(oracle_state *s, uint32_t seed);
/* hash_u32 (reaction_kernel/shared.rs:1015-1023) and
 * Element_FILT_interactWavelengths (reaction_kernel/rays/emitter.rs:28-80),
 * used by the e == 2 pair pass; exposed for known-answer tests. */
uint32_t oracle_hash_u32(uint32_t x);
uint32_t oracle_interact_wavelengths(int32_t filt_tmp, uint32_t filt_ctype, float filt_temp, uint32_t orig_wl);

/* Of reaction_kernel::step_reactions_gpu (reaction_kernel/mod.rs:1123-1237),
 * in the driver's order: the modext clear of the mass-override channel, the
 * life pass (life.rs:47-205), the combustion pass's FIRE/PLSM heating and
 * ignition and WTRV branches (combustion/fire.rs:26-408, 983-1013) and the
 * fluid pass's WATR/DSTW/SLTW/ICEI/SNOW branches (fluid.rs:149-400), the
 * last two under the canonical rule stated in reactions.c.  The other
 * reaction passes and element branches are not restated. */
void oracle_life_pass(oracle_state *s, uint32_t seed);
void oracle_combustion_pass(oracle_state *s, uint32_t seed);
void oracle_fluid_pass(oracle_state *s, uint32_t seed);
void oracle_step_reactions(oracle_state *s, uint32_t seed);

/* SimulationEngine::step (lib.rs:303-459) over the passes above. */
void oracle_step(oracle_state *s, uint64_t frame_count);

#ifdef __cplusplus
}
#endif
#endif
