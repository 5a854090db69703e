/*
 * tpt_hip.h -- C ABI of the MI355X (gfx950) implementation of tpt-sim's
 * per-tick simulation step.
 *
 * Every entry point replaces one Rust function of the reference's tick path;
 * the citation on each declaration names the function (crate-relative path,
 * file:line) whose behaviour it reproduces.  Buffers keep the reference's
 * structure-of-arrays layout and word formats (SimBuffers,
 * crates/tpt-sim/src/buffers/mod.rs:81-393): slot arrays are indexed by slot
 * id, pixel maps by y*xres+x, cell grids by cy*xcells+cx (cells are 4x4 px),
 * floats are IEEE-754 binary32, pmap words are ((slot+1)<<9)|type.
 *
 * Conventions
 *  - All functions return 0 on success or a negative TPT_E* code.  Invalid
 *    dimensions, lengths or ids are rejected with TPT_EINVAL and leave the
 *    state untouched (the reference's step functions silently no-op on the
 *    same inputs, e.g. crates/tpt-sim/src/air_kernel/mod.rs:124-126).
 *  - One context owns one HIP stream; calls on one context are not
 *    thread-safe.  Device work is asynchronous: tpt_download and
 *    tpt_synchronize wait for it.
 *  - No PyTorch, CUDA or other framework types cross this boundary.
 */
#ifndef TPT_HIP_H
#define TPT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#define TPT_API __attribute__((visibility("default")))
#else
#define TPT_API
#endif

#define TPT_OK 0
#define TPT_EINVAL (-1)   /* bad argument, dimension, length or buffer id   */
#define TPT_ENOMEM (-2)   /* device allocation failed                        */
#define TPT_EDEVICE (-3)  /* HIP runtime error (launch, copy, sync)          */
#define TPT_ESTATE (-4)   /* call not valid in the context's current state   */

#define TPT_PT_NUM 256            /* element slots (constants.rs:60)           */
#define TPT_ELEM_STRIDE 32        /* u32 words per GpuElementProps row         */
#define TPT_MOVE_META_WORDS 12    /* u32 words per element in move_lut         */
#define TPT_MOVE_LUT_WORDS (TPT_PT_NUM * TPT_MOVE_META_WORDS + TPT_PT_NUM * TPT_PT_NUM)
#define TPT_MAX_CAPACITY 0x7FFFFFu /* 23-bit pmap slot field (layout.rs:21-31) */
#define TPT_MAX_DIM 2896u          /* largest canvas side (layout.rs:33-40)    */
#define TPT_SPAWN_DEFAULT_WORDS 8  /* u32 words per spawn-defaults row          */
#define TPT_GPU_SPAWN_START 1024u  /* device allocation window, slots           */
#define TPT_GPU_SPAWN_END 17408u   /* [START, END) (buffers/mod.rs:502-514)     */
#define TPT_GPU_CREATE_COUNT 16384u /* creations drained per tick (mod.rs:514)  */

/* Mirror of the simulation-relevant fields of SimulationConfig
 * (crates/tpt-core/src/particle.rs:166-346).  Enums use the reference's
 * discriminants: edge_mode Void=0 Solid=1 Loop=2; gravity_mode Vertical=0
 * Off=1 Radial=2 Custom=3; air_mode On=0 PressureOff=1 VelocityOff=2 Off=3
 * NoUpdate=4; convection_mode 0=none 1=legacy 2=Boussinesq. */
typedef struct tpt_config {
    uint32_t xres;              /* canvas width in px, multiple of 4        */
    uint32_t yres;              /* canvas height in px, multiple of 4       */
    uint32_t capacity;          /* particle pool slots (<= TPT_MAX_CAPACITY) */
    uint32_t aheat_enable;
    uint32_t heat_enable;
    uint32_t newton_grav;
    uint32_t edge_mode;
    uint32_t gravity_mode;
    uint32_t air_mode;
    uint32_t convection_mode;
    float ambient_air_temp;
    float edge_pressure;
    float edge_velocity_x;
    float edge_velocity_y;
    float custom_gravity_x;
    float custom_gravity_y;
    float vorticity_coeff;
    /* 0: reference gravity-mass table (only GRVT/NBHL/NWHL/GBMB/GPMP deposit,
     *    gravity_kernel.rs:275-362);
     * 1: extension -- every live particle deposits its element weight
     *    (GpuElementProps.weight) as mass.  No reference counterpart. */
    uint32_t grav_mass_mode;
    /* SimulationConfig::water_equal_test (particle.rs:209): run the water
     * equalisation pass after movement (lib.rs:418-420). */
    uint32_t water_equal_test;
} tpt_config;

/* Device buffers, named after the SimBuffers fields they mirror. */
typedef enum tpt_buf_id {
    /* slot arrays: capacity words each (buffers/mod.rs:81-140) */
    TPT_BUF_PART_TY = 0, TPT_BUF_PART_LIFE, TPT_BUF_PART_CTYPE, TPT_BUF_PART_X,
    TPT_BUF_PART_Y, TPT_BUF_PART_VX, TPT_BUF_PART_VY, TPT_BUF_PART_TEMP,
    TPT_BUF_PART_EQUIL, TPT_BUF_PART_FLAGS, TPT_BUF_PART_TMP, TPT_BUF_PART_TMP2,
    TPT_BUF_PART_TMP3, TPT_BUF_PART_TMP4, TPT_BUF_PART_DCOLOUR,
    TPT_BUF_INTENT_TARGET,
    /* pixel maps: xres*yres words each */
    TPT_BUF_PMAP, TPT_BUF_PHOTONS, TPT_BUF_PMAP_COUNT, TPT_BUF_CELL_WINNER,
    /* cell grids: xcells*ycells words each */
    TPT_BUF_AIR_VX, TPT_BUF_AIR_VY, TPT_BUF_AIR_PV, TPT_BUF_AIR_HV,
    TPT_BUF_AIR_OVX, TPT_BUF_AIR_OVY, TPT_BUF_AIR_OPV, TPT_BUF_AIR_OHV,
    TPT_BUF_FAN_VX, TPT_BUF_FAN_VY, TPT_BUF_WALL_BMAP, TPT_BUF_WALL_EMAP,
    TPT_BUF_BLOCK_AIR, TPT_BUF_BLOCK_AIRH, TPT_BUF_ELEM_BLOCK_AIR,
    TPT_BUF_ELEM_BLOCK_AIRH, TPT_BUF_GRAV_MASK, TPT_BUF_GRAV_MASS,
    TPT_BUF_GRAV_MASS_PREV, TPT_BUF_GRAV_FX, TPT_BUF_GRAV_FY,
    /* PGRV/NGRV tool wells, f32 mass per cell (SimBuffers::grav_tool_mass,
       buffers/mod.rs:316-325).  tpt_step_gravity folds every non-zero cell
       into the snapshot its solve reads and zeroes it
       (seed_gravity_tool_mass_kernel, gravity_kernel.rs:241-262). */
    TPT_BUF_GRAV_TOOL_MASS,
    /* One-shot element mass assignments (CEXP, GSNS), i32 in 1/65536 units,
       and their write flags (gpu_grav_mass_override{,_set},
       buffers/mod.rs:326-347).  tpt_step_gravity writes each flagged value
       over the mass it built (apply_grav_mass_override_kernel,
       gravity_kernel.rs:379-391); tpt_step zeroes both grids where the
       reference's element pass clears them, between the thermal and
       movement passes (reaction_kernel/modext/mod.rs:53-99). */
    TPT_BUF_GRAV_MASS_OVERRIDE, TPT_BUF_GRAV_MASS_OVERRIDE_SET,
    /* tables */
    TPT_BUF_ELEM_PROPS,  /* TPT_PT_NUM*TPT_ELEM_STRIDE words   */
    TPT_BUF_MOVE_LUT,    /* TPT_MOVE_LUT_WORDS words            */
    TPT_BUF_PHOTON_MASK, /* TPT_PT_NUM words: PhotonReflectWavelengths
                            (build_photon_mask, move_kernel/mod.rs:1592-1606) */
    TPT_BUF_SPAWN_DEFAULTS, /* TPT_PT_NUM*TPT_SPAWN_DEFAULT_WORDS words:
                            DefaultProperties rows {temp bits, life, ctype,
                            tmp, tmp2, tmp3, tmp4, 0}
                            (ElementRegistry::to_gpu_spawn_defaults_u32,
                            tpt-core/src/elements/mod.rs:905-936) */
    TPT_BUF_COUNT
} tpt_buf_id;

typedef struct tpt_ctx tpt_ctx;

/* Error text of the last failing call on this thread ("" if none). */
TPT_API const char *tpt_last_error(void);

/* SimulationEngine::new (crates/tpt-sim/src/lib.rs:140) together with
 * SimBuffers::new (crates/tpt-sim/src/buffers/layout.rs:42-257): allocates
 * every buffer for the canvas, zero-filled, and creates the context's
 * stream.  The reference's host shadows stay with the caller. */
TPT_API int tpt_create(const tpt_config *cfg, tpt_ctx **out);
TPT_API int tpt_destroy(tpt_ctx *ctx);

/* Replaces SimBuffers::upload_element_tables + move_kernel::refresh_move_tables
 * (crates/tpt-sim/src/move_kernel/mod.rs:1578-1590): the caller passes the
 * GpuElementProps rows (elements/mod.rs:482-556) and the packed movement LUT
 * (build_move_lut, move_kernel/mod.rs:1534-1576). */
TPT_API int tpt_set_element_table(tpt_ctx *ctx, const uint32_t *props, size_t props_words,
                          const uint32_t *move_lut, size_t lut_words);

/* SimBuffers::sync_host_to_gpu / sync_gpu_to_host for one buffer
 * (crates/tpt-sim/src/buffers/gpu_sync.rs:31, 244).  Byte offset and length
 * must lie inside the buffer and be multiples of 4. */
TPT_API int tpt_upload(tpt_ctx *ctx, int buf, const void *src, size_t offset_bytes, size_t bytes);
TPT_API int tpt_download(tpt_ctx *ctx, int buf, void *dst, size_t offset_bytes, size_t bytes);
TPT_API int tpt_buffer_bytes(tpt_ctx *ctx, int buf, size_t *bytes);

/* SimBuffers::used_slot_bound (crates/tpt-sim/src/buffers/mod.rs:552-559):
 * slot passes scan [0, bound).  The host owns slot allocation, so it sets
 * the bound after loading particles; clamped to [1, capacity]. */
TPT_API int tpt_set_slot_bound(tpt_ctx *ctx, uint32_t bound);
/* The current bound.  A movement pass raises it to cover the device
 * allocation window [TPT_GPU_SPAWN_START, TPT_GPU_SPAWN_END) (clamped to the
 * capacity), as the reference's `gpu_spawn_used` does once the movement
 * step has armed its creation queue (move_kernel/mod.rs:1212-1216). */
TPT_API int tpt_get_slot_bound(tpt_ctx *ctx, uint32_t *bound);

/* Per-pass entry points, in tick order. */

/* air_kernel::step_air_gpu (crates/tpt-sim/src/air_kernel/mod.rs:120-299). */
TPT_API int tpt_step_air(tpt_ctx *ctx);

/* gravity_kernel::step_gravity_gpu (crates/tpt-sim/src/gravity_kernel.rs:826-1016)
 * including GravityFft::solve (crates/tpt-sim/src/gravity_gpu/mod.rs:114).
 * When newton_grav is 0 the force grids are zeroed
 * (SimBuffers::zero_gravity_forces, lib.rs:344-352). */
TPT_API int tpt_step_gravity(tpt_ctx *ctx, uint32_t seed);

/* GravityFft::solve alone (gravity_gpu/mod.rs:114-135): solves the force
 * grids from TPT_BUF_GRAV_MASS_PREV (fixed point, 1/65536 units). */
TPT_API int tpt_gravity_solve(tpt_ctx *ctx);

/* No counterpart in the reference.  tpt_gravity_solve for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): it allocates nothing and
 * takes no context.  Solves fx, fy (xcells*ycells floats each) from the
 * fixed-point mass grid (i32, 1/65536 units) and the gravity mask (u32), all
 * device pointers, on `stream` (a hipStream_t; NULL is the default stream).
 * The FFT's buffers live in a 16.8 MB static device array that each GPU the
 * library runs on holds until the process ends; padded grids beyond
 * 1024x1024 (xcells or ycells above 512, canvases beyond 2048x2048 px) are
 * rejected with TPT_EINVAL.  Calls that overlap on one GPU race on that
 * array. */
TPT_API int tpt_gravity_solve_noalloc(const int32_t *mass, const uint32_t *mask, float *fx, float *fy,
                                      uint32_t xcells, uint32_t ycells, void *stream);

/* SimBuffers::recalc_free_particles_device
 * (crates/tpt-sim/src/buffers/spatial.rs:371-472). */
TPT_API int tpt_recalc_free_particles(tpt_ctx *ctx, uint32_t seed, int stacking);

/* thermal_kernel::step_thermal_and_transitions_gpu
 * (crates/tpt-sim/src/thermal_kernel.rs:1574-1717). */
TPT_API int tpt_step_thermal(tpt_ctx *ctx, uint32_t seed, int legacy);

/* reaction_kernel::step_reactions_gpu (crates/tpt-sim/src/reaction_kernel/mod.rs:1123-1237)
 * as far as this library implements it, in the driver's order: the modext
 * clear of the mass override channel (modext/mod.rs:53-99); the life pass
 * (life_and_local_reaction_kernel, reaction_kernel/life.rs:47-205:
 * fire/plasma/TRON burn-in lifetimes, fire burning out to smoke or steam,
 * plasma to noble gas or steam, sparks settling into their conductor, the
 * element kill at zero life, the stasis flag); the combustion pass's
 * FIRE/PLSM heating and 5x5 ignition with the COAL/BCOL timer, SPNG and its
 * WTRV branch (combustion/fire.rs:26-408, 983-1013); the fluid pass's WATR,
 * DSTW, SLTW, ICEI and SNOW branches (fluid.rs:149-400).  The two racing
 * passes follow one canonical order (ascending acting pixel, actors decide
 * from the pass's starting state, a target keeps its first conversion).
 * The other element branches of these kernels (wood, H2 fusion, yeast, ...)
 * and the energy, electronics and other element passes are not
 * implemented. */
TPT_API int tpt_step_reactions(tpt_ctx *ctx, uint32_t seed);

/* move_kernel::step_movement_gpu (crates/tpt-sim/src/move_kernel/mod.rs:1176-1227),
 * including the e == 2 pair pass between the intent and commit launches
 * (reaction_kernel/energy_entry.rs:90-393) and the drain of the photons it
 * creates (begin/drain_pending_creates, reaction_kernel/mod.rs:138-157,
 * create_alloc/create_fields_kernel, reaction_kernel/shared.rs:516-912).
 * The pair pass needs TPT_BUF_SPAWN_DEFAULTS. */
TPT_API int tpt_step_movement(tpt_ctx *ctx, uint32_t seed);

/* water_equal_gpu::step_water_equal_gpu (crates/tpt-sim/src/water_equal_gpu/mod.rs:67-166):
 * the communicating-vessel pass after movement, a no-op unless
 * water_equal_test is set.  Reads the cell grids as the preceding movement
 * pass left them. */
TPT_API int tpt_step_water_equal(tpt_ctx *ctx, uint32_t seed);

/* SimulationEngine::step (crates/tpt-sim/src/lib.rs:303-459) for the passes
 * above (the reaction step between thermal and movement), with seed = rotl(frame_count*0x9E3779B9, 13) ^ 0x85EBCA6B
 * (lib.rs:315-318) and CheckStacking on every 10th frame (lib.rs:359). */
TPT_API int tpt_step(tpt_ctx *ctx, uint64_t frame_count);

/* `ticks` calls of tpt_step for frames first_frame, first_frame+1, ...
 * (the loop of SimulationEngine::step callers, tpt-app main.rs:656-672).
 * The first tick runs as tpt_step; the steady-state tick is then captured
 * once as a hipGraph (re-captured after any upload, table or slot-bound
 * change) and replayed, with the frame counter, seed and CheckStacking
 * flag kept in device memory.  Results equal the tpt_step loop. */
TPT_API int tpt_run(tpt_ctx *ctx, uint64_t first_frame, uint32_t ticks);

/* The batched form (no counterpart in the reference, which steps one
 * canvas per SimulationEngine): n independent canvases, each a tpt_ctx
 * with its own buffers and stream, advanced together.  tpt_batch_step runs
 * one eager tick per canvas, then replays captured ticks: by default each
 * canvas's tick is its own hipGraph, replayed on the canvas's stream with
 * the canvases side by side (tpt_run per canvas); with lanes set
 * (tpt_batch_set_lanes), one hipGraph holds every canvas's tick, the
 * canvases dealt to that many serial chains (fork/join across the canvases'
 * streams).  Canvases are loaded and read through tpt_batch_context and the
 * single-context calls. */
typedef struct tpt_batch tpt_batch;
TPT_API int tpt_batch_create(uint32_t n, const tpt_config *cfgs, tpt_batch **out);
TPT_API int tpt_batch_destroy(tpt_batch *batch);
TPT_API int tpt_batch_context(tpt_batch *batch, uint32_t k, tpt_ctx **ctx);
TPT_API int tpt_batch_step(tpt_batch *batch, uint64_t first_frame, uint32_t ticks);
/* The same ticks as per-canvas tpt_step launches, each canvas on its own
 * stream (canvases overlap on the GPU). */
TPT_API int tpt_batch_step_launches(tpt_batch *batch, uint64_t first_frame, uint32_t ticks);
/* The same ticks one canvas at a time, as the reference steps one
 * SimulationEngine after another (tpt-app main.rs:656-672): each canvas's
 * tick waits on the GPU for the previous canvas's. */
TPT_API int tpt_batch_step_serial(tpt_batch *batch, uint64_t first_frame, uint32_t ticks);
/* 0 (the default): one graph per canvas, on the canvas's stream.  Otherwise
 * one graph for the whole batch whose canvases go, largest first, to
 * `lanes` serial chains that run side by side (lanes >= n: one chain per
 * canvas). */
TPT_API int tpt_batch_set_lanes(tpt_batch *batch, uint32_t lanes);
/* Autotuner: one eager tick, then trial_ticks graph replays at each
 * candidate lane count of the whole-batch graph (1, 2, 4, ... and one per
 * canvas), timed with HIP events; keeps the fastest (*lanes_out) and its
 * graph.  The canvases
 * advance through all of these ticks.  Currently switched off: it returns
 * TPT_ESTATE (one run aborted inside the HIP runtime with no cause found;
 * tpt_batch_set_lanes sets a split by hand). */
TPT_API int tpt_batch_autotune(tpt_batch *batch, uint64_t first_frame, uint32_t trial_ticks, uint32_t *lanes_out);
/* The autotuner's record: lane counts and milliseconds per batch tick. */
TPT_API int tpt_batch_tuning(tpt_batch *batch, uint32_t *lanes, float *ms_per_tick, uint32_t max_entries,
                             uint32_t *n_entries);
TPT_API int tpt_batch_synchronize(tpt_batch *batch);

/* ---------------------------------------------------------------------------
 * The learned air surrogate (no counterpart in the reference): a small
 * convolutional U-Net (32/64/128 channels, 3x3 convolutions, GELU; bf16
 * weights on MFMA, fp32 accumulation) that advances a context's air grids
 * (pressure, velocity, ambient heat) by 8 ticks per call in place of 8 calls
 * of air_kernel::step_air_gpu (air_kernel/mod.rs:120-299).  It is trained on
 * trajectories of that plain air step by its own forward and backward
 * kernels.  Scenes for training and evaluation carry heat sources: cells
 * whose ambient heat the plain ticks (tpt_sur_plain_ticks) and the
 * surrogate both hold at a fixed temperature. */
#define TPT_SUR_MAX_BLOBS 16
#define TPT_SUR_MAX_RECTS 32
#define TPT_SUR_MAX_SOURCES 8
typedef struct tpt_sur_scene_desc {
    uint32_t n_blobs, n_rects, n_sources, reserved;
    float blob[TPT_SUR_MAX_BLOBS][4];        /* centre x, y (cells), amplitude, sigma (cells) */
    uint32_t rect[TPT_SUR_MAX_RECTS][4];     /* WL_WALL cells [x0, x1) x [y0, y1): x0, y0, x1, y1 */
    uint32_t source[TPT_SUR_MAX_SOURCES][4]; /* heat-source cells, same form */
    float source_temp[TPT_SUR_MAX_SOURCES];  /* K */
} tpt_sur_scene_desc;

typedef struct tpt_sur tpt_sur;
/* A network for an xcells x ycells air grid (multiples of 4), He-initialised. */
TPT_API int tpt_sur_create(uint32_t xcells, uint32_t ycells, tpt_sur **out);
TPT_API int tpt_sur_destroy(tpt_sur *sur);
/* Parameters: one fp32 vector (every layer's [cout][k] matrix, bias in
 * column 9*cin); set resets the optimiser state. */
TPT_API int tpt_sur_param_count(tpt_sur *sur, size_t *n);
TPT_API int tpt_sur_init_params(tpt_sur *sur, uint64_t seed);
TPT_API int tpt_sur_set_params(tpt_sur *sur, const float *params, size_t n);
TPT_API int tpt_sur_get_params(tpt_sur *sur, float *params, size_t n);
TPT_API int tpt_sur_get_grads(tpt_sur *sur, float *grads, size_t n);
/* Normalisation of pressure, velocity and ambient heat: input channel c is
 * field / in_scales[c] (heat as (hv - ambient) / in_scales[3]); output
 * channel c is the 8-tick change / out_scales[c].  vx and vy use index 1. */
TPT_API int tpt_sur_set_scales(tpt_sur *sur, const float in_scales[4], const float out_scales[4], float ambient);
/* Host-buffer forward (x0: H x W x 8 bf16 bits in, out: H x W x 8 fp32) and
 * backward (g_out: H x W x 8 bf16 bits; the weight gradients replace the
 * gradient vector) of the network alone. */
TPT_API int tpt_sur_forward_raw(tpt_sur *sur, const uint16_t *x0, float *out);
TPT_API int tpt_sur_backward_raw(tpt_sur *sur, const uint16_t *g_out);
/* A scene on a context with the surrogate's grid: blobs of pressure, zero
 * velocity, ambient heat, WL_WALL rectangles and border ring, heat sources. */
TPT_API int tpt_sur_scene(tpt_sur *sur, tpt_ctx *ctx, const tpt_sur_scene_desc *scene);
/* `ticks` plain air steps (tpt_step_air) with the heat sources held. */
TPT_API int tpt_sur_plain_ticks(tpt_sur *sur, tpt_ctx *ctx, uint32_t ticks);
/* Training set: n_traj trajectories of n_snap snapshots 8 plain ticks apart. */
TPT_API int tpt_sur_dataset_alloc(tpt_sur *sur, uint32_t n_traj, uint32_t n_snap);
TPT_API int tpt_sur_dataset_capture(tpt_sur *sur, tpt_ctx *ctx, uint32_t traj, uint32_t snap);
/* Adam training on (trajectory, snapshot) pairs, predicting snapshot + 1;
 * gradients summed over `accum` pairs per step. */
TPT_API int tpt_sur_train(tpt_sur *sur, const uint32_t *pairs, uint32_t n_pairs, uint32_t accum, float lr,
                          float *mean_loss);
TPT_API int tpt_sur_eval_loss(tpt_sur *sur, const uint32_t *pairs, uint32_t n_pairs, float *mean_loss);
/* Bound on the global L2 norm of each step's mean gradient (default 1; 0
 * turns clipping off). */
TPT_API int tpt_sur_set_clip(tpt_sur *sur, float clip);
/* Training horizon of tpt_sur_train, 1 .. 8: 1 (the default) trains each
 * pair (traj, k) on snapshot k -> k+1; r > 1 first advances snapshot k
 * r - 1 times by the network itself, without gradient, and trains on that
 * prediction -> snapshot k+r (pairs need k+r < n_snap).  No counterpart in
 * the reference. */
TPT_API int tpt_sur_set_rollout(tpt_sur *sur, uint32_t steps);
/* Inference with the pools, upsample-adds and input preparation folded into
 * the convolutions that read them (fused = 1, the default) or run as
 * separate kernels (0); the two give identical bits.  No counterpart in
 * the reference (the surrogate has none). */
TPT_API int tpt_sur_set_fused(tpt_sur *sur, int fused);
/* 8 air ticks of the context by the network.  With `check`, a prediction
 * with a non-finite value or whose energy sum(p^2 + vx^2 + vy^2) exceeds
 * energy_ratio times the input's is replaced by 8 plain ticks from the
 * input; *fell_back (nullable) says which ran. */
TPT_API int tpt_sur_step_air(tpt_sur *sur, tpt_ctx *ctx, int check, float energy_ratio, int *fell_back);
/* The last checked call's energy before and after and its non-finite count. */
TPT_API int tpt_sur_last_check(tpt_sur *sur, float out[3]);

/* The tick seed of lib.rs:315-318. */
TPT_API uint32_t tpt_frame_seed(uint64_t frame_count);

/* Waits for all work queued on the context's stream. */
TPT_API int tpt_synchronize(tpt_ctx *ctx);

/* Pass timing with HIP events on the context's stream (no reference
 * counterpart; the reference only times whole frames, tpt-app main.rs:637-726).
 * When enabled every kernel launch is bracketed by events; totals accumulate
 * until tpt_reset_pass_times.  Names are static strings. */
TPT_API int tpt_set_profiling(tpt_ctx *ctx, int enable);
TPT_API int tpt_reset_pass_times(tpt_ctx *ctx);
TPT_API int tpt_get_pass_times(tpt_ctx *ctx, const char **names, double *total_ms,
                       uint64_t *launches, int max_entries, int *n_entries);

#ifdef __cplusplus
}
#endif
#endif /* TPT_HIP_H */
