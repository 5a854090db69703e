// Internal declarations shared by the HIP translation units of libtpt_hip.
//
// Numerics contract for every kernel in this library: no FMA contraction
// (built with -ffp-contract=off), IEEE binary32 division and square root
// (hipcc's correctly rounded defaults), expressions evaluated in the order
// the reference writes them, and float->int conversions through tpt_f2i
// (truncation toward zero, saturating, NaN -> 0).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/tpt_hip.h"

#define TPT_CELL 4
#define TPT_CREATE_REC_WORDS 6  // requester slot, x | y << 16, ctype, temp, vx, vy
#define TPT_SPAWN_BITMAP_WORDS ((TPT_GPU_SPAWN_END - TPT_GPU_SPAWN_START) / 32u)

// Slots of the small device counter block (tpt_ctx::counters).
#define TPT_CTR_INVIS 0   // INVIS/FILT placement queue length
#define TPT_CTR_SAW 1     // deferred WOOD->SAWD list length
#define TPT_CTR_OFFMAP 2  // off-map list length (see spatial.hip)
#define TPT_CTR_EMAP_GATE 3  // wire closure gate bits (move.hip)
#define TPT_CTR_PHOT 4       // blocked-photon hit list length (move.hip)
#define TPT_CTR_PAIR 5       // e == 2 pair candidate list length (move.hip)
#define TPT_CTR_CREATE 6     // e == 2 creation records queued this tick (move.hip)
#define TPT_CTR_WEQ 7        // water equalisation bidder list length (move.hip)
#define TPT_CTR_FRAME_LO 8   // next frame_count, low / high word (tick.hip)
#define TPT_CTR_FRAME_HI 9
#define TPT_CTR_SEED 10      // seed of the pass(es) running now
#define TPT_CTR_STACK 11     // 1: CheckStacking this tick
#define TPT_CTR_RX_STAMP 12  // element passes: stamp of the current step's burner records (reactions.hip)
#define TPT_CTR_CONFLICT 13  // map scatter: slots whose pixel another slot also wrote (spatial.hip)

// Passes that carry their own timing slot (see tpt_get_pass_times).
enum tpt_pass {
    PASS_AIR_PV = 0,      // pressure + velocity, fused
    PASS_AIR_ADVECT,      // smoothing, far advection, vorticity, fans, caps
    PASS_AIR_HEAT,        // ambient heat diffusion/advection + convection
    PASS_AIR_SETTLE,      // copy-back, wall masks, emap decay, HotAir
    PASS_GRAV_MASS,       // tool seed / mass snapshot bookkeeping
    PASS_GRAV_FFT_ROWS,   // padded forward row FFT
    PASS_GRAV_FFT_COLS,   // forward column FFT * kernel spectra, inverse column FFT
    PASS_GRAV_FFT_INV,    // inverse row FFT + force extraction
    PASS_GRAV_APPLY,      // Loss + Newtonian force on particles
    PASS_GRAV_BUILD,      // mass snapshot copy (element-weight mass mode; the deposit runs in grav_apply)
    PASS_RECALC_CLEAR,    // pmap/photons/pmap_count clear
    PASS_RECALC_SCATTER,  // life decay + map scatter
    PASS_STACKING,        // CheckStacking
    PASS_CONDUCT,         // conduction gate, ambient exchange, equilibrium
    PASS_CONDUCT_CELL,    // per-cell resolution of air_hv / block_airh writes
    PASS_CONDUCT_APPLY,   // last-writer temperature apply
    PASS_TRANSITION,      // phase transitions
    PASS_TRANSITION_CELL, // per-cell resolution of air_pv writes
    PASS_REFRESH_PMAP,    // pmap type refresh
    PASS_AIR_DEPOSIT,     // particle -> air drag
    PASS_AIR_DEPOSIT_CELL,
    PASS_MOVE_DETECT,     // WL_DETECT charging + wire closure
    PASS_MOVE_CLEAR,      // cell_winner clear + planner type map + cell records
    PASS_MOVE_INTENT,     // velocity integration + destination + claim
    PASS_MOVE_SAW,        // deferred WOOD -> SAWD + e == 2 pair resolution
    PASS_MOVE_COMMIT,     // winner check + move/swap
    PASS_MOVE_DRAIN,      // blocked-photon hits + creations queued by the pair pass
    PASS_MOVE_REBUILD,    // pmap/photons rebuild
    PASS_WATER_EQUAL,     // water equalisation (when enabled)
    PASS_REACTION_LIFE,   // reaction life pass (burn-out, spark settling, life kill)
    PASS_REACTION_COMBUSTION,  // fire heating and ignition, steam
    PASS_REACTION_FLUID,  // water family and ice
    PASS_COUNT
};

extern const char *const tpt_pass_names[PASS_COUNT];

struct tpt_fft_state {
    uint32_t width = 0, height = 0;   // padded transform size (powers of two)
    uint32_t xcells = 0, ycells = 0;
    float2 *work = nullptr;           // padded mass spectrum, width*height
    float2 *spec = nullptr;           // column-transformed product (x force + i y force)
    float2 *kernel = nullptr;         // cached kernel spectrum of kx + i ky
    float2 *twiddle = nullptr;        // exp(-2 pi i k / N), N = max(width,height)
    bool kernels_ready = false;
};

// Scratch of the water equalisation pass.
struct tpt_water_bufs {
    uint32_t *col_min = nullptr, *col_max = nullptr;  // xres: chunk accumulators, kept reset
    uint32_t *col_surf = nullptr;                      // xres
    unsigned long long *col_target = nullptr;          // xres: (row + 1) << 32 | (target pixel + 1)
    uint32_t *winner = nullptr;                        // pixels
};

struct tpt_ctx {
    tpt_config cfg;
    uint32_t xcells = 0, ycells = 0, cells = 0, pixels = 0;
    uint32_t slot_bound = 1;
    int device = 0;
    hipStream_t stream = nullptr;
    void *buf[TPT_BUF_COUNT] = {};
    size_t buf_bytes[TPT_BUF_COUNT] = {};
    bool tables_set = false;

    // Library-internal scratch (not part of the reference layout).
    uint32_t *cell_scratch_u = nullptr;   // cells words
    float *cell_scratch_f = nullptr;      // cells words
    uint32_t *cell_scratch_u2 = nullptr;  // cells words
    float *cell_scratch_f2 = nullptr;     // cells words
    uint32_t *slot_scratch_u = nullptr;   // capacity words
    uint32_t *slot_scratch_u2 = nullptr;  // capacity words
    float *slot_scratch_f = nullptr;      // capacity words
    uint32_t *counters = nullptr;         // small device counters
    // Pixel-major view of the maps built by RecalcFreeParticles: one 64-bit
    // record per pixel, ((slot+1)<<9 | flag<<8 | type) << 32 | temp bits,
    // the off-map slot list, and the conduction results per pixel
    // (equilibrium temperature bits, insulator/tmp flags).
    uint64_t *pix_rec = nullptr;          // pixels
    uint2 *pix_heat = nullptr;            // pixels
    uint8_t *pix_type = nullptr;          // pixels: pmap type byte for the movement planner
    uint32_t *offmap = nullptr;           // capacity words
    uint32_t *map_conflicts = nullptr;    // capacity words: the map scatter's shared-pixel slots
    // True while pix_rec / offmap describe the current particle state: set
    // by RecalcFreeParticles, cleared by anything that may change slots,
    // positions, temperatures or the maps afterwards.
    bool maps_fresh = false;
    // True while every live slot is either the owner of its rounded pixel's
    // record or on the off-map list, with positions unchanged since
    // RecalcFreeParticles (types may have changed since): set there,
    // cleared by movement, uploads and slot-bound changes.
    bool rec_cover = false;
    // True from the movement pass's map rebuild until anything other than
    // air or gravity runs: the next RecalcFreeParticles may then update the
    // rebuilt maps instead of rebuilding them.
    bool rebuild_fresh = false;
    // Wall facts the host learns from its uploads (the device never edits
    // wall_bmap): the canvas has conductor walls; an uploaded emap carries
    // span-end marker bits.
    bool has_wire = false;
    bool has_stasis = false;  // some wall cell may be WL_STASIS (18)
    bool emap_markers = false;
    // The gravity tool grid / override channel may hold non-zero cells:
    // set by their uploads, cleared once the gravity step consumed the tool
    // grid / the tick cleared the channel.  The device only ever zeroes
    // them, so the host knows when their kernels have work.
    bool tool_armed = false;
    bool override_armed = false;

    // Graph capture of the steady-state tick (tick.hip).  `capturing` makes
    // the passes record every kernel whose launch is otherwise decided on
    // the host per frame; `epoch` counts changes that invalidate a capture.
    bool capturing = false;
    uint64_t epoch = 0;
    uint64_t graph_epoch = ~0ull;
    uint32_t graph_slots = 0;
    hipGraphExec_t graph = nullptr;
    uint32_t *pix_count_mv = nullptr;     // pixels: particle counts of the movement rebuild
    uint32_t *slot_owner = nullptr;       // capacity words, all ones at rest (photon hit rounds)
    // e == 2 creations: TPT_CREATE_REC_WORDS words per record, one record
    // at most per pair mover (so `capacity` records), then the free bitmap
    // of the device allocation window (TPT_SPAWN_BITMAP_WORDS words).
    uint32_t *create_rec = nullptr;
    // Per-cell 16-byte records, so that a random cell access costs one line:
    // {grav_fx, grav_fy, wall_bmap, wall_emap} from the gravity solve to
    // grav_apply, {grav_fx, grav_fy, air_pv, 0} for convection and the
    // transitions, then {air_vx, air_vy, wall_bmap, wall_emap} for the
    // movement planner (rebuilt by each pass before use).
    uint4 *cell_rec = nullptr;
    // Element passes: per-cell flags of the element classes present this
    // tick, per-pixel pending writes (type word, temperature bits) between a
    // pass's decide and apply kernels (zero at rest), per-pixel records of
    // the burners (valid when their stamp is the step's), per-cell counts of
    // explosive ignitions (zero at rest).
    uint32_t *rx_flags = nullptr;         // cells
    uint2 *rx_side = nullptr;             // pixels
    uint2 *rx_burner = nullptr;           // pixels: {stamp | surround << 28, temperature bits} of burner owners
    uint32_t *rx_pvcount = nullptr;       // cells
    uint32_t *rx_list = nullptr;          // 64 per wave of a 64x4 tile: changed pixels of the current element pass
    uint32_t *rx_wcount = nullptr;        // per wave of a 64x4 tile: entries in its rx_list segment
    uint32_t rx_tiles = 0;                // 64x4-pixel tiles of the canvas
    tpt_water_bufs water;
    tpt_fft_state fft;

    // Profiling: a ring of event pairs per pass, folded lazily so that
    // timing never stalls the launch stream.
    static constexpr int RING = 32;
    bool profiling = false;
    hipEvent_t ev_start[PASS_COUNT][RING] = {};
    hipEvent_t ev_stop[PASS_COUNT][RING] = {};
    int ev_head[PASS_COUNT] = {};     // next ring slot to record
    int ev_pending[PASS_COUNT] = {};  // recorded, not yet folded
    double pass_ms[PASS_COUNT] = {};
    uint64_t pass_launches[PASS_COUNT] = {};
};

// Error plumbing ----------------------------------------------------------
int tpt_set_error(int code, const char *fmt, ...);
#define TPT_HIP_CHECK(expr)                                                       \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess)                                                     \
            return tpt_set_error(TPT_EDEVICE, "%s failed: %s (%s:%d)", #expr,     \
                                 hipGetErrorString(_e), __FILE__, __LINE__);      \
    } while (0)

// Pass timing: bracket launches with events when profiling is on.
void tpt_pass_begin(tpt_ctx *ctx, int pass);
void tpt_pass_end(tpt_ctx *ctx, int pass);
int tpt_check_launch(const char *what);

template <typename T> static inline T *tpt_buf(tpt_ctx *ctx, int id) {
    return reinterpret_cast<T *>(ctx->buf[id]);
}

// Pass drivers implemented per translation unit.
// The start of a tick (frame counter, seed, CheckStacking flag in the
// counter block), run by the first kernel of the tick's air pass.
struct tpt_tick_begin {
    uint32_t *counters;  // null: not a tick
    uint32_t lo, hi;     // frame_count, when not taken from the device
    uint32_t from_device;
};

__device__ __forceinline__ uint32_t tpt_frame_seed_dev(uint64_t f) {
    uint32_t v = (uint32_t)f * 0x9E3779B9u;
    v = (v << 13) | (v >> 19);
    return v ^ 0x85EBCA6Bu;
}

// The tick's seed and CheckStacking flag from frame `f` (the argument, or
// the device counter when `from_device`), and the counter advanced; one
// thread.
__device__ __forceinline__ void tpt_tick_begin_dev(const tpt_tick_begin &tb) {
    uint32_t *c = tb.counters;
    // Both frames are formed and one is selected by value: a selected
    // address (the counters or the by-value argument) would put the argument
    // in scratch memory for every thread of the calling kernel.
    const uint64_t on_device = ((uint64_t)c[TPT_CTR_FRAME_HI] << 32) | c[TPT_CTR_FRAME_LO];
    const uint64_t given = ((uint64_t)tb.hi << 32) | tb.lo;
    uint64_t f = tb.from_device ? on_device : given;
    c[TPT_CTR_SEED] = tpt_frame_seed_dev(f);
    c[TPT_CTR_STACK] = (f % 10u) == 0u ? 1u : 0u;
    f += 1u;
    c[TPT_CTR_FRAME_LO] = (uint32_t)f;
    c[TPT_CTR_FRAME_HI] = (uint32_t)(f >> 32);
}

// air_kernel::step_air_gpu; `tb` set when the pass opens a tick.
int tpt_air_step(tpt_ctx *ctx, const tpt_tick_begin *tb = nullptr);
int tpt_gravity_step(tpt_ctx *ctx, uint32_t seed);
int tpt_gravity_solve_only(tpt_ctx *ctx);
int tpt_gravity_free(tpt_ctx *ctx);
// Zeroes the element mass-override channel (value and flag grids) at the
// point of the tick where the reference's element pass clears it.
int tpt_gravity_override_clear(tpt_ctx *ctx);
// The reaction step as far as this library runs it: the override clear and
// the life pass.
int tpt_reactions_step(tpt_ctx *ctx);
int tpt_recalc_step(tpt_ctx *ctx, int stacking);
int tpt_thermal_step(tpt_ctx *ctx, int legacy);
int tpt_movement_step(tpt_ctx *ctx);
int tpt_move_rebuild_maps(tpt_ctx *ctx);
int tpt_water_equal_step(tpt_ctx *ctx);

// Device helpers ------------------------------------------------------------
__device__ __forceinline__ float tpt_clampf(float v, float lo, float hi) {
    float out = v;
    if (out < lo) out = lo;
    if (out > hi) out = hi;
    return out;
}

// Truncation toward zero with saturation; NaN -> 0.
__host__ __device__ __forceinline__ int tpt_f2i(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int)0x80000000u;
    return (int)f;
}

__host__ __device__ __forceinline__ uint32_t tpt_f2u(float f) {
    if (!(f == f) || f <= 0.0f) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}
