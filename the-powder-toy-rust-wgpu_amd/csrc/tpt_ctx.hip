// Context, buffers, tables, pass timing and the whole-tick driver of
// libtpt_hip (C ABI in include/tpt_hip.h).
//
// The buffer set mirrors SimBuffers (crates/tpt-sim/src/buffers/mod.rs:81-393,
// allocated by layout.rs:42-257): one device allocation per logical array so
// a caller can upload/download exactly the reference's arrays.

#include "tpt_internal.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>

const char *const tpt_pass_names[PASS_COUNT] = {
    "air_pressure_velocity", "air_smooth_advect", "air_heat_convect", "air_settle",
    "grav_mass_seed", "grav_fft_rows", "grav_fft_cols", "grav_fft_inverse_extract",
    "grav_apply", "grav_build_mass", "recalc_clear", "recalc_scatter", "check_stacking",
    "thermal_conduction", "thermal_conduction_cells", "thermal_apply", "phase_transition",
    "phase_transition_cells", "refresh_pmap_types", "air_deposit", "air_deposit_cells", "move_detect",
    "move_prepare", "move_velocity_intent", "move_saw_pair", "move_commit", "move_photon_hits_drain",
    "move_rebuild_maps", "water_equal", "reaction_life", "reaction_combustion", "reaction_fluid",
};

static thread_local char g_err[512];

extern "C" const char *tpt_last_error(void) { return g_err; }

int tpt_set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int tpt_check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return tpt_set_error(TPT_EDEVICE, "launch of %s failed: %s", what, hipGetErrorString(e));
    return TPT_OK;
}

// Folds the oldest pending measurement of `pass` into its total.
static void fold_oldest(tpt_ctx *ctx, int pass) {
    const int R = tpt_ctx::RING;
    int slot = (ctx->ev_head[pass] - ctx->ev_pending[pass] + R) % R;
    float ms = 0.0f;
    // A failed wait or query drops the measurement; timing never fails a pass.
    if (hipEventSynchronize(ctx->ev_stop[pass][slot]) == hipSuccess &&
        hipEventElapsedTime(&ms, ctx->ev_start[pass][slot], ctx->ev_stop[pass][slot]) == hipSuccess)
        ctx->pass_ms[pass] += ms;
    ctx->ev_pending[pass] -= 1;
}

void tpt_pass_begin(tpt_ctx *ctx, int pass) {
    if (!ctx->profiling || ctx->capturing) return;
    if (ctx->ev_pending[pass] == tpt_ctx::RING) fold_oldest(ctx, pass);
    (void)hipEventRecord(ctx->ev_start[pass][ctx->ev_head[pass]], ctx->stream);
}

void tpt_pass_end(tpt_ctx *ctx, int pass) {
    if (!ctx->profiling || ctx->capturing) return;
    int slot = ctx->ev_head[pass];
    (void)hipEventRecord(ctx->ev_stop[pass][slot], ctx->stream);
    ctx->ev_head[pass] = (slot + 1) % tpt_ctx::RING;
    ctx->ev_pending[pass] += 1;
    ctx->pass_launches[pass] += 1;
}

static size_t slot_words(const tpt_ctx *c) { return c->cfg.capacity; }

static int alloc_buffers(tpt_ctx *c) {
    const size_t cap = slot_words(c);
    for (int id = 0; id < TPT_BUF_COUNT; ++id) {
        size_t words = 0;
        if (id <= TPT_BUF_INTENT_TARGET) words = cap;
        else if (id <= TPT_BUF_CELL_WINNER) words = c->pixels;
        else if (id <= TPT_BUF_GRAV_MASS_OVERRIDE_SET) words = c->cells;
        else if (id == TPT_BUF_ELEM_PROPS) words = (size_t)TPT_PT_NUM * TPT_ELEM_STRIDE;
        else if (id == TPT_BUF_MOVE_LUT) words = TPT_MOVE_LUT_WORDS;
        else if (id == TPT_BUF_PHOTON_MASK) words = TPT_PT_NUM;
        else if (id == TPT_BUF_SPAWN_DEFAULTS) words = (size_t)TPT_PT_NUM * TPT_SPAWN_DEFAULT_WORDS;
        c->buf_bytes[id] = words * 4;
        if (hipMalloc(&c->buf[id], words * 4) != hipSuccess)
            return tpt_set_error(TPT_ENOMEM, "hipMalloc of buffer %d (%zu bytes) failed", id, words * 4);
        if (hipMemsetAsync(c->buf[id], 0, words * 4, c->stream) != hipSuccess)
            return tpt_set_error(TPT_EDEVICE, "hipMemsetAsync of buffer %d failed", id);
    }
    // Every slot starts as "no intent" (move_kernel/mod.rs:124).
    if (hipMemsetAsync(c->buf[TPT_BUF_INTENT_TARGET], 0xFF, cap * 4, c->stream) != hipSuccess)
        return tpt_set_error(TPT_EDEVICE, "intent clear failed");
    size_t cb = (size_t)c->cells * 4;
    c->rx_tiles = ((c->cfg.xres + 63u) / 64u) * (c->cfg.yres / 4u);  // the element passes' 64x4 tiles
    if (hipMalloc(&c->cell_scratch_u, cb) != hipSuccess || hipMalloc(&c->cell_scratch_f, cb) != hipSuccess ||
        hipMalloc(&c->cell_scratch_u2, cb) != hipSuccess || hipMalloc(&c->cell_scratch_f2, cb) != hipSuccess ||
        hipMalloc(&c->slot_scratch_u, cap * 4) != hipSuccess || hipMalloc(&c->slot_scratch_u2, cap * 4) != hipSuccess ||
        hipMalloc(&c->slot_scratch_f, cap * 4) != hipSuccess || hipMalloc(&c->counters, 256) != hipSuccess ||
        hipMalloc(&c->pix_rec, (size_t)c->pixels * 8) != hipSuccess ||
        hipMalloc(&c->pix_heat, (size_t)c->pixels * 8) != hipSuccess || hipMalloc(&c->offmap, cap * 4) != hipSuccess ||
        hipMalloc(&c->map_conflicts, cap * 4) != hipSuccess ||
        hipMalloc(&c->pix_type, (size_t)c->pixels) != hipSuccess ||
        hipMalloc(&c->pix_count_mv, (size_t)c->pixels * 4) != hipSuccess ||
        hipMalloc(&c->slot_owner, cap * 4) != hipSuccess ||
        hipMalloc(&c->create_rec, (cap * TPT_CREATE_REC_WORDS + TPT_SPAWN_BITMAP_WORDS) * 4) != hipSuccess ||
        hipMalloc(&c->cell_rec, (size_t)c->cells * sizeof(uint4)) != hipSuccess ||
        hipMalloc(&c->rx_flags, (size_t)c->cells * 4) != hipSuccess ||
        hipMalloc(&c->rx_side, (size_t)c->pixels * sizeof(uint2)) != hipSuccess ||
        hipMalloc(&c->rx_burner, (size_t)c->pixels * sizeof(uint2)) != hipSuccess ||
        hipMalloc(&c->rx_pvcount, (size_t)c->cells * 4) != hipSuccess ||
        hipMalloc(&c->rx_list, (size_t)c->rx_tiles * 256 * 4) != hipSuccess ||
        hipMalloc(&c->rx_wcount, (size_t)c->rx_tiles * 4 * 4) != hipSuccess ||
        hipMalloc(&c->water.col_min, (size_t)c->cfg.xres * 4) != hipSuccess ||
        hipMalloc(&c->water.col_max, (size_t)c->cfg.xres * 4) != hipSuccess ||
        hipMalloc(&c->water.col_surf, (size_t)c->cfg.xres * 4) != hipSuccess ||
        hipMalloc(&c->water.col_target, (size_t)c->cfg.xres * 8) != hipSuccess ||
        hipMalloc(&c->water.winner, (size_t)c->pixels * 4) != hipSuccess)
        return tpt_set_error(TPT_ENOMEM, "scratch allocation failed");
    if (hipMemsetAsync(c->water.col_min, 0xFF, (size_t)c->cfg.xres * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(c->water.col_max, 0, (size_t)c->cfg.xres * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(c->cell_scratch_u, 0, cb, c->stream) != hipSuccess ||
        hipMemsetAsync(c->cell_scratch_u2, 0, cb, c->stream) != hipSuccess ||
        hipMemsetAsync(c->counters, 0, 256, c->stream) != hipSuccess ||
        hipMemsetAsync(c->rx_flags, 0, (size_t)c->cells * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(c->rx_side, 0, (size_t)c->pixels * sizeof(uint2), c->stream) != hipSuccess ||
        hipMemsetAsync(c->rx_burner, 0, (size_t)c->pixels * sizeof(uint2), c->stream) != hipSuccess ||
        hipMemsetAsync(c->rx_pvcount, 0, (size_t)c->cells * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(c->rx_wcount, 0, (size_t)c->rx_tiles * 4 * 4, c->stream) != hipSuccess ||
        hipMemsetAsync(c->slot_owner, 0xFF, cap * 4, c->stream) != hipSuccess)
        return tpt_set_error(TPT_EDEVICE, "scratch clear failed");
    return TPT_OK;
}

static void free_all(tpt_ctx *c) {
    for (int id = 0; id < TPT_BUF_COUNT; ++id)
        if (c->buf[id]) (void)hipFree(c->buf[id]);
    void *extra[] = {c->cell_scratch_u, c->cell_scratch_f, c->cell_scratch_u2, c->cell_scratch_f2,
                     c->slot_scratch_u, c->slot_scratch_u2, c->slot_scratch_f, c->counters,
                     c->pix_rec, c->pix_heat, c->offmap, c->map_conflicts, c->pix_type, c->pix_count_mv, c->slot_owner, c->create_rec, c->cell_rec,
                     c->rx_flags, c->rx_side, c->rx_burner, c->rx_pvcount, c->rx_list, c->rx_wcount, c->water.col_min, c->water.col_max, c->water.col_surf, c->water.col_target,
                     c->water.winner};
    for (void *p : extra)
        if (p) (void)hipFree(p);
    tpt_gravity_free(c);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    for (int p = 0; p < PASS_COUNT; ++p)
        for (int r = 0; r < tpt_ctx::RING; ++r) {
            if (c->ev_start[p][r]) (void)hipEventDestroy(c->ev_start[p][r]);
            if (c->ev_stop[p][r]) (void)hipEventDestroy(c->ev_stop[p][r]);
        }
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

extern "C" int tpt_create(const tpt_config *cfg, tpt_ctx **out) {
    g_err[0] = 0;
    if (!cfg || !out) return tpt_set_error(TPT_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->xres < TPT_CELL || cfg->yres < TPT_CELL || cfg->xres % TPT_CELL || cfg->yres % TPT_CELL ||
        cfg->xres > TPT_MAX_DIM || cfg->yres > TPT_MAX_DIM)
        return tpt_set_error(TPT_EINVAL, "canvas %ux%u must be a multiple of 4 in [4, %u]", cfg->xres,
                             cfg->yres, TPT_MAX_DIM);
    if (cfg->capacity == 0 || cfg->capacity > TPT_MAX_CAPACITY)
        return tpt_set_error(TPT_EINVAL, "capacity %u outside [1, 0x7FFFFF]", cfg->capacity);
    if (cfg->edge_mode > 2 || cfg->gravity_mode > 3 || cfg->air_mode > 4 || cfg->convection_mode > 2 ||
        cfg->grav_mass_mode > 1 || cfg->water_equal_test > 1)
        return tpt_set_error(TPT_EINVAL, "enum field out of range");
    tpt_ctx *c = new (std::nothrow) tpt_ctx();
    if (!c) return tpt_set_error(TPT_ENOMEM, "host allocation failed");
    c->cfg = *cfg;
    c->xcells = cfg->xres / TPT_CELL;
    c->ycells = cfg->yres / TPT_CELL;
    c->cells = c->xcells * c->ycells;
    c->pixels = cfg->xres * cfg->yres;
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return tpt_set_error(TPT_EDEVICE, "no HIP device or stream creation failed");
    }
    int rc = alloc_buffers(c);
    if (rc == TPT_OK && hipStreamSynchronize(c->stream) != hipSuccess)
        rc = tpt_set_error(TPT_EDEVICE, "initial clear failed");
    if (rc != TPT_OK) {
        free_all(c);
        delete c;
        return rc;
    }
    *out = c;
    return TPT_OK;
}

extern "C" int tpt_destroy(tpt_ctx *ctx) {
    if (!ctx) return tpt_set_error(TPT_EINVAL, "null context");
    // The context is released either way; a failed wait is still reported.
    const bool synced = hipStreamSynchronize(ctx->stream) == hipSuccess;
    free_all(ctx);
    delete ctx;
    return synced ? TPT_OK : tpt_set_error(TPT_EDEVICE, "stream synchronize before destroy failed");
}

extern "C" int tpt_set_element_table(tpt_ctx *ctx, const uint32_t *props, size_t props_words,
                                     const uint32_t *move_lut, size_t lut_words) {
    if (!ctx || !props || !move_lut) return tpt_set_error(TPT_EINVAL, "null argument");
    if (props_words != (size_t)TPT_PT_NUM * TPT_ELEM_STRIDE || lut_words != (size_t)TPT_MOVE_LUT_WORDS)
        return tpt_set_error(TPT_EINVAL, "table sizes %zu/%zu, expected %u/%u", props_words, lut_words,
                             TPT_PT_NUM * TPT_ELEM_STRIDE, TPT_MOVE_LUT_WORDS);
    ctx->maps_fresh = ctx->rec_cover = ctx->rebuild_fresh = false;
    ctx->epoch++;
    TPT_HIP_CHECK(hipMemcpyAsync(ctx->buf[TPT_BUF_ELEM_PROPS], props, props_words * 4, hipMemcpyHostToDevice, ctx->stream));
    TPT_HIP_CHECK(hipMemcpyAsync(ctx->buf[TPT_BUF_MOVE_LUT], move_lut, lut_words * 4, hipMemcpyHostToDevice, ctx->stream));
    TPT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->tables_set = true;
    return TPT_OK;
}

static int check_range(tpt_ctx *ctx, int buf, size_t off, size_t bytes) {
    if (!ctx) return tpt_set_error(TPT_EINVAL, "null context");
    if (buf < 0 || buf >= TPT_BUF_COUNT) return tpt_set_error(TPT_EINVAL, "buffer id %d out of range", buf);
    if (off % 4 || bytes % 4) return tpt_set_error(TPT_EINVAL, "offset/length must be multiples of 4");
    if (off > ctx->buf_bytes[buf] || bytes > ctx->buf_bytes[buf] - off)
        return tpt_set_error(TPT_EINVAL, "range [%zu, %zu) outside buffer %d of %zu bytes", off, off + bytes, buf,
                             ctx->buf_bytes[buf]);
    return TPT_OK;
}

extern "C" int tpt_upload(tpt_ctx *ctx, int buf, const void *src, size_t off, size_t bytes) {
    int rc = check_range(ctx, buf, off, bytes);
    if (rc) return rc;
    if (!src && bytes) return tpt_set_error(TPT_EINVAL, "null source");
    if (!bytes) return TPT_OK;
    // The gravity tool and override grids touch no particle or map state;
    // uploading them arms the one-shot kernels that consume them.
    if (buf == TPT_BUF_GRAV_TOOL_MASS)
        ctx->tool_armed = true;
    else if (buf == TPT_BUF_GRAV_MASS_OVERRIDE || buf == TPT_BUF_GRAV_MASS_OVERRIDE_SET)
        ctx->override_armed = true;
    else
        ctx->maps_fresh = ctx->rec_cover = ctx->rebuild_fresh = false;
    ctx->epoch++;
    if (buf == TPT_BUF_WALL_BMAP || buf == TPT_BUF_WALL_EMAP) {
        const uint32_t *w = (const uint32_t *)src;
        bool found = false, stasis = false;
        for (size_t k = 0; k < bytes / 4; ++k) {
            const uint32_t v = w[k];
            found = found || (buf == TPT_BUF_WALL_BMAP
                                  ? (v == 1u || v == 2u || v == 3u || v == 5u || v == 6u || v == 12u || v == 18u)
                                  : (v & 0xC0u) != 0u);
            stasis = stasis || (buf == TPT_BUF_WALL_BMAP && v == 18u);
        }
        const bool whole = off == 0 && bytes == ctx->buf_bytes[buf];
        bool &flag = buf == TPT_BUF_WALL_BMAP ? ctx->has_wire : ctx->emap_markers;
        flag = whole ? found : (flag || found);
        if (buf == TPT_BUF_WALL_BMAP) ctx->has_stasis = whole ? stasis : (ctx->has_stasis || stasis);
    }
    TPT_HIP_CHECK(hipMemcpyAsync((char *)ctx->buf[buf] + off, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    TPT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (buf == TPT_BUF_ELEM_PROPS || buf == TPT_BUF_MOVE_LUT) ctx->tables_set = true;
    return TPT_OK;
}

extern "C" int tpt_download(tpt_ctx *ctx, int buf, void *dst, size_t off, size_t bytes) {
    int rc = check_range(ctx, buf, off, bytes);
    if (rc) return rc;
    if (!dst && bytes) return tpt_set_error(TPT_EINVAL, "null destination");
    if (!bytes) return TPT_OK;
    TPT_HIP_CHECK(hipMemcpyAsync(dst, (char *)ctx->buf[buf] + off, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TPT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return TPT_OK;
}

extern "C" int tpt_buffer_bytes(tpt_ctx *ctx, int buf, size_t *bytes) {
    if (!ctx || !bytes) return tpt_set_error(TPT_EINVAL, "null argument");
    if (buf < 0 || buf >= TPT_BUF_COUNT) return tpt_set_error(TPT_EINVAL, "buffer id %d out of range", buf);
    *bytes = ctx->buf_bytes[buf];
    return TPT_OK;
}

extern "C" int tpt_set_slot_bound(tpt_ctx *ctx, uint32_t bound) {
    if (!ctx) return tpt_set_error(TPT_EINVAL, "null context");
    if (bound < 1) bound = 1;
    if (bound > ctx->cfg.capacity) bound = ctx->cfg.capacity;
    if (bound != ctx->slot_bound) {
        ctx->maps_fresh = ctx->rec_cover = ctx->rebuild_fresh = false;
        ctx->epoch++;
    }
    ctx->slot_bound = bound;
    return TPT_OK;
}

extern "C" int tpt_get_slot_bound(tpt_ctx *ctx, uint32_t *bound) {
    if (!ctx || !bound) return tpt_set_error(TPT_EINVAL, "null argument");
    *bound = ctx->slot_bound;
    return TPT_OK;
}

extern "C" int tpt_synchronize(tpt_ctx *ctx) {
    if (!ctx) return tpt_set_error(TPT_EINVAL, "null context");
    TPT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return TPT_OK;
}

extern "C" uint32_t tpt_frame_seed(uint64_t frame_count) {
    uint32_t v = (uint32_t)frame_count * 0x9E3779B9u;
    v = (v << 13) | (v >> 19);
    return v ^ 0x85EBCA6Bu;
}

extern "C" int tpt_step_air(tpt_ctx *ctx) {
    if (!ctx) return tpt_set_error(TPT_EINVAL, "null context");
    if (!ctx->tables_set) return tpt_set_error(TPT_ESTATE, "element table not set");
    return tpt_air_step(ctx);
}

extern "C" int tpt_set_profiling(tpt_ctx *ctx, int enable) {
    if (!ctx) return tpt_set_error(TPT_EINVAL, "null context");
    // Events are created on first use: a batch of many contexts needs none.
    if (enable && !ctx->ev_start[0][0]) {
        for (int p = 0; p < PASS_COUNT; ++p)
            for (int r = 0; r < tpt_ctx::RING; ++r)
                if (hipEventCreate(&ctx->ev_start[p][r]) != hipSuccess ||
                    hipEventCreate(&ctx->ev_stop[p][r]) != hipSuccess)
                    return tpt_set_error(TPT_EDEVICE, "event creation failed");
    }
    ctx->profiling = enable != 0;
    return TPT_OK;
}

extern "C" int tpt_reset_pass_times(tpt_ctx *ctx) {
    if (!ctx) return tpt_set_error(TPT_EINVAL, "null context");
    TPT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < PASS_COUNT; ++p) {
        ctx->pass_ms[p] = 0.0;
        ctx->pass_launches[p] = 0;
        ctx->ev_pending[p] = 0;
    }
    return TPT_OK;
}

extern "C" int tpt_get_pass_times(tpt_ctx *ctx, const char **names, double *total_ms, uint64_t *launches,
                                  int max_entries, int *n_entries) {
    if (!ctx || !n_entries) return tpt_set_error(TPT_EINVAL, "null argument");
    TPT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < PASS_COUNT; ++p)
        while (ctx->ev_pending[p] > 0) fold_oldest(ctx, p);
    int n = PASS_COUNT < max_entries ? PASS_COUNT : max_entries;
    for (int p = 0; p < n; ++p) {
        if (names) names[p] = tpt_pass_names[p];
        if (total_ms) total_ms[p] = ctx->pass_ms[p];
        if (launches) launches[p] = ctx->pass_launches[p];
    }
    *n_entries = n;
    return TPT_OK;
}
