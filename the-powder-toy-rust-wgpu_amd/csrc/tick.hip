// Whole-tick driver, the per-pass C entry points, graph replay and the
// batched form.
//
// tpt_step follows SimulationEngine::step (crates/tpt-sim/src/lib.rs:303-459)
// for the passes this library implements: air -> Newtonian gravity (or the
// force-grid clear) -> RecalcFreeParticles (+ CheckStacking every 10th
// frame) -> conduction/transitions -> the element passes
// (reaction_kernel::step_reactions_gpu, lib.rs:379: the mass-override clear,
// the life, combustion and fluid passes for the elements this library
// covers) -> movement -> water equalisation (when enabled).
//
// The frame counter, the tick seed and the CheckStacking flag live in the
// context's device counter block, so that a captured tick replays with the
// right seed every frame: the tick's first kernel derives them from the
// frame (tpt_tick_begin_dev), and every seeded kernel reads the seed from
// there.

#include "tpt_internal.h"

#include <algorithm>
#include <new>
#include <utility>
#include <vector>

namespace {

// A per-pass call's seed (and CheckStacking flag).
__global__ void set_seed(uint32_t *__restrict__ counters, uint32_t seed, uint32_t stacking) {
    if (threadIdx.x != 0) return;
    counters[TPT_CTR_SEED] = seed;
    counters[TPT_CTR_STACK] = stacking;
}

int require(tpt_ctx *ctx) {
    if (!ctx) return tpt_set_error(TPT_EINVAL, "null context");
    if (!ctx->tables_set) return tpt_set_error(TPT_ESTATE, "element table not set");
    return TPT_OK;
}

int put_seed(tpt_ctx *ctx, uint32_t seed, int stacking) {
    hipLaunchKernelGGL(set_seed, dim3(1), dim3(64), 0, ctx->stream, ctx->counters, seed, stacking ? 1u : 0u);
    return tpt_check_launch("seed");
}

// One tick.  `from_device`: take the frame from the device counter (a
// captured tick); otherwise from `frame`.
int tick(tpt_ctx *ctx, uint64_t frame, bool from_device) {
    const tpt_tick_begin tb{ctx->counters, (uint32_t)frame, (uint32_t)(frame >> 32), from_device ? 1u : 0u};
    int rc;
    if ((rc = tpt_air_step(ctx, &tb))) return rc;
    if ((rc = tpt_gravity_step(ctx, 0u))) return rc;
    if ((rc = tpt_recalc_step(ctx, !from_device && (frame % 10u) == 0u))) return rc;
    if ((rc = tpt_thermal_step(ctx, ctx->cfg.heat_enable ? 0 : 1))) return rc;
    // The element passes (step_reactions_gpu, lib.rs:379).
    if ((rc = tpt_reactions_step(ctx))) return rc;
    if ((rc = tpt_movement_step(ctx))) return rc;
    return ctx->cfg.water_equal_test ? tpt_water_equal_step(ctx) : TPT_OK;
}

// Records the steady-state tick of `ctx` on its stream, which is already
// capturing.  Profiling events are not recorded into graphs.
int record_tick(tpt_ctx *ctx) {
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    ctx->capturing = true;
    const int rc = tick(ctx, 0, true);
    ctx->capturing = false;
    ctx->profiling = prof;
    return rc;
}

int capture(tpt_ctx *ctx) {
    if (ctx->graph) {
        (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
    TPT_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = record_tick(ctx);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return tpt_set_error(TPT_EDEVICE, "tick capture: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return tpt_set_error(TPT_EDEVICE, "graph instantiate: %s", hipGetErrorString(ei));
    ctx->graph_epoch = ctx->epoch;
    ctx->graph_slots = ctx->slot_bound;
    return TPT_OK;
}

}  // namespace

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
    if (!rc) rc = put_seed(ctx, seed, stacking);
    return rc ? rc : tpt_recalc_step(ctx, stacking);
}

extern "C" int tpt_step_thermal(tpt_ctx *ctx, uint32_t seed, int legacy) {
    int rc = require(ctx);
    if (!rc) rc = put_seed(ctx, seed, 0);
    return rc ? rc : tpt_thermal_step(ctx, legacy);
}

extern "C" int tpt_step_movement(tpt_ctx *ctx, uint32_t seed) {
    int rc = require(ctx);
    if (!rc) rc = put_seed(ctx, seed, 0);
    return rc ? rc : tpt_movement_step(ctx);
}

extern "C" int tpt_step_reactions(tpt_ctx *ctx, uint32_t seed) {
    int rc = require(ctx);
    if (!rc) rc = put_seed(ctx, seed, 0);
    return rc ? rc : tpt_reactions_step(ctx);
}

extern "C" int tpt_step_water_equal(tpt_ctx *ctx, uint32_t seed) {
    int rc = require(ctx);
    if (rc || !ctx->cfg.water_equal_test) return rc;
    if ((rc = put_seed(ctx, seed, 0))) return rc;
    return tpt_water_equal_step(ctx);
}

extern "C" int tpt_step(tpt_ctx *ctx, uint64_t frame_count) {
    int rc = require(ctx);
    return rc ? rc : tick(ctx, frame_count, false);
}

extern "C" int tpt_run(tpt_ctx *ctx, uint64_t first_frame, uint32_t ticks) {
    int rc = require(ctx);
    if (rc || ticks == 0) return rc;
    // The eager first tick brings the context to the steady state the
    // captured tick assumes (maps rebuilt by the previous movement pass).
    if ((rc = tick(ctx, first_frame, false))) return rc;
    if (ticks == 1) return TPT_OK;
    if (!ctx->graph || ctx->graph_epoch != ctx->epoch || ctx->graph_slots != ctx->slot_bound)
        if ((rc = capture(ctx))) return rc;
    for (uint32_t t = 1; t < ticks; ++t) TPT_HIP_CHECK(hipGraphLaunch(ctx->graph, ctx->stream));
    return TPT_OK;
}

// ---------------------------------------------------------------------------
// Batched form.

struct tpt_batch {
    std::vector<tpt_ctx *> ctx;
    hipStream_t main = nullptr;
    hipEvent_t fork = nullptr, done = nullptr;  // fork: recorded inside the capture only
    std::vector<hipEvent_t> join, pre;           // join: inside the capture; pre: outside
    hipGraphExec_t graph = nullptr;
    std::vector<uint64_t> epochs;
    std::vector<uint32_t> slots;
    // 0: every canvas replays its own captured tick on its own stream;
    // otherwise one graph holds every canvas's tick, the canvases dealt to
    // `lanes` serial chains that run side by side.
    uint32_t lanes = 0;
    std::vector<std::pair<uint32_t, float>> tuning;  // autotuner: (lanes, ms per batch tick)
};

static void batch_free(tpt_batch *b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->pre)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->main) (void)hipStreamDestroy(b->main);
    for (tpt_ctx *c : b->ctx)
        if (c) tpt_destroy(c);
    delete b;
}

extern "C" int tpt_batch_create(uint32_t n, const tpt_config *cfgs, tpt_batch **out) {
    if (!out || !cfgs || n == 0) return tpt_set_error(TPT_EINVAL, "batch needs n >= 1 configs and an output");
    *out = nullptr;
    tpt_batch *b = new (std::nothrow) tpt_batch();
    if (!b) return tpt_set_error(TPT_ENOMEM, "host allocation failed");
    b->ctx.assign(n, nullptr);
    b->join.assign(n, nullptr);
    b->pre.assign(n, nullptr);
    int rc = TPT_OK;
    for (uint32_t k = 0; k < n && rc == TPT_OK; ++k) rc = tpt_create(&cfgs[k], &b->ctx[k]);
    if (rc == TPT_OK && (hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking) != hipSuccess ||
                         hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
                         hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess))
        rc = tpt_set_error(TPT_EDEVICE, "batch stream/event creation failed");
    for (uint32_t k = 0; k < n && rc == TPT_OK; ++k)
        if (hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b->pre[k], hipEventDisableTiming) != hipSuccess)
            rc = tpt_set_error(TPT_EDEVICE, "batch event creation failed");
    if (rc != TPT_OK) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return TPT_OK;
}

extern "C" int tpt_batch_destroy(tpt_batch *b) {
    if (!b) return tpt_set_error(TPT_EINVAL, "null batch");
    (void)tpt_batch_synchronize(b);
    batch_free(b);
    return TPT_OK;
}

extern "C" int tpt_batch_context(tpt_batch *b, uint32_t k, tpt_ctx **ctx) {
    if (!b || !ctx || k >= b->ctx.size()) return tpt_set_error(TPT_EINVAL, "batch context %u out of range", k);
    *ctx = b->ctx[k];
    return TPT_OK;
}

// Makes `waiter` wait for everything queued on `on` so far.
static int stream_after(hipStream_t waiter, hipStream_t on, hipEvent_t ev) {
    TPT_HIP_CHECK(hipEventRecord(ev, on));
    TPT_HIP_CHECK(hipStreamWaitEvent(waiter, ev, 0));
    return TPT_OK;
}

// The canvases of each lane, largest first, dealt greedily to the lane
// with the least work so far (work = pixels + particles' slot range).
static std::vector<std::vector<size_t>> batch_lanes(const tpt_batch *b) {
    const size_t n = b->ctx.size();
    const size_t L = b->lanes >= n ? n : b->lanes;
    std::vector<size_t> order(n);
    for (size_t k = 0; k < n; ++k) order[k] = k;
    auto work = [b](size_t k) { return (uint64_t)b->ctx[k]->pixels + b->ctx[k]->slot_bound; };
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work(x) > work(y); });
    std::vector<std::vector<size_t>> lanes(L);
    std::vector<uint64_t> load(L, 0);
    for (size_t k : order) {
        const size_t l = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        lanes[l].push_back(k);
        load[l] += work(k);
    }
    return lanes;
}

static int batch_capture(tpt_batch *b) {
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    const std::vector<std::vector<size_t>> lanes = batch_lanes(b);
    TPT_HIP_CHECK(hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
    int rc = TPT_OK;
    hipError_t e = hipEventRecord(b->fork, b->main);
    for (size_t l = 0; l < lanes.size() && rc == TPT_OK && e == hipSuccess; ++l) {
        hipEvent_t after = b->fork;  // each lane's first canvas starts at the fork
        for (size_t k : lanes[l]) {
            if (rc != TPT_OK || e != hipSuccess) break;
            tpt_ctx *c = b->ctx[k];
            e = hipStreamWaitEvent(c->stream, after, 0);
            if (e == hipSuccess) rc = record_tick(c);
            if (rc == TPT_OK && e == hipSuccess) e = hipEventRecord(b->join[k], c->stream);
            after = b->join[k];
        }
        if (rc == TPT_OK && e == hipSuccess) e = hipStreamWaitEvent(b->main, after, 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->main, &g);
    if (rc != TPT_OK || e != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc) return rc;
        return tpt_set_error(TPT_EDEVICE, "batch capture: %s", hipGetErrorString(e != hipSuccess ? e : ee));
    }
    const hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return tpt_set_error(TPT_EDEVICE, "batch graph instantiate: %s", hipGetErrorString(ei));
    b->epochs.resize(b->ctx.size());
    b->slots.resize(b->ctx.size());
    for (size_t k = 0; k < b->ctx.size(); ++k) {
        b->epochs[k] = b->ctx[k]->epoch;
        b->slots[k] = b->ctx[k]->slot_bound;
    }
    return TPT_OK;
}

extern "C" int tpt_batch_step(tpt_batch *b, uint64_t first_frame, uint32_t ticks) {
    if (!b) return tpt_set_error(TPT_EINVAL, "null batch");
    if (ticks == 0) return TPT_OK;
    int rc = TPT_OK;
    if (b->lanes == 0) {
        // Every canvas's tick captured once as its own graph (tpt_run) and
        // replayed on the canvas's stream; the canvases' streams run side by
        // side.  Measured faster than one graph holding every canvas as a
        // branch, whose joins order the canvases' work more tightly.
        for (tpt_ctx *c : b->ctx)
            if ((rc = tpt_run(c, first_frame, ticks))) return rc;
        return TPT_OK;
    }
    for (tpt_ctx *c : b->ctx)
        if ((rc = require(c)) || (rc = tick(c, first_frame, false))) return rc;
    if (ticks == 1) return TPT_OK;
    bool stale = !b->graph || b->epochs.size() != b->ctx.size();
    for (size_t k = 0; k < b->ctx.size() && !stale; ++k)
        stale = b->epochs[k] != b->ctx[k]->epoch || b->slots[k] != b->ctx[k]->slot_bound;
    // The replays on the batch stream follow each canvas's eager tick ...
    for (size_t k = 0; k < b->ctx.size(); ++k)
        if ((rc = stream_after(b->main, b->ctx[k]->stream, b->pre[k]))) return rc;
    if (stale && (rc = batch_capture(b))) return rc;
    for (uint32_t t = 1; t < ticks; ++t) TPT_HIP_CHECK(hipGraphLaunch(b->graph, b->main));
    // ... and later work on a canvas's own stream follows the replays.
    for (tpt_ctx *c : b->ctx)
        if ((rc = stream_after(c->stream, b->main, b->done))) return rc;
    return TPT_OK;
}

extern "C" int tpt_batch_step_launches(tpt_batch *b, uint64_t first_frame, uint32_t ticks) {
    if (!b) return tpt_set_error(TPT_EINVAL, "null batch");
    for (uint32_t t = 0; t < ticks; ++t)
        for (tpt_ctx *c : b->ctx) {
            int rc = require(c);
            if (rc || (rc = tick(c, first_frame + t, false))) return rc;
        }
    return TPT_OK;
}

// One canvas at a time, the way the reference steps its engines: canvas k's
// tick waits on the GPU for canvas k-1's (no host synchronisation).
extern "C" int tpt_batch_step_serial(tpt_batch *b, uint64_t first_frame, uint32_t ticks) {
    if (!b) return tpt_set_error(TPT_EINVAL, "null batch");
    const size_t n = b->ctx.size();
    for (uint32_t t = 0; t < ticks; ++t)
        for (size_t k = 0; k < n; ++k) {
            tpt_ctx *c = b->ctx[k];
            int rc = require(c);
            const size_t prev = k ? k - 1 : n - 1;
            if (!rc && (t || k) && n > 1) rc = stream_after(c->stream, b->ctx[prev]->stream, b->pre[prev]);
            if (rc || (rc = tick(c, first_frame + t, false))) return rc;
        }
    return TPT_OK;
}

// Times the captured tick at several work splits on live replays and keeps
// the fastest.  Candidates: 1, 2, 4, ... lanes up to the canvas count, and
// one lane per canvas.  Advances every canvas by 1 + sum(trial_ticks + 1)
// ticks; results equal tpt_batch_step's for the same number of ticks.
// Switched off: one run of this entry point on an MI355X ended in a silent
// abort inside the HIP runtime (no GPU fault reported) while the same test
// had passed before; it is the only path that captures and destroys several
// graphs on one batch, reusing the batch's events, and no cause has been
// shown.  It returns TPT_ESTATE until that is understood.
constexpr bool kBatchAutotune = false;

extern "C" int tpt_batch_autotune(tpt_batch *b, uint64_t first_frame, uint32_t trial_ticks, uint32_t *lanes_out) {
    if (!kBatchAutotune)
        return tpt_set_error(TPT_ESTATE, "tpt_batch_autotune is switched off (an unexplained abort on the GPU); "
                                         "use tpt_batch_set_lanes");
    if (!b || trial_ticks == 0) return tpt_set_error(TPT_EINVAL, "batch autotune needs a batch and trial ticks");
    int rc = TPT_OK;
    for (tpt_ctx *c : b->ctx)
        if ((rc = require(c)) || (rc = tick(c, first_frame, false))) return rc;
    for (size_t k = 0; k < b->ctx.size(); ++k)
        if ((rc = stream_after(b->main, b->ctx[k]->stream, b->pre[k]))) return rc;
    const uint32_t n = (uint32_t)b->ctx.size();
    std::vector<uint32_t> cands;
    for (uint32_t L = 1; L < n; L *= 2) cands.push_back(L);
    cands.push_back(0);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    TPT_HIP_CHECK(hipEventCreate(&t0));
    TPT_HIP_CHECK(hipEventCreate(&t1));
    b->tuning.clear();
    uint32_t best = 0;
    float best_ms = 0.0f;
    for (uint32_t L : cands) {
        b->lanes = L;
        if ((rc = batch_capture(b))) break;
        hipError_t e = hipGraphLaunch(b->graph, b->main);  // warm
        if (e == hipSuccess) e = hipEventRecord(t0, b->main);
        for (uint32_t t = 0; t < trial_ticks && e == hipSuccess; ++t) e = hipGraphLaunch(b->graph, b->main);
        if (e == hipSuccess) e = hipEventRecord(t1, b->main);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        if (e != hipSuccess) {
            rc = tpt_set_error(TPT_EDEVICE, "batch autotune: %s", hipGetErrorString(e));
            break;
        }
        ms /= (float)trial_ticks;
        b->tuning.push_back({L, ms});
        if (b->tuning.size() == 1 || ms < best_ms) {
            best = L;
            best_ms = ms;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return rc;
    b->lanes = best;
    if ((rc = batch_capture(b))) return rc;
    for (tpt_ctx *c : b->ctx)
        if ((rc = stream_after(c->stream, b->main, b->done))) return rc;
    if (lanes_out) *lanes_out = best;
    return TPT_OK;
}

extern "C" int tpt_batch_tuning(tpt_batch *b, uint32_t *lanes, float *ms_per_tick, uint32_t max_entries,
                                uint32_t *n_entries) {
    if (!b || !n_entries) return tpt_set_error(TPT_EINVAL, "null argument");
    uint32_t n = 0;
    for (const auto &t : b->tuning) {
        if (n < max_entries) {
            if (lanes) lanes[n] = t.first;
            if (ms_per_tick) ms_per_tick[n] = t.second;
        }
        ++n;
    }
    *n_entries = n;
    return TPT_OK;
}

extern "C" int tpt_batch_set_lanes(tpt_batch *b, uint32_t lanes) {
    if (!b) return tpt_set_error(TPT_EINVAL, "null batch");
    if (lanes != b->lanes) {
        b->lanes = lanes;
        b->epochs.clear();  // recapture on the next tpt_batch_step
    }
    return TPT_OK;
}

extern "C" int tpt_batch_synchronize(tpt_batch *b) {
    if (!b) return tpt_set_error(TPT_EINVAL, "null batch");
    TPT_HIP_CHECK(hipStreamSynchronize(b->main));
    for (tpt_ctx *c : b->ctx) TPT_HIP_CHECK(hipStreamSynchronize(c->stream));
    return TPT_OK;
}
