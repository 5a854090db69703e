// Water equalisation for gfx950: water_equal_gpu::step_water_equal_gpu
// (crates/tpt-sim/src/water_equal_gpu/mod.rs:67-166), the optional
// communicating-vessel pass after movement (lib.rs:418-420).

#include "move_common.h"

namespace {

using namespace tpt_move;

constexpr int BLK = 256;

// ---- water equalisation (water_equal_gpu/, lib.rs:418-420) --------------
//
// The reference scans each column and each row with one thread; here the
// column pass folds 64-row chunks into per-column atomics, and each row is
// one workgroup that finds its liquid spans with a block-wide max-scan and
// reduces them in LDS.  Canonical rule WEQ: a column target that several
// rows assign keeps the highest row's (64-bit atomicMax on (row, target)).

constexpr int WEQ_CHUNK = 64;            // rows per column-pass block (16 per wave)
constexpr int WEQ_MAX_X = (int)TPT_MAX_DIM;
constexpr int WEQ_BLK = 256;
constexpr int WEQ_PER = (WEQ_MAX_X + WEQ_BLK - 1) / WEQ_BLK;  // columns per thread in a row pass

// Whether a pmap word holds a liquid (element state 2, props word 3), off a
// 256-bit table of liquid types that every wave builds with four ballots, so
// the per-pixel test needs no dependent load.  Every lane of the wave must
// construct it.
struct WeqLiquidTable {
    unsigned long long w0, w1, w2, w3;  // separate values: an array would be indexed through scratch
    __device__ explicit WeqLiquidTable(const uint32_t *__restrict__ props) {
        const uint32_t lane = threadIdx.x & 63u;
        w0 = __ballot(props[lane * 32u + 3u] == 2u);
        w1 = __ballot(props[(64u + lane) * 32u + 3u] == 2u);
        w2 = __ballot(props[(128u + lane) * 32u + 3u] == 2u);
        w3 = __ballot(props[(192u + lane) * 32u + 3u] == 2u);
    }
    __device__ bool operator()(uint32_t p) const {
        const uint32_t ty = p & 0x1FFu;
        if (p == 0u || ty >= 256u) return false;
        const uint32_t q = ty >> 6;
        const unsigned long long word = (w0 & (0ull - (q == 0u))) | (w1 & (0ull - (q == 1u))) |
                                        (w2 & (0ull - (q == 2u))) | (w3 & (0ull - (q == 3u)));
        return (word >> (ty & 63u)) & 1ull;
    }
};

// Column pass, part 1: the first and last liquid row of each (column,
// 64-row chunk), folded into the column's min / max (max stored as row + 1).
// Each of a block's four waves takes 16 rows of the chunk, all 16 loads in
// flight at once; the waves combine in LDS and one atomic pair per column
// and chunk goes out.
__global__ __launch_bounds__(256) void weq_column_chunks(const uint32_t *__restrict__ pmap,
                                                         const uint32_t *__restrict__ props,
                                                         uint32_t *__restrict__ col_min, uint32_t *__restrict__ col_max,
                                                         int xres, int yres) {
    __shared__ uint32_t s_lo[4][64], s_hi[4][64];
    const WeqLiquidTable liquid(props);
    const int tx = threadIdx.x, w = threadIdx.y;
    const int x = blockIdx.x * 64 + tx;
    const int y0 = blockIdx.y * WEQ_CHUNK + w * 16, y1 = min(y0 + 16, yres);
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    if (x < xres) {
        uint32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = y0 + u < y1 ? pmap[(y0 + u) * xres + x] : 0u;
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (liquid(v[u])) {
                if (lo == 0xFFFFFFFFu) lo = (uint32_t)(y0 + u);
                hi = (uint32_t)(y0 + u) + 1u;
            }
    }
    s_lo[w][tx] = lo;
    s_hi[w][tx] = hi;
    __syncthreads();
    if (w != 0 || x >= xres) return;
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        lo = min(lo, s_lo[k][tx]);
        hi = max(hi, s_hi[k][tx]);
    }
    if (hi) {
        atomicMin(&col_min[x], lo);
        atomicMax(&col_max[x], hi);
    }
}

// Column pass, part 2 (water_column_analyze_kernel, column.rs:20-82): the
// column's surface word, the target and accumulators reset for the next use.
__global__ void weq_column_surfaces(const uint32_t *__restrict__ pmap, const uint4 *__restrict__ cellrec,
                                    uint32_t *__restrict__ col_min, uint32_t *__restrict__ col_max,
                                    uint32_t *__restrict__ col_surf, unsigned long long *__restrict__ col_target,
                                    uint32_t *__restrict__ counters, int xres, int xc) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x == 0) counters[TPT_CTR_WEQ] = 0u;
    if (x >= xres) return;
    const uint32_t lo = col_min[x], hi = col_max[x];
    col_min[x] = 0xFFFFFFFFu;
    col_max[x] = 0u;
    col_target[x] = 0ull;
    if (hi == 0u) {
        col_surf[x] = 0xFFFFFFFFu;
        return;
    }
    bool free_above = false;
    if (lo > 0u && pmap[(lo - 1u) * (uint32_t)xres + (uint32_t)x] == 0u) {
        const uint4 r = cellrec[((lo - 1u) / 4u) * (uint32_t)xc + (uint32_t)x / 4u];
        free_above = !wall_blocks_particle(r.z, r.w, 0x2u);
    }
    col_surf[x] = (lo << 16) | (free_above ? 1u << 15 : 0u) | ((hi - 1u) & 0x7FFFu);
}

// Row pass (water_row_connect_kernel, row.rs:22-150), one workgroup per row.
// Lane t of wave v owns pixels x = 256 u + 64 v + t, so each (wave, u) pair
// holds one 64-pixel word of the row: a ballot gives the word's liquid and
// linkable bits, one wave turns them into span-start bits and an exclusive
// max-scan of the words' last starts, and every liquid pixel reads its span
// start off those.  The span reductions are LDS atomics indexed by the
// span's start column; only start columns are initialised.
constexpr int WEQ_WORDS = WEQ_PER * (WEQ_BLK / 64);
static_assert(WEQ_WORDS <= 64, "one wave scans the row's words");

__global__ __launch_bounds__(WEQ_BLK) void weq_rows(const uint32_t *__restrict__ pmap,
                                                    const uint32_t *__restrict__ props,
                                                    const uint4 *__restrict__ cellrec,
                                                    const uint32_t *__restrict__ col_surf,
                                                    unsigned long long *__restrict__ col_target,
                                                    uint32_t *__restrict__ winner, int xres, int yres, int xc,
                                                    const uint32_t *__restrict__ seedp) {
    extern __shared__ uint32_t weq_lds[];  // min_top, max_free, dmin, dmax: xres words each
    uint32_t *min_top = weq_lds, *max_free = weq_lds + xres, *dmin = weq_lds + 2 * xres, *dmax = weq_lds + 3 * xres;
    __shared__ unsigned long long word_liq[WEQ_WORDS], word_link[WEQ_WORDS], word_start[WEQ_WORDS];
    __shared__ int word_carry[WEQ_WORDS], row_any;
    const int y = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int nwords = (xres + 63) >> 6;
    // All of the row's loads are issued before any is used.
    // Four adjacent lanes share a cell: its wall and emap words are loaded
    // by the first of them and passed on by a shuffle.
    uint32_t pix[WEQ_PER], sf[WEQ_PER];
    uint2 we[WEQ_PER];
    const uint2 *__restrict__ cell_we = reinterpret_cast<const uint2 *>(cellrec);
#pragma unroll
    for (int u = 0; u < WEQ_PER; ++u) {
        const int x = u * WEQ_BLK + t;
        const bool in = x < xres;
        pix[u] = in ? pmap[y * xres + x] : 0u;
        sf[u] = in ? col_surf[x] : 0xFFFFFFFFu;
        we[u] = in && (lane & 3) == 0 ? cell_we[((y >> 2) * xc + (x >> 2)) * 2 + 1] : make_uint2(0u, 0u);
    }
    const WeqLiquidTable liquid(props);
    bool liq[WEQ_PER];
#pragma unroll
    for (int u = 0; u < WEQ_PER; ++u) {
        liq[u] = liquid(pix[u]);
        const uint32_t wall = (uint32_t)__shfl((int)we[u].x, lane & ~3), emap = (uint32_t)__shfl((int)we[u].y, lane & ~3);
        const bool link = !wall_blocks_particle(wall, emap, 0x2u);
        const unsigned long long lw = __ballot(liq[u]), kw = __ballot(link);
        const int w = u * (WEQ_BLK / 64) + wv;
        if (lane == 0 && w < nwords) {
            word_liq[w] = lw;
            word_link[w] = kw;
        }
    }
    __syncthreads();
    // A liquid pixel starts a span unless its left neighbour is liquid and
    // its own cell lets liquids through.
    if (wv == 0) {
        const int w = lane;
        unsigned long long st = 0ull;
        int last = -1;
        if (w < nwords) {
            const unsigned long long l = word_liq[w];
            const unsigned long long prev = (l << 1) | (w > 0 ? word_liq[w - 1] >> 63 : 0ull);
            st = l & ~(prev & word_link[w]);
            if (st) last = w * 64 + 63 - __clzll((long long)st);
            word_start[w] = st;
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(last, off);
            if (lane >= off && v > last) last = v;
        }
        const int excl = __shfl_up(last, 1);
        if (w < nwords) word_carry[w] = lane > 0 ? excl : -1;
        if (lane == 63) row_any = last >= 0;
    }
    __syncthreads();
    if (!row_any) return;
    int start[WEQ_PER];
#pragma unroll
    for (int u = 0; u < WEQ_PER; ++u) {
        const int x = u * WEQ_BLK + t, w = u * (WEQ_BLK / 64) + wv;
        start[u] = -1;
        if (!liq[u]) continue;
        const unsigned long long below = word_start[w] & ((2ull << lane) - 1ull);
        start[u] = below ? w * 64 + 63 - __clzll((long long)below) : word_carry[w];
        if (start[u] == x) {
            min_top[x] = (uint32_t)yres;
            max_free[x] = 0u;
            dmin[x] = (uint32_t)xres;
            dmax[x] = 0u;
        }
    }
    __syncthreads();
    const uint32_t seed = *seedp;
    // Per span: the highest surface, then the lowest free surface and the
    // columns that have it.
    bool inrow[WEQ_PER];
#pragma unroll
    for (int u = 0; u < WEQ_PER; ++u) {
        const uint32_t top = sf[u] >> 16, bot = sf[u] & 0x7FFFu;
        inrow[u] = start[u] >= 0 && sf[u] != 0xFFFFFFFFu && (uint32_t)y >= top && (uint32_t)y <= bot;
        if (inrow[u]) {
            atomicMin(&min_top[start[u]], top);
            if (sf[u] & (1u << 15)) atomicMax(&max_free[start[u]], top);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < WEQ_PER; ++u) {
        const int k = u * WEQ_BLK + t;
        if (inrow[u] && (sf[u] & (1u << 15)) && (sf[u] >> 16) == max_free[start[u]]) {
            atomicMin(&dmin[start[u]], (uint32_t)k);
            atomicMax(&dmax[start[u]], (uint32_t)k);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < WEQ_PER; ++u) {
        const int k = u * WEQ_BLK + t;
        if (start[u] < 0 || sf[u] == 0xFFFFFFFFu) continue;
        const uint32_t s0 = (uint32_t)start[u], mf = max_free[s0];
        if (!(mf > min_top[s0] + 1u)) continue;
        const uint32_t dest_row = mf - 1u;
        if ((sf[u] >> 16) >= dest_row) continue;
        const uint32_t span_len = dmax[s0] - dmin[s0] + 1u;
        const uint32_t dest_x = dmin[s0] + ((uint32_t)k + seed) % span_len;
        const uint32_t dest = dest_row * (uint32_t)xres + dest_x;
        atomicMax(&col_target[k], ((unsigned long long)(y + 1) << 32) | (unsigned long long)(dest + 1u));
        winner[dest] = 0xFFFFFFFFu;
    }
}

// The 1-in-200 gate (flood_chance, intent.rs:46-49).
__device__ __forceinline__ bool weq_chance(uint32_t frame_seed, uint32_t id) {
    return fmix32(frame_seed ^ (id * 0x9E3779B9u)) % 200u == 0u;
}

// water_intent_kernel (intent.rs:50-129): eligible liquids bid for their
// column's target with the lowest slot winning; bidders are listed for the
// commit.
__global__ __launch_bounds__(BLK) void weq_intent(Slots s, MoveEnv e, const uint32_t *__restrict__ props,
                                                  const unsigned long long *__restrict__ col_target,
                                                  uint32_t *__restrict__ winner, uint32_t *__restrict__ list,
                                                  uint32_t *__restrict__ counters, uint32_t slots,
                                                  const uint32_t *__restrict__ seedp) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= slots) return;
    const uint32_t ty = s.ty[id];
    if (ty == 0u || ty >= 256u || props[ty * 32u + 3u] != 2u) return;
    if (s.flags[id] & (0x10u | 0x02u)) return;
    if (!weq_chance(*seedp, id)) return;
    const int xi = tpt_f2i(s.x[id] + 0.5f), yi = tpt_f2i(s.y[id] + 0.5f);
    if (xi < 0 || xi >= e.xr || yi < 0 || yi >= e.yr) return;
    if ((e.pmap[yi * e.xr + xi] >> 9) != id + 1u) return;
    const uint32_t packed = (uint32_t)col_target[xi];
    if (packed == 0u) return;
    const uint32_t target = packed - 1u, dest_y = target / (uint32_t)e.xr, dest_x = target % (uint32_t)e.xr;
    const uint32_t occ = e.pmap[target];
    const bool can_enter = eval_move_ty(e, props[ty * 32u + 1u], ty, (int)dest_x, (int)dest_y, occ & 0x1FFu) != 0u;
    if ((uint32_t)yi < dest_y && occ == 0u && can_enter) {
        atomicMin(&winner[target], id + 1u);
        list[atomicAdd(&counters[TPT_CTR_WEQ], 1u)] = id;
    }
}

// water_commit_kernel (commit.rs:15-63) over the bidders: only a bidder
// can hold a target.  Targets are empty pixels, so no winner's old pixel is
// another's target.
__global__ __launch_bounds__(BLK) void weq_commit(Slots s, uint32_t *__restrict__ pmap,
                                                  const unsigned long long *__restrict__ col_target,
                                                  const uint32_t *__restrict__ winner,
                                                  const uint32_t *__restrict__ list,
                                                  const uint32_t *__restrict__ counters, int xres) {
    const uint32_t n = counters[TPT_CTR_WEQ];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t id = list[k], ty = s.ty[id];
        const int xi = tpt_f2i(s.x[id] + 0.5f), yi = tpt_f2i(s.y[id] + 0.5f);
        const uint32_t target = (uint32_t)col_target[xi] - 1u, dest_y = target / (uint32_t)xres;
        if (winner[target] != id + 1u) continue;
        const uint32_t old_idx = (uint32_t)yi * (uint32_t)xres + (uint32_t)xi;
        if ((pmap[old_idx] >> 9) == id + 1u) pmap[old_idx] = 0u;
        pmap[target] = ((id + 1u) << 9) | (ty & 0x1FFu);
        s.x[id] = (float)(target % (uint32_t)xres);
        s.y[id] = (float)dest_y;
    }
}

}  // namespace

int tpt_water_equal_step(tpt_ctx *ctx) {
    const tpt_config &c = ctx->cfg;
    const int xres = (int)c.xres, yres = (int)c.yres, xc = (int)ctx->xcells;
    const uint32_t slots = ctx->slot_bound;
    uint32_t *pmap = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP);
    const uint32_t *props = tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS);
    const uint32_t *seed = ctx->counters + TPT_CTR_SEED;
    tpt_water_bufs &w = ctx->water;
    // The pass moves particles and rewrites pmap: the movement rebuild's
    // pixel records no longer describe the state.
    ctx->maps_fresh = ctx->rebuild_fresh = ctx->rec_cover = false;
    tpt_pass_begin(ctx, PASS_WATER_EQUAL);
    hipLaunchKernelGGL(weq_column_chunks, dim3((xres + 63) / 64, (yres + WEQ_CHUNK - 1) / WEQ_CHUNK), dim3(64, 4),
                       0, ctx->stream, pmap, props, w.col_min, w.col_max, xres, yres);
    hipLaunchKernelGGL(weq_column_surfaces, dim3((xres + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream, pmap,
                       ctx->cell_rec, w.col_min, w.col_max, w.col_surf, w.col_target, ctx->counters, xres, xc);
    hipLaunchKernelGGL(weq_rows, dim3(yres), dim3(WEQ_BLK), 16u * (uint32_t)xres, ctx->stream, pmap, props,
                       ctx->cell_rec, w.col_surf,
                       w.col_target, w.winner, xres, yres, xc, seed);
    Slots s{tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<float>(ctx, TPT_BUF_PART_X),
            tpt_buf<float>(ctx, TPT_BUF_PART_Y),   tpt_buf<float>(ctx, TPT_BUF_PART_VX),
            tpt_buf<float>(ctx, TPT_BUF_PART_VY),  tpt_buf<uint32_t>(ctx, TPT_BUF_PART_FLAGS),
            tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP), tpt_buf<uint32_t>(ctx, TPT_BUF_INTENT_TARGET)};
    MoveEnv e{pmap, ctx->pix_type, ctx->cell_rec, tpt_buf<uint32_t>(ctx, TPT_BUF_MOVE_LUT), xres, yres, xc,
              c.edge_mode, c.gravity_mode, c.custom_gravity_x, c.custom_gravity_y, c.newton_grav, nullptr, nullptr,
              nullptr, nullptr, nullptr};
    hipLaunchKernelGGL(weq_intent, dim3((slots + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream, s, e, props, w.col_target,
                       w.winner, ctx->slot_scratch_u, ctx->counters, slots, seed);
    hipLaunchKernelGGL(weq_commit, dim3(64), dim3(BLK), 0, ctx->stream, s, pmap, w.col_target, w.winner,
                       ctx->slot_scratch_u, ctx->counters, xres);
    tpt_pass_end(ctx, PASS_WATER_EQUAL);
    return tpt_check_launch("water equalisation");
}

