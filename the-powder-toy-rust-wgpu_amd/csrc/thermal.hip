// Conduction and phase transitions for gfx950:
// thermal_kernel::step_thermal_and_transitions_gpu
// (crates/tpt-sim/src/thermal_kernel.rs:1574-1717).
//
// The reference's non-atomic read-modify-writes on shared cells and
// particles are replaced by deterministic two-step forms (a slot pass that
// records per-slot results plus a per-cell pass that resolves them):
//   R2  air_hv: every gated particle reads the pre-pass value; the cell keeps
//       the write of its highest gated slot (atomicMax on slot+1).
//   R3  block_airh: gate failures are counted per cell (atomicAdd) and the
//       saturating "+1 while bit 3 is clear" is applied n times at once.
//   R5  air_pv / wall_emap: transitions read the pre-pass grids; the cell
//       keeps the pv of its highest changed slot; spark charges land after.
//   R6  liquid convection: one round of a priority matching over the
//       proposals (i -> partner) built from pre-pass temperatures: a swap
//       happens iff i is the lowest slot of every proposal touching i or its
//       partner, so the swaps executed are disjoint.
//
// Kernels: th_conduct, th_conduct_cells, th_apply, th_convect_propose,
// th_convect_touch, th_convect_swap, th_transition, th_transition_cells,
// th_refresh.

#include "tpt_internal.h"

namespace {

constexpr int BLK = 256;
constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t PARTICLE_HASH_SALT = 2654435761u;
constexpr uint32_t TRANSITION_HASH_SALT = 0x51ED2701u;

enum : uint32_t {
    T_WATR = 2, T_FIRE = 4, T_LAVA = 6, T_STNE = 5, T_SPRK = 12, T_SNOW = 13, T_ICEI = 14, T_SAND = 17,
    T_SALT = 18, T_SLTW = 19, T_ACID = 20, T_WTRV = 21, T_GLAS = 24, T_PLSM = 28, T_PHOT = 29, T_ELEC = 31,
    T_DSTW = 35, T_LN2 = 38, T_BMTL = 46, T_BRMT = 47, T_HSWC = 68, T_RIME = 79, T_FOG = 80, T_PIPE = 82,
    T_FRZW = 84, T_BIZR = 85, T_BIZRG = 86, T_PLUT = 89, T_BGLA = 95, T_THRM = 99, T_CFLM = 100, T_LIFE = 105,
    T_DEUT = 111, T_FILT = 134, T_BRAY = 136, T_QRTZ = 141, T_PQRT = 142, T_GEL = 150, T_PPIP = 162,
    T_VIBR = 165, T_BVBR = 166, T_TUNG = 171, T_CRMC = 178, T_HEAC = 179, T_LITH = 190, T_BASE = 193,
    T_SEED = 194, T_BISM = 239
};
enum : int {
    COL_PROPERTIES = 1, COL_FLAMMABLE = 12, COL_EXPLOSIVE = 13, COL_HEAT_CONDUCT = 15, COL_DEFAULT_TEMP = 16,
    COL_LOW_PRESSURE = 18, COL_LOW_PRESSURE_TRANSITION = 19, COL_HIGH_PRESSURE = 20,
    COL_HIGH_PRESSURE_TRANSITION = 21, COL_LOW_TEMP = 22, COL_LOW_TEMP_TRANSITION = 23, COL_HIGH_TEMP = 24,
    COL_HIGH_TEMP_TRANSITION = 25
};
constexpr int32_t NT = -1, ST = 512;

struct Parts {
    uint32_t *ty;
    int32_t *life;
    uint32_t *ctype;
    float *x, *y, *temp, *equil;
    int32_t *tmp, *tmp2;
};

struct Dims {
    int xres, yres, xc, yc;
    uint32_t slots;
};

__device__ __forceinline__ uint32_t gpu_hash(uint32_t seed) {
    uint32_t h = seed ^ 0x9E3779B9u;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    return h ^ (h >> 16);
}

__device__ __forceinline__ uint32_t EP(const uint32_t *__restrict__ p, uint32_t t, int col) { return p[t * 32u + col]; }
__device__ __forceinline__ float EPF(const uint32_t *__restrict__ p, uint32_t t, int col) {
    return __uint_as_float(p[t * 32u + col]);
}

__device__ __forceinline__ bool is_heat_insulator(uint32_t ty, int32_t life, int32_t tmp, uint32_t hc) {
    return hc == 0u || (ty == T_HSWC && life != 10) || ((ty == T_PIPE || ty == T_PPIP) && (tmp & 1) == 0);
}

__device__ __forceinline__ float heat_capacity(uint32_t ty, uint32_t ctype) {
    float hc = 1.0f;
    if (ty == T_BASE) hc = 1.5f;
    if ((ty == T_PIPE || ty == T_PPIP) && ctype > 0u && ctype < 256u) hc += ctype == T_BASE ? 1.5f : 1.0f;
    return hc;
}

__device__ __forceinline__ bool pair_excluded(uint32_t t, uint32_t rt, int32_t st, int32_t rtt) {
    return (t == T_FILT && (rt == T_BRAY || rt == T_BIZR || rt == T_BIZRG)) ||
           (rt == T_FILT && (t == T_BRAY || t == T_PHOT || t == T_BIZR || t == T_BIZRG)) ||
           (t == T_ELEC && rt == T_DEUT) || (t == T_DEUT && rt == T_ELEC) ||
           (t == T_HSWC && rt == T_FILT && st == 1) || (t == T_FILT && rt == T_HSWC && rtt == 1);
}

__device__ __forceinline__ uint32_t slot_at(const uint32_t *__restrict__ pmap, int x, int y, Dims d) {
    if (x < 0 || y < 0 || x >= d.xres || y >= d.yres) return NONE;
    const uint32_t packed = pmap[y * d.xres + x];
    return (packed >> 9) ? (packed >> 9) - 1u : NONE;
}

// thermal_conduction_kernel (thermal_kernel.rs:459-605), rules R2/R3.
// One slot, reading its neighbours through pmap and the slot arrays.
__device__ __forceinline__ void conduct_slot(const Parts &p, const uint32_t *__restrict__ pmap,
                                             const float *__restrict__ air_hv, const uint32_t *__restrict__ props,
                                             float *__restrict__ hv_new, uint32_t *__restrict__ hv_key,
                                             uint32_t *__restrict__ fails, const Dims &d, uint32_t aheat,
                                             uint32_t seed, uint32_t i) {
    const uint32_t ty = p.ty[i];
    if (ty == 0u || ty >= 256u) return;
    const uint32_t hcond = EP(props, ty, COL_HEAT_CONDUCT);
    const int x = tpt_f2i(p.x[i]), y = tpt_f2i(p.y[i]);
    if (x < 0 || x >= d.xres || y < 0 || y >= d.yres) return;
    const int ci = (y / TPT_CELL) * d.xc + x / TPT_CELL;
    const int32_t life = p.life[i], tmp = p.tmp[i];
    const uint32_t ctype = p.ctype[i];
    float gel = 1.0f;
    if (ty == T_GEL) gel = (float)tmp * 2.55f;
    float scaled = (float)hcond * gel;
    if (scaled < 0.0f) scaled = 0.0f;
    const uint32_t threshold = tpt_f2u(scaled);
    const uint32_t roll = gpu_hash(seed ^ (i * PARTICLE_HASH_SALT) ^ ty) % 250u;
    if (!is_heat_insulator(ty, life, tmp, hcond) && roll < threshold) {
        const float t0 = p.temp[i];
        float self_temp = t0;
        if (aheat && (EP(props, ty, COL_PROPERTIES) & 0x00200000u) == 0u) {
            const float hv = air_hv[ci];
            const float dtemp = hv - t0;
            const float hc = heat_capacity(ty, ctype);
            float alpha = 0.4f * hc;
            if (alpha > 0.04f) alpha = 0.04f;
            self_temp = tpt_clampf(t0 + alpha * dtemp / hc, 0.0f, 9999.0f);
            hv_new[i] = tpt_clampf(hv - alpha * dtemp, 0.0f, 9999.0f);
            atomicMax(&hv_key[ci], i + 1u);
        }
        const float hc_self = heat_capacity(ty, ctype);
        float c_heat = self_temp * hc_self, hc_total = hc_self;
        for (int ny = 0; ny < 3; ++ny)
            for (int nx = 0; nx < 3; ++nx) {
                if (nx == 1 && ny == 1) continue;
                const uint32_t nid = slot_at(pmap, x + nx - 1, y + ny - 1, d);
                if (nid >= d.slots) continue;
                const uint32_t nty = p.ty[nid];
                if (nty == 0u || nty >= 256u) continue;
                const int32_t ntmp = p.tmp[nid];
                if (is_heat_insulator(nty, p.life[nid], ntmp, EP(props, nty, COL_HEAT_CONDUCT))) continue;
                if (pair_excluded(ty, nty, tmp, ntmp)) continue;
                const float n_hc = heat_capacity(nty, p.ctype[nid]);
                c_heat += p.temp[nid] * n_hc;
                hc_total += n_hc;
            }
        p.equil[i] = tpt_clampf(c_heat / hc_total, 0.0f, 9999.0f);
    } else {
        atomicAdd(&fails[ci], 1u);
        p.equil[i] = -1.0f;
    }
}

__global__ __launch_bounds__(BLK) void th_conduct(Parts p, const uint32_t *__restrict__ pmap,
                                                  const float *__restrict__ air_hv,
                                                  const uint32_t *__restrict__ props, float *__restrict__ hv_new,
                                                  uint32_t *__restrict__ hv_key, uint32_t *__restrict__ fails,
                                                  Dims d, uint32_t aheat, const uint32_t *__restrict__ seedp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t seed = *seedp;
    if (i < d.slots) conduct_slot(p, pmap, air_hv, props, hv_new, hv_key, fails, d, aheat, seed, i);
}

// The off-map list's slots (energy particles, INVIS/FILT, stacked losers,
// off-pixel positions), same per-slot form.
// Also writes the per-cell {grav_fx, grav_fy, air_pv, 0} records the
// convection proposal and the transitions read (th_cell_pack's job on the
// other paths): conduction changes none of the three grids.
__global__ __launch_bounds__(BLK) void th_conduct_list(Parts p, const uint32_t *__restrict__ pmap,
                                                       const float *__restrict__ air_hv,
                                                       const uint32_t *__restrict__ props, float *__restrict__ hv_new,
                                                       uint32_t *__restrict__ hv_key, uint32_t *__restrict__ fails,
                                                       const uint32_t *__restrict__ list,
                                                       const uint32_t *__restrict__ counters, Dims d, uint32_t aheat,
                                                       const uint32_t *__restrict__ seedp,
                                                       const uint32_t *__restrict__ gfx,
                                                       const uint32_t *__restrict__ gfy,
                                                       const uint32_t *__restrict__ pv, uint4 *__restrict__ crec,
                                                       uint32_t cells) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x)
        crec[c] = make_uint4(gfx[c], gfy[c], pv[c], 0u);
    const uint32_t n = counters[TPT_CTR_OFFMAP], seed = *seedp;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t i = list[k];
        if (i < d.slots) conduct_slot(p, pmap, air_hv, props, hv_new, hv_key, fails, d, aheat, seed, i);
    }
}

__global__ void th_conduct_cells(float *__restrict__ air_hv, uint32_t *__restrict__ block_airh,
                                 const float *__restrict__ hv_new, uint32_t *__restrict__ hv_key,
                                 uint32_t *__restrict__ fails, uint32_t cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const uint32_t key = hv_key[c];
    if (key) {
        air_hv[c] = hv_new[key - 1u];
        hv_key[c] = 0u;
    }
    const uint32_t n = fails[c];
    if (n) {
        uint32_t v = block_airh[c];
        if ((v & 8u) == 0u) {
            const uint32_t room = 8u - (v & 7u);
            v += n < room ? n : room;
            block_airh[c] = v;
        }
        fails[c] = 0u;
    }
}

// thermal_apply_kernel (thermal_kernel.rs:616-708).
// With `heat` set, the neighbours' equilibria come from the per-pixel
// conduction results (the pixel-major path leaves them there) instead of
// the slot array.
__device__ __forceinline__ void apply_slot(const Parts &p, const uint32_t *__restrict__ pmap,
                                           const uint32_t *__restrict__ props, const Dims &d, uint32_t j,
                                           const uint2 *__restrict__ heat = nullptr) {
    const uint32_t jty = p.ty[j];
    if (jty == 0u || jty >= 256u) return;
    const int x = tpt_f2i(p.x[j]), y = tpt_f2i(p.y[j]);
    if (x < 0 || x >= d.xres || y < 0 || y >= d.yres) return;
    uint32_t best = j;
    const float own = p.temp[j];
    float best_temp = own;
    bool has = false;
    const float eq = p.equil[j];
    if (eq >= 0.0f) {
        best_temp = eq;
        has = true;
    }
    const int32_t jtmp = p.tmp[j];
    const bool j_ins = is_heat_insulator(jty, p.life[j], jtmp, EP(props, jty, COL_HEAT_CONDUCT));
    if (!j_ins) {
        for (int ny = 0; ny < 3; ++ny)
            for (int nx = 0; nx < 3; ++nx) {
                if (nx == 1 && ny == 1) continue;
                const uint32_t nid = slot_at(pmap, x + nx - 1, y + ny - 1, d);
                if (nid >= d.slots || nid == j) continue;
                const uint32_t nty = p.ty[nid];
                if (nty == 0u || nty >= 256u || !(!has || nid > best)) continue;
                const float neq = heat ? __uint_as_float(heat[(y + ny - 1) * d.xres + (x + nx - 1)].x) : p.equil[nid];
                if (!(neq >= 0.0f)) continue;
                const int32_t ntmp = p.tmp[nid];
                if (is_heat_insulator(nty, p.life[nid], ntmp, EP(props, nty, COL_HEAT_CONDUCT))) continue;
                if (pair_excluded(nty, jty, ntmp, jtmp)) continue;
                best = nid;
                best_temp = neq;
                has = true;
            }
    }
    p.temp[j] = tpt_clampf(has ? best_temp : own, 0.0f, 9999.0f);
}

__global__ __launch_bounds__(BLK) void th_apply(Parts p, const uint32_t *__restrict__ pmap,
                                                const uint32_t *__restrict__ props, Dims d) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < d.slots) apply_slot(p, pmap, props, d, j);
}

// The off-map list's slots; an off-map slot that owns the record of its
// rounded pixel (a flagged owner) also refreshes the record's temperature.
// The apply of the off-map list: entries first, first + stride, ...
__device__ __forceinline__ void apply_list_part(const Parts &p, const uint32_t *__restrict__ pmap,
                                                const uint32_t *__restrict__ props,
                                                const uint32_t *__restrict__ list,
                                                const uint32_t *__restrict__ counters,
                                                const uint2 *__restrict__ heat, unsigned long long *rec,
                                                const Dims &d, uint32_t first, uint32_t stride) {
    const uint32_t n = counters[TPT_CTR_OFFMAP];
    for (uint32_t k = first; k < n; k += stride) {
        const uint32_t j = list[k];
        if (j >= d.slots) continue;
        apply_slot(p, pmap, props, d, j, heat);
        const int rx = tpt_f2i(p.x[j] + 0.5f), ry = tpt_f2i(p.y[j] + 0.5f);
        if (rx < 0 || ry < 0 || rx >= d.xres || ry >= d.yres) continue;
        const size_t pix = (size_t)ry * d.xres + rx;
        const unsigned long long v = rec[pix];
        if ((uint32_t)(v >> 41) == j + 1u)
            rec[pix] = (v & 0xFFFFFFFF00000000ull) | (unsigned long long)__float_as_uint(p.temp[j]);
    }
}

// Pixel-major conduction and apply over the records RecalcFreeParticles
// left (unflagged record owners only; the rest go through the off-map
// list).  Conduction leaves each pixel's equilibrium and flags in `heat`
// for the apply tile; apply leaves the new temperature in the record for
// the convection pass.  A 64x16 pixel tile (16x4 cells) per 256-thread block, each thread
// one 4-pixel column of one cell, the tile plus a one-pixel halo in LDS.
// Each cell's 16 pixels live in 4 adjacent lanes of one wave, so the R2/R3
// cell results are reduced with lane shuffles and written by the block,
// merged with what the off-map list recorded in hv_key / fails.
constexpr int TW = 64, TH = 16, LW = TW + 2, LH = TH + 2;
constexpr int LOAD_ITERS = (LH * LW + TW * 4 - 1) / (TW * 4);  // halo-tile loads per thread

// Types whose conduction reads life / tmp / ctype besides the type itself.
__device__ __forceinline__ bool special_type(uint32_t t) {
    return t == T_HSWC || t == T_PIPE || t == T_PPIP || t == T_GEL;
}

// Per-type flags the tiles look up instead of re-deriving them per
// neighbour.  Types without TF_SPECIAL conduct unless HeatConduct is 0, with
// heat capacity 1 (1.5 for BASE); the pair exclusions only ever pair two
// TF_PAIR types.
constexpr uint32_t TF_HC0 = 1u, TF_SPECIAL = 2u, TF_PAIR = 4u, TF_BASE = 8u, TF_NOAMB = 16u;

__device__ __forceinline__ uint32_t type_flags(const uint32_t *__restrict__ props, uint32_t t) {
    uint32_t f = 0u;
    if (EP(props, t, COL_HEAT_CONDUCT) == 0u) f |= TF_HC0;
    if (special_type(t)) f |= TF_SPECIAL;
    if (t == T_FILT || t == T_BRAY || t == T_PHOT || t == T_BIZR || t == T_BIZRG || t == T_ELEC || t == T_DEUT ||
        t == T_HSWC)
        f |= TF_PAIR;
    if (t == T_BASE) f |= TF_BASE;
    if (EP(props, t, COL_PROPERTIES) & 0x00200000u) f |= TF_NOAMB;
    return f;
}

// Appends v to an LDS list for the lanes where pred holds; every lane of
// the wave calls it.
__device__ __forceinline__ void lds_append(uint32_t *list, uint32_t *count, uint32_t v, bool pred) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = __ffsll((unsigned long long)m) - 1u;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (pred) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
}

// Conduction for the unflagged owners of a 64x16 tile.  Phase 1 stages the
// records (+halo), the type tables and the 16x4 cells' pre-pass air_hv;
// phase 2 writes the pixels with nothing to compute and compacts the rest
// into an LDS list (wave ballot + prefix count); phase 3 runs the owners,
// two or so per thread, reducing R2 (highest gated slot carries its hv,
// a 64-bit LDS max) and R3 (failure count) per cell in LDS; phase 4 merges
// the off-map list's cell results and writes the cells.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void th_conduct_px(Parts p, const unsigned long long *__restrict__ rec,
                                                     float *__restrict__ air_hv, uint32_t *__restrict__ block_airh,
                                                     const uint32_t *__restrict__ props,
                                                     const float *__restrict__ hv_new, uint32_t *__restrict__ hv_key,
                                                     uint32_t *__restrict__ fails, uint2 *__restrict__ heat, Dims d,
                                                     uint32_t aheat, const uint32_t *__restrict__ seedp) {
    const uint32_t seed = *seedp;
    __shared__ uint32_t skey[LH][LW];
    __shared__ float stemp[LH][LW];
    __shared__ uint32_t shc[256], sflag[256];
    __shared__ uint32_t list[TW * TH];
    __shared__ unsigned long long cbest[64];
    __shared__ uint32_t cfail[64];
    __shared__ float chv[64];
    __shared__ uint32_t count;
    const int tx = threadIdx.x, row = threadIdx.y, t = row * TW + tx;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    shc[t] = EP(props, (uint32_t)t, COL_HEAT_CONDUCT);
    sflag[t] = type_flags(props, (uint32_t)t);
    if (t < 64) {
        const int cx = (x0 >> 2) + (t & 15), cy = (y0 >> 2) + (t >> 4);
        cbest[t] = 0ull;
        cfail[t] = 0u;
        chv[t] = (aheat && cx < d.xc && cy < d.yc) ? air_hv[cy * d.xc + cx] : 0.0f;  // pre-pass (R2)
    }
    if (t == 0) count = 0u;
    {
        // All halo loads in flight before the first LDS store.
        unsigned long long v[LOAD_ITERS];
#pragma unroll
        for (int u = 0; u < LOAD_ITERS; ++u) {
            const int e = t + u * TW * 4, ly = e / LW, lx = e - ly * LW;
            const int gx = x0 + lx - 1, gy = y0 + ly - 1;
            v[u] = 0ull;
            if (e < LH * LW && gx >= 0 && gy >= 0 && gx < d.xres && gy < d.yres) v[u] = rec[(size_t)gy * d.xres + gx];
        }
#pragma unroll
        for (int u = 0; u < LOAD_ITERS; ++u) {
            const int e = t + u * TW * 4, ly = e / LW, lx = e - ly * LW;
            if (e < LH * LW) {
                skey[ly][lx] = (uint32_t)(v[u] >> 32);
                stemp[ly][lx] = __uint_as_float((uint32_t)v[u]);
            }
        }
    }
    __syncthreads();
    const int x = x0 + tx;
    for (int k = 0; k < 4; ++k) {
        const int ly = row * 4 + k + 1, y = y0 + row * 4 + k;
        const bool on = x < d.xres && y < d.yres;
        const uint32_t key = on ? skey[ly][tx + 1] : 0u;
        const bool work = key != 0u && (key & 0x100u) == 0u;
        if (on && !work) {
            uint2 out = make_uint2(__float_as_uint(-1.0f), 0u);
            if (key != 0u) {  // flagged owner: its equilibrium came from the off-map list
                const uint32_t s = (key >> 9) - 1u, ty = key & 0xFFu;
                int32_t life = 0, tmp = 0;
                if (sflag[ty] & TF_SPECIAL) {
                    life = p.life[s];
                    tmp = p.tmp[s];
                }
                const bool ins = is_heat_insulator(ty, life, tmp, shc[ty]);
                out = make_uint2(__float_as_uint(p.equil[s]), (ins ? 1u : 0u) | (tmp == 1 ? 2u : 0u));
            }
            heat[(size_t)y * d.xres + x] = out;
        }
        lds_append(list, &count, (uint32_t)((ly - 1) * TW + tx), work);
    }
    __syncthreads();
    const uint32_t n = count;
    for (uint32_t j = (uint32_t)t; j < n; j += TW * 4) {
        const uint32_t li = list[j];
        const int ly = (int)(li / TW) + 1, lx = (int)(li % TW) + 1;
        const int px = x0 + lx - 1, py = y0 + ly - 1;
        const int cl = ((ly - 1) >> 2) * 16 + ((lx - 1) >> 2);
        const uint32_t key = skey[ly][lx];
        const uint32_t s = (key >> 9) - 1u, ty = key & 0xFFu, fo = sflag[ty];
        int32_t life = 0, tmp = 0;
        uint32_t ctype = 0u;
        if (fo & TF_SPECIAL) {
            life = p.life[s];
            tmp = p.tmp[s];
            ctype = p.ctype[s];
        }
        const uint32_t hcond = shc[ty];
        const bool ins = (fo & TF_SPECIAL) ? is_heat_insulator(ty, life, tmp, hcond) : (fo & TF_HC0) != 0u;
        float eq = -1.0f;
        if (s < d.slots) {
            float gel = 1.0f;
            if (ty == T_GEL) gel = (float)tmp * 2.55f;
            float scaled = (float)hcond * gel;
            if (scaled < 0.0f) scaled = 0.0f;
            const uint32_t threshold = tpt_f2u(scaled);
            const uint32_t roll = gpu_hash(seed ^ (s * PARTICLE_HASH_SALT) ^ ty) % 250u;
            if (!ins && roll < threshold) {
                const float hc = (fo & TF_SPECIAL) ? heat_capacity(ty, ctype) : ((fo & TF_BASE) ? 1.5f : 1.0f);
                const float t0 = stemp[ly][lx];
                float self_temp = t0;
                if (aheat && (fo & TF_NOAMB) == 0u) {
                    const float hv0 = chv[cl];
                    const float dtemp = hv0 - t0;
                    float alpha = 0.4f * hc;
                    if (alpha > 0.04f) alpha = 0.04f;
                    self_temp = tpt_clampf(t0 + alpha * dtemp / hc, 0.0f, 9999.0f);
                    const float hn = tpt_clampf(hv0 - alpha * dtemp, 0.0f, 9999.0f);
                    atomicMax(&cbest[cl], ((unsigned long long)(s + 1u) << 32) | (unsigned long long)__float_as_uint(hn));
                }
                float c_heat = self_temp * hc, hc_total = hc;
                for (int ny = 0; ny < 3; ++ny)
                    for (int nx = 0; nx < 3; ++nx) {
                        if (nx == 1 && ny == 1) continue;
                        const uint32_t nk = skey[ly + ny - 1][lx + nx - 1];
                        if (nk == 0u) continue;
                        const uint32_t nid = (nk >> 9) - 1u, nty = nk & 0xFFu;
                        if (nid >= d.slots) continue;
                        const uint32_t fn = sflag[nty];
                        float n_hc;
                        if (fn & TF_SPECIAL) {
                            const int32_t nlife = p.life[nid], ntmp = p.tmp[nid];
                            if (is_heat_insulator(nty, nlife, ntmp, shc[nty])) continue;
                            if ((fo & fn & TF_PAIR) && pair_excluded(ty, nty, tmp, ntmp)) continue;
                            n_hc = heat_capacity(nty, p.ctype[nid]);
                        } else {
                            if (fn & TF_HC0) continue;
                            if ((fo & fn & TF_PAIR) && pair_excluded(ty, nty, tmp, 0)) continue;
                            n_hc = (fn & TF_BASE) ? 1.5f : 1.0f;
                        }
                        c_heat += stemp[ly + ny - 1][lx + nx - 1] * n_hc;
                        hc_total += n_hc;
                    }
                eq = tpt_clampf(c_heat / hc_total, 0.0f, 9999.0f);
            } else {
                atomicAdd(&cfail[cl], 1u);
            }
        }
        if (s < d.slots) p.equil[s] = eq;
        heat[(size_t)py * d.xres + px] = make_uint2(__float_as_uint(eq), (ins ? 1u : 0u) | (tmp == 1 ? 2u : 0u));
    }
    __syncthreads();
    if (t < 64) {
        const int cx = (x0 >> 2) + (t & 15), cy = (y0 >> 2) + (t >> 4);
        if (cx < d.xc && cy < d.yc) {
            const int ci = cy * d.xc + cx;
            uint32_t best_key = (uint32_t)(cbest[t] >> 32);
            float best_hv = __uint_as_float((uint32_t)cbest[t]);
            const uint32_t lk = hv_key[ci];
            if (lk) {
                if (lk > best_key) {
                    best_key = lk;
                    best_hv = hv_new[lk - 1u];
                }
                hv_key[ci] = 0u;
            }
            if (best_key) air_hv[ci] = best_hv;
            uint32_t nfail = cfail[t];
            const uint32_t lf = fails[ci];
            if (lf) {
                nfail += lf;
                fails[ci] = 0u;
            }
            if (nfail) {
                uint32_t v = block_airh[ci];
                if ((v & 8u) == 0u) {
                    const uint32_t room = 8u - (v & 7u);
                    v += nfail < room ? nfail : room;
                    block_airh[ci] = v;
                }
            }
        }
    }
}

// A pixel as an apply candidate, packed in one word from its record key,
// its conduction equilibrium and the conduction flags (bit 0 insulator,
// bit 1 tmp == 1): slot + 1 (bits 0-22) when it can lend its equilibrium
// (a slot below the bound, equilibrium >= 0, not an insulator), else 0;
// TF_PAIR (bit 23) with the type's index among the TF_PAIR types (bits
// 24-26); tmp == 1 (bit 27) and insulator (bit 28) for the pixel's own use.
constexpr uint32_t CAND_ID = 0x7FFFFFu, CAND_PAIR = 1u << 23, CAND_TMP1 = 1u << 27, CAND_INS = 1u << 28;
__constant__ uint32_t cand_pair_types[8] = {T_FILT, T_BRAY, T_PHOT, T_BIZR, T_BIZRG, T_ELEC, T_DEUT, T_HSWC};

__device__ __forceinline__ uint32_t cand_pair_type(uint32_t c) { return cand_pair_types[(c >> 24) & 7u]; }

__device__ __forceinline__ uint32_t cand_info(uint32_t nk, float eq, uint32_t hinfo, const uint32_t *sflag,
                                              uint32_t slots) {
    uint32_t c = ((hinfo & 2u) ? CAND_TMP1 : 0u) | ((hinfo & 1u) ? CAND_INS : 0u);
    if (nk == 0u) return c;
    const uint32_t nid = (nk >> 9) - 1u, nty = nk & 0xFFu;
    if (nid < slots && eq >= 0.0f && (hinfo & 1u) == 0u) c |= nid + 1u;
    if (sflag[nty] & TF_PAIR) {
        uint32_t k = 0u;
        while (cand_pair_types[k] != nty) ++k;
        c |= CAND_PAIR | (k << 24);
    }
    return c;
}

// Apply for the unflagged owners of a tile: phase 2 compacts the owners,
// phase 3 computes them.
// Blocks past the tile rows apply the off-map list: the listed particles
// and the unflagged owners the tiles handle are disjoint, and neither side
// reads what the other writes (a listed particle's temperature, or its
// record's temperature word).
__global__ __launch_bounds__(256) void th_apply_px(Parts p, const unsigned long long *__restrict__ rec,
                                                   uint32_t *__restrict__ rec_temp, const uint2 *__restrict__ heat,
                                                   const uint32_t *__restrict__ props, Dims d,
                                                   const uint32_t *__restrict__ pmap,
                                                   const uint32_t *__restrict__ offmap,
                                                   const uint32_t *__restrict__ counters,
                                                   unsigned long long *rec_list, int tile_rows, int list_rows,
                                                   uint32_t *__restrict__ conv_tgt, uint32_t *__restrict__ conv_touch) {
    if ((int)blockIdx.y >= tile_rows) {
        const int extra = (int)blockIdx.y - tile_rows;
        const uint32_t t = threadIdx.y * TW + threadIdx.x;
        if (extra < list_rows) {
            const uint32_t nb = (uint32_t)list_rows * gridDim.x, b = (uint32_t)extra * gridDim.x + blockIdx.x;
            apply_list_part(p, pmap, props, offmap, counters, heat, rec_list, d, b * 256u + t, nb * 256u);
        } else {
            // The convection pass's proposal and touch words start at NONE
            // for every slot (neither buffer is read by the apply).
            const uint32_t nb = (gridDim.y - (uint32_t)(tile_rows + list_rows)) * gridDim.x;
            const uint32_t b = (uint32_t)(extra - list_rows) * gridDim.x + blockIdx.x;
            for (uint32_t k = b * 256u + t; k < d.slots; k += nb * 256u) {
                conv_tgt[k] = NONE;
                conv_touch[k] = NONE;
            }
        }
        return;
    }
    __shared__ uint32_t skey[LH][LW];
    __shared__ float seq[LH][LW];
    __shared__ uint32_t scand[LH][LW];  // each pixel as an apply candidate: cand_info()
    __shared__ float sown[TH][TW];  // record temperatures of the tile itself
    __shared__ uint32_t sflag[256];
    __shared__ uint32_t list[TW * TH];
    __shared__ uint32_t count;
    const int tx = threadIdx.x, row = threadIdx.y, t = row * TW + tx;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    sflag[t] = type_flags(props, (uint32_t)t);
    if (t == 0) count = 0u;
    {
        // All halo loads in flight before the first LDS store (every canvas
        // pixel's heat entry was written by the conduction tile).
        unsigned long long v[LOAD_ITERS];
        uint2 h[LOAD_ITERS];
#pragma unroll
        for (int u = 0; u < LOAD_ITERS; ++u) {
            const int e = t + u * TW * 4, ly = e / LW, lx = e - ly * LW;
            const int gx = x0 + lx - 1, gy = y0 + ly - 1;
            v[u] = 0ull;
            h[u] = make_uint2(__float_as_uint(-1.0f), 0u);
            if (e < LH * LW && gx >= 0 && gy >= 0 && gx < d.xres && gy < d.yres) {
                const size_t pix = (size_t)gy * d.xres + gx;
                v[u] = rec[pix];
                h[u] = heat[pix];
            }
        }
        __syncthreads();  // sflag
#pragma unroll
        for (int u = 0; u < LOAD_ITERS; ++u) {
            const int e = t + u * TW * 4, ly = e / LW, lx = e - ly * LW;
            if (e < LH * LW) {
                if (ly >= 1 && ly <= TH && lx >= 1 && lx <= TW) sown[ly - 1][lx - 1] = __uint_as_float((uint32_t)v[u]);
                const uint32_t nk = (uint32_t)(v[u] >> 32);
                const float eq = __uint_as_float(h[u].x);
                skey[ly][lx] = nk;
                seq[ly][lx] = eq;
                scand[ly][lx] = cand_info(nk, eq, h[u].y, sflag, d.slots);
            }
        }
    }
    __syncthreads();
    const int x = x0 + tx;
    for (int k = 0; k < 4; ++k) {
        const int ly = row * 4 + k + 1, y = y0 + row * 4 + k;
        const bool on = x < d.xres && y < d.yres;
        const uint32_t key = on ? skey[ly][tx + 1] : 0u;
        const bool work = key != 0u && (key & 0x100u) == 0u && ((key >> 9) - 1u) < d.slots;
        lds_append(list, &count, (uint32_t)((ly - 1) * TW + tx), work);
    }
    __syncthreads();
    const uint32_t n = count;
    for (uint32_t jj = (uint32_t)t; jj < n; jj += TW * 4) {
        const uint32_t li = list[jj];
        const int ly = (int)(li / TW) + 1, lx = (int)(li % TW) + 1;
        const uint32_t key = skey[ly][lx];
        const uint32_t j = (key >> 9) - 1u, jty = key & 0xFFu, fo = sflag[jty];
        const float own = sown[ly - 1][lx - 1];
        uint32_t best = j;
        float best_temp = own;
        bool has = false;
        const float eq = seq[ly][lx];
        if (eq >= 0.0f) {
            best_temp = eq;
            has = true;
        }
        const uint32_t jinfo = scand[ly][lx];
        if ((jinfo & CAND_INS) == 0u) {
            const int32_t jtmp1 = (jinfo & CAND_TMP1) ? 1 : 0;
            for (int ny = 0; ny < 3; ++ny)
                for (int nx = 0; nx < 3; ++nx) {
                    if (nx == 1 && ny == 1) continue;
                    const uint32_t c = scand[ly + ny - 1][lx + nx - 1];
                    const uint32_t nid1 = c & CAND_ID;
                    if (nid1 == 0u || nid1 - 1u == j) continue;
                    const uint32_t nid = nid1 - 1u;
                    if (!(!has || nid > best)) continue;
                    if ((fo & TF_PAIR) && (c & CAND_PAIR) &&
                        pair_excluded(cand_pair_type(c), jty, (c & CAND_TMP1) ? 1 : 0, jtmp1))
                        continue;
                    best = nid;
                    best_temp = seq[ly + ny - 1][lx + nx - 1];
                    has = true;
                }
        }
        const float nt = tpt_clampf(has ? best_temp : own, 0.0f, 9999.0f);
        // The record's temperature is the slot's current one, so an unchanged
        // value needs no store (a random store costs a whole line).
        if (__float_as_uint(nt) != __float_as_uint(own)) p.temp[j] = nt;
        // The record keeps the post-apply temperature for the convection pass.
        rec_temp[2 * ((size_t)(y0 + ly - 1) * d.xres + (x0 + lx - 1))] = __float_as_uint(nt);
    }
}

struct GravCfg {
    uint32_t mode;
    float cgx, cgy;
};

// The liquid's convection partner (thermal_kernel.rs:755-820) from the
// pre-pass state, or NONE.
__device__ uint32_t convect_target(const Parts &p, const uint32_t *__restrict__ pmap,
                                   const uint32_t *__restrict__ props, const float *__restrict__ gfx,
                                   const float *__restrict__ gfy, Dims d, GravCfg g, uint32_t i, uint32_t seed) {
    const uint32_t t = p.ty[i];
    if (t == 0u || t >= 256u) return NONE;
    const int x = tpt_f2i(p.x[i]), y = tpt_f2i(p.y[i]);
    if (x < 0 || x >= d.xres || y < 0 || y >= d.yres) return NONE;
    if ((EP(props, t, COL_PROPERTIES) & 0x02u) == 0u) return NONE;
    if (t == T_GEL) {
        const uint32_t roll = gpu_hash(seed ^ (i * PARTICLE_HASH_SALT) ^ t ^ TRANSITION_HASH_SALT);
        const float gel = (float)p.tmp[i] * 2.55f;
        const uint32_t gate_roll = gpu_hash(roll ^ 0x1B873593u) % 255u;
        if (gel <= 1.0f + (float)gate_roll) return NONE;
    }
    const int ci = (y / TPT_CELL) * d.xc + x / TPT_CELL;
    const float pg = -2.0f;
    float gx = 0.0f, gy = 0.0f;
    if (g.mode == 0u) {
        gy = pg;
    } else if (g.mode == 2u) {
        const float dx = (float)x - (float)d.xres * 0.5f, dy = (float)y - (float)d.yres * 0.5f;
        if (dx != 0.0f || dy != 0.0f) {
            const float dd = 0.01f - sqrtf(dx * dx + dy * dy);
            gx = pg * (dx / dd);
            gy = pg * (dy / dd);
        }
    } else if (g.mode == 3u) {
        gx = pg * g.cgx;
        gy = pg * g.cgy;
    }
    gx += -2.0f * gfx[ci];
    gy += -2.0f * gfy[ci];
    const float rxf = gx + (float)x, ryf = gy + (float)y;
    const int rx = rxf >= 0.0f ? tpt_f2i(rxf + 0.5f) : tpt_f2i(rxf - 0.5f);
    const int ry = ryf >= 0.0f ? tpt_f2i(ryf + 0.5f) : tpt_f2i(ryf - 0.5f);
    const int ox = min(max(rx, x - 1), x + 1), oy = min(max(ry, y - 1), y + 1);
    if (ox == x && oy == y) return NONE;
    const uint32_t nid = slot_at(pmap, ox, oy, d);
    if (nid >= d.slots || p.ty[nid] != t || !(p.temp[i] > p.temp[nid])) return NONE;
    return nid;
}

// convect_target with the partner read from the pixel records (slot, type
// and the post-apply temperature in one 8-byte load) instead of pmap plus
// two slot-array gathers.  Valid while the records are current: types as
// at RecalcFreeParticles, temperatures as written by the apply kernels.
// The convection proposal of liquid slot i of type t at pixel (x, y) with
// temperature own_temp, against the pixel records (convect_target with the
// partner read from one 8-byte record).
__device__ __forceinline__ uint32_t propose_rec(const Parts &p, const unsigned long long *__restrict__ rec,
                                                const uint4 *__restrict__ crec, const Dims &d, const GravCfg &g,
                                                uint32_t i, uint32_t t, int x, int y, float own_temp,
                                                uint32_t seed) {
    if (t == T_GEL) {
        const uint32_t roll = gpu_hash(seed ^ (i * PARTICLE_HASH_SALT) ^ t ^ TRANSITION_HASH_SALT);
        const float gel = (float)p.tmp[i] * 2.55f;
        const uint32_t gate_roll = gpu_hash(roll ^ 0x1B873593u) % 255u;
        if (gel <= 1.0f + (float)gate_roll) return NONE;
    }
    const int ci = (y / TPT_CELL) * d.xc + x / TPT_CELL;
    const float pg = -2.0f;
    float gx = 0.0f, gy = 0.0f;
    if (g.mode == 0u) {
        gy = pg;
    } else if (g.mode == 2u) {
        const float dx = (float)x - (float)d.xres * 0.5f, dy = (float)y - (float)d.yres * 0.5f;
        if (dx != 0.0f || dy != 0.0f) {
            const float dd = 0.01f - sqrtf(dx * dx + dy * dy);
            gx = pg * (dx / dd);
            gy = pg * (dy / dd);
        }
    } else if (g.mode == 3u) {
        gx = pg * g.cgx;
        gy = pg * g.cgy;
    }
    const uint2 gf = *(const uint2 *)&crec[ci];
    gx += -2.0f * __uint_as_float(gf.x);
    gy += -2.0f * __uint_as_float(gf.y);
    const float rxf = gx + (float)x, ryf = gy + (float)y;
    const int rx = rxf >= 0.0f ? tpt_f2i(rxf + 0.5f) : tpt_f2i(rxf - 0.5f);
    const int ry = ryf >= 0.0f ? tpt_f2i(ryf + 0.5f) : tpt_f2i(ryf - 0.5f);
    const int ox = min(max(rx, x - 1), x + 1), oy = min(max(ry, y - 1), y + 1);
    if ((ox != x || oy != y) && ox >= 0 && oy >= 0 && ox < d.xres && oy < d.yres) {
        const unsigned long long v = rec[(size_t)oy * d.xres + ox];
        const uint32_t key = (uint32_t)(v >> 32), nid = (key >> 9) - 1u;
        if (key != 0u && nid < d.slots && (key & 0xFFu) == t && own_temp > __uint_as_float((uint32_t)v))
            return nid;
    }
    return NONE;
}

// Pixel-major convection proposals over the records RecalcFreeParticles left:
// one thread per pixel, its unflagged owner proposing (the record holds its
// type, and its temperature as the apply left it; for such owners the
// rounded pixel is the truncated one the proposal uses).  Blocks past the
// pixels take the off-map list's slots.  Every slot's tgt / touch was set
// to NONE by the apply launch, so only proposals are stored.
__global__ __launch_bounds__(BLK) void th_convect_propose_px(Parts p, const unsigned long long *__restrict__ rec,
                                                             const uint32_t *__restrict__ props,
                                                             const uint4 *__restrict__ crec,
                                                             const uint32_t *__restrict__ offmap,
                                                             const uint32_t *__restrict__ counters,
                                                             uint32_t *__restrict__ tgt, Dims d, GravCfg g,
                                                             const uint32_t *__restrict__ seedp,
                                                             uint32_t pixel_blocks) {
    const uint32_t seed = *seedp;
    if (blockIdx.x >= pixel_blocks) {
        const uint32_t n = counters[TPT_CTR_OFFMAP];
        const uint32_t stride = (gridDim.x - pixel_blocks) * blockDim.x;
        for (uint32_t k = (blockIdx.x - pixel_blocks) * blockDim.x + threadIdx.x; k < n; k += stride) {
            const uint32_t i = offmap[k];
            if (i >= d.slots) continue;
            const uint32_t t = p.ty[i];
            if (t == 0u || t >= 256u || (EP(props, t, COL_PROPERTIES) & 0x02u) == 0u) continue;
            const int x = tpt_f2i(p.x[i]), y = tpt_f2i(p.y[i]);
            if (x < 0 || x >= d.xres || y < 0 || y >= d.yres) continue;
            const uint32_t target = propose_rec(p, rec, crec, d, g, i, t, x, y, p.temp[i], seed);
            if (target != NONE) tgt[i] = target;
        }
        return;
    }
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint32_t)(d.xres * d.yres)) return;
    const unsigned long long v = rec[idx];
    const uint32_t key = (uint32_t)(v >> 32);
    if (key == 0u || (key & 0x100u)) return;
    const uint32_t i = (key >> 9) - 1u, t = key & 0xFFu;
    if (i >= d.slots || (EP(props, t, COL_PROPERTIES) & 0x02u) == 0u) return;
    const int x = (int)(idx % (uint32_t)d.xres), y = (int)(idx / (uint32_t)d.xres);
    const uint32_t target = propose_rec(p, rec, crec, d, g, i, t, x, y, __uint_as_float((uint32_t)v), seed);
    if (target != NONE) tgt[i] = target;
}

__global__ __launch_bounds__(BLK) void th_convect_propose(Parts p, const uint32_t *__restrict__ pmap,
                                                          const uint32_t *__restrict__ props,
                                                          const float *__restrict__ gfx,
                                                          const float *__restrict__ gfy, uint32_t *__restrict__ tgt,
                                                          uint32_t *__restrict__ touch, Dims d, GravCfg g,
                                                          const uint32_t *__restrict__ seedp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.slots) return;
    tgt[i] = convect_target(p, pmap, props, gfx, gfy, d, g, i, *seedp);
    touch[i] = NONE;
}

__global__ __launch_bounds__(BLK) void th_convect_touch(const uint32_t *__restrict__ tgt, uint32_t *__restrict__ touch,
                                                        uint32_t slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint32_t t = tgt[i];
    if (t == NONE) return;
    atomicMin(&touch[i], i);
    atomicMin(&touch[t], i);
}

__global__ __launch_bounds__(BLK) void th_convect_swap(const uint32_t *__restrict__ tgt,
                                                       const uint32_t *__restrict__ touch, float *__restrict__ temp,
                                                       uint32_t slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint32_t t = tgt[i];
    if (t == NONE || touch[i] != i || touch[t] != i) return;
    const float hot = temp[i];
    temp[i] = temp[t];
    temp[t] = hot;
}

__device__ float pres_sum(const float *__restrict__ pv, int x, int y, Dims d) {
    const int cy = y / TPT_CELL, cx = x / TPT_CELL;
    const int ym = max(y - 2, 0) / TPT_CELL, xm = max(x - 2, 0) / TPT_CELL;
    const int yp = min(y + 2, d.yres - 1) / TPT_CELL, xp = min(x + 2, d.xres - 1) / TPT_CELL;
    const float v = (pv[cy * d.xc + cx] + pv[ym * d.xc + cx] + pv[yp * d.xc + cx] + pv[cy * d.xc + xm] +
                     pv[cy * d.xc + xp]) *
                    2.0f;
    return v > 0.0f ? v : 0.0f;
}

// phase_transition_kernel after the convection swap (thermal_kernel.rs:822-1315), rule R5.
// Per-cell {grav_fx, grav_fy, air_pv, 0} for the convection proposal and
// the transitions: one random line per particle instead of three.
__global__ void th_cell_pack(const uint32_t *__restrict__ gfx, const uint32_t *__restrict__ gfy,
                             const uint32_t *__restrict__ pv, uint4 *__restrict__ rec, uint32_t cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells) rec[c] = make_uint4(gfx[c], gfy[c], pv[c], 0u);
}

// crec: the per-cell {grav_fx, grav_fy, air_pv} record th_cell_pack wrote.
__global__ __launch_bounds__(BLK) void th_transition(Parts p, const float *__restrict__ air_pv,
                                                     const uint4 *__restrict__ crec,
                                                     const uint32_t *__restrict__ wall_bmap,
                                                     const uint32_t *__restrict__ wall_emap,
                                                     const uint32_t *__restrict__ props, float *__restrict__ pv_new,
                                                     uint32_t *__restrict__ pv_key, uint32_t *__restrict__ spark,
                                                     uint32_t *__restrict__ pmap_refresh, Dims d, uint32_t legacy,
                                                     const uint32_t *__restrict__ seedp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t seed = *seedp;
    if (i >= d.slots) return;
    // The slot's fields are loaded together up front; testing them one by
    // one would chain a memory latency per branch.
    uint32_t t = p.ty[i];
    const float fx = p.x[i], fy = p.y[i];
    const uint32_t ctype0 = p.ctype[i];
    const int32_t tmp0 = p.tmp[i], tmp20 = p.tmp2[i], life0 = p.life[i];
    const float eq0 = p.equil[i];
    if (t == 0u || t >= 256u) return;
    const uint32_t t_entry = t;
    const int x = tpt_f2i(fx), y = tpt_f2i(fy);
    if (x < 0 || x >= d.xres || y < 0 || y >= d.yres) return;
    const int ci = (y / TPT_CELL) * d.xc + x / TPT_CELL;
    const uint4 cr = crec[ci];
    const float pv0 = __uint_as_float(cr.z);
    float pv = pv0;
    const uint32_t roll = gpu_hash(seed ^ (i * PARTICLE_HASH_SALT) ^ t ^ TRANSITION_HASH_SALT);
    if (!legacy) {
        const uint32_t tprops0 = EP(props, t, COL_PROPERTIES);
        uint32_t ctype = ctype0;
        const int32_t tmp = tmp0, tmp2 = tmp20;
        const float eq = eq0;
        if (eq >= 0.0f) {
            const float pt = eq;
            if ((t == T_ICEI || t == T_SNOW) && (ctype == 0u || ctype >= 256u || ctype == T_ICEI || ctype == T_SNOW)) {
                ctype = T_WATR;
                p.ctype[i] = ctype;
            }
            const float high_temp = EPF(props, t, COL_HIGH_TEMP);
            const int32_t high_tr = (int32_t)EP(props, t, COL_HIGH_TEMP_TRANSITION);
            const float low_temp = EPF(props, t, COL_LOW_TEMP);
            const int32_t low_tr = (int32_t)EP(props, t, COL_LOW_TEMP_TRANSITION);
            const float pv_shift = 2.0f * pv;
            bool boil = (tprops0 & 0x02u) && high_tr >= 0 && high_tr < 256 &&
                        (EP(props, (uint32_t)high_tr, COL_PROPERTIES) & 0x08u);
            if (t == T_LN2 || t == T_SLTW) boil = true;
            bool cond = (tprops0 & 0x08u) && low_tr >= 0 && low_tr < 256 &&
                        (EP(props, (uint32_t)low_tr, COL_PROPERTIES) & 0x02u);
            if (t == T_WTRV) cond = true;
            const float ctemph = boil ? pt - pv_shift : pt;
            const float ctempl = cond ? pt - pv_shift : pt;
            uint32_t old_ty = t, nt = t, new_ctype = ctype;
            int32_t new_life = life0, new_tmp = tmp, new_tmp2 = tmp2;
            bool s = false;
            if (high_tr != NT && ctemph >= high_temp) {
                if (high_tr != ST) {
                    if (t == T_FOG) new_ctype = 0u;
                    nt = (uint32_t)high_tr;
                    s = true;
                } else if (t == T_ICEI || t == T_SNOW) {
                    bool melts = false;
                    if (new_ctype > 0u && new_ctype < 256u && new_ctype != t) {
                        const int32_t ct_tr = (int32_t)EP(props, new_ctype, COL_LOW_TEMP_TRANSITION);
                        const float ct_low = EPF(props, new_ctype, COL_LOW_TEMP);
                        if (ct_tr == (int32_t)T_ICEI || ct_tr == (int32_t)T_SNOW) {
                            if (pt >= ct_low) melts = true;
                        } else if (pt >= 273.15f) {
                            melts = true;
                        }
                    }
                    if (melts) {
                        nt = new_ctype;
                        new_ctype = 0u;
                        new_life = 0;
                        s = true;
                    }
                } else if (t == T_SLTW) {
                    nt = (roll % 4u == 0u) ? T_SALT : T_WTRV;
                    s = true;
                } else if (t == T_BRMT) {
                    if (new_ctype == T_TUNG) {
                        if (ctemph >= EPF(props, T_TUNG, COL_HIGH_TEMP)) {
                            nt = T_LAVA;
                            old_ty = T_TUNG;
                            s = true;
                        }
                    } else if (ctemph >= high_temp) {
                        nt = T_LAVA;
                        s = true;
                    }
                } else if (t == T_CRMC) {
                    if (ctemph >= pres_sum(air_pv, x, y, d) + high_temp) {
                        nt = T_LAVA;
                        s = true;
                    }
                } else if (t == T_RIME) {
                    if (new_tmp > 5) {
                        nt = T_ACID;
                        new_life = 25 + 5 * new_tmp;
                        new_tmp = 0;
                    } else {
                        nt = new_ctype == T_DSTW ? T_DSTW : T_WATR;
                        new_ctype = 0u;
                    }
                    s = true;
                }
            } else if (low_tr != NT && ctempl < low_temp) {
                if (low_tr != ST) {
                    nt = (uint32_t)low_tr;
                    s = true;
                } else if (t == T_WTRV) {
                    nt = pt < 273.0f ? T_RIME : T_DSTW;
                    s = true;
                } else if (t == T_LAVA) {
                    if (new_ctype > 0u && new_ctype < 256u && new_ctype != T_LAVA) {
                        const float c_high = EPF(props, new_ctype, COL_HIGH_TEMP);
                        const int32_t c_high_tr = (int32_t)EP(props, new_ctype, COL_HIGH_TEMP_TRANSITION);
                        bool freeze = false;
                        if (new_ctype == T_THRM) {
                            freeze = pt < EPF(props, T_BMTL, COL_HIGH_TEMP);
                        } else if (new_ctype == T_VIBR || new_ctype == T_BVBR) {
                            freeze = pt < 273.15f;
                        } else if (new_ctype == T_TUNG) {
                            freeze = pt < c_high;
                        } else if (new_ctype == T_CRMC) {
                            freeze = !(ctemph >= pres_sum(air_pv, x, y, d) + c_high);
                        } else if (c_high_tr == (int32_t)T_LAVA || new_ctype == T_HEAC) {
                            freeze = pt < c_high && new_ctype != T_BISM;
                        } else {
                            freeze = pt < 973.0f;
                        }
                        if (freeze) {
                            if (new_ctype == T_THRM) {
                                nt = T_BMTL;
                                new_tmp = 0;
                            } else if (new_ctype == T_PLUT) {
                                nt = T_LAVA;
                                new_tmp = 0;
                            } else {
                                nt = new_ctype;
                            }
                            new_ctype = 0u;
                            s = true;
                        }
                    } else if (pt < 973.0f) {
                        nt = T_STNE;
                        s = true;
                    }
                }
            }
            if (s) {
                if (nt == T_ICEI || nt == T_LAVA || nt == T_SNOW) new_ctype = old_ty;
                if (nt == T_RIME) new_ctype = T_DSTW;
                if (!(nt == T_ICEI && new_ctype == T_FRZW) && nt != T_ACID) new_life = 0;
                if (nt == T_FIRE) {
                    new_tmp = 0;
                    if (old_ty == T_SEED) {
                        new_ctype = 0u;
                        new_tmp2 = 0;
                    }
                }
                if ((EP(props, nt, COL_PROPERTIES) & 0x08u) && !(EP(props, old_ty, COL_PROPERTIES) & 0x08u))
                    pv += 0.50f;
                if (nt == 0u) {
                    p.ty[i] = 0u;
                    t = 0u;
                } else {
                    p.ty[i] = nt;
                    t = nt;
                    if (nt == T_FIRE || nt == T_PLSM || nt == T_CFLM) new_life = 120 + (int32_t)(roll % 50u);
                    if (nt == T_LAVA) {
                        if (new_ctype == T_BRMT) new_ctype = T_BMTL;
                        else if (new_ctype == T_SAND || new_ctype == T_BGLA || (new_ctype == T_LITH && new_tmp2 > 3))
                            new_ctype = T_GLAS;
                        else if (new_ctype == T_PQRT) new_ctype = T_QRTZ;
                        new_life = 240 + (int32_t)(roll % 120u);
                    }
                }
                p.ctype[i] = new_ctype;
                p.life[i] = new_life;
                p.tmp[i] = new_tmp;
                p.tmp2[i] = new_tmp2;
            }
            float pt2 = tpt_clampf(pt, 0.0f, 9999.0f);
            if (t == T_LAVA) {
                p.life[i] = min(max(tpt_f2i((pt2 - 700.0f) / 7.0f), 0), 400);
                if (p.ctype[i] == T_THRM && p.tmp[i] > 0) {
                    p.tmp[i] -= 1;
                    pt2 = 3500.0f;
                }
                if (p.ctype[i] == T_PLUT && p.tmp[i] > 0) {
                    p.tmp[i] -= 1;
                    pt2 = 9999.0f;
                }
                p.temp[i] = pt2;
            }
        } else {
            p.temp[i] = tpt_clampf(p.temp[i], 0.0f, 9999.0f);
        }
    }
    if (t != 0u && t < 256u) {
        if (t == T_LIFE) p.temp[i] = tpt_clampf(p.temp[i] - 50.0f, 0.0f, 9999.0f);
        const uint32_t tprops = EP(props, t, COL_PROPERTIES);
        if ((tprops & 0x20u) || t == T_SPRK) {
            int nx = x % TPT_CELL, ny = y % TPT_CELL;
            nx = nx == 0 ? x / TPT_CELL - 1 : (nx == TPT_CELL - 1 ? x / TPT_CELL + 1 : x / TPT_CELL);
            ny = ny == 0 ? y / TPT_CELL - 1 : (ny == TPT_CELL - 1 ? y / TPT_CELL + 1 : y / TPT_CELL);
            if (nx >= 0 && ny >= 0 && nx < d.xc && ny < d.yc) {
                const int wi = ny * d.xc + nx;
                if (t != T_SPRK) {
                    if ((wall_emap[wi] & 0xFFu) == 12u && p.life[i] == 0 && (wall_bmap[wi] & 0xFFu) != 18u) {
                        p.ty[i] = T_SPRK;
                        p.life[i] = 4;
                        p.ctype[i] = t;
                        t = T_SPRK;
                    }
                } else {
                    const uint32_t wall = wall_bmap[wi] & 0xFFu;
                    // WL_DETECT, WL_EWALL, WL_ALLOWLIQUID, WL_WALLELEC, WL_ALLOWALLELEC, WL_EHOLE
                    if (wall == 3u || wall == 2u || wall == 6u || wall == 1u || wall == 11u || wall == 12u)
                        spark[wi] = 1u;
                }
            }
        }
        const int32_t explosive = (int32_t)EP(props, t, COL_EXPLOSIVE);
        if ((explosive & 2) && pv > 2.5f) {
            const float fire_temp = EPF(props, T_FIRE, COL_DEFAULT_TEMP);
            const int32_t flam = (int32_t)EP(props, t, COL_FLAMMABLE);
            p.temp[i] = tpt_clampf(fire_temp + (float)(flam / 2), 0.0f, 9999.0f);
            p.life[i] = 180 + (int32_t)(roll % 80u);
            p.ty[i] = T_FIRE;
            t = T_FIRE;
            pv += 0.25f;
        }
        bool s = false;
        uint32_t nt = t;
        const float gx = __uint_as_float(cr.x), gy = __uint_as_float(cr.y);
        const float gravtot = (gx < 0.0f ? 0.0f - gx : gx) + (gy < 0.0f ? 0.0f - gy : gy);
        const int32_t hpt = (int32_t)EP(props, t, COL_HIGH_PRESSURE_TRANSITION);
        const float hpv = EPF(props, t, COL_HIGH_PRESSURE);
        const int32_t lpt = (int32_t)EP(props, t, COL_LOW_PRESSURE_TRANSITION);
        const float lpv = EPF(props, t, COL_LOW_PRESSURE);
        const int32_t tmp = p.tmp[i];
        if (hpt != NT && pv > hpv) {
            if (hpt != ST) {
                nt = (uint32_t)hpt;
                s = true;
            } else if (t == T_BMTL && (pv > 2.5f || (pv > 1.0f && tmp == 1))) {
                nt = T_BRMT;
                s = true;
            }
        } else if (lpt != NT && pv < lpv && gravtot <= lpv / 4.0f) {
            if (lpt != ST) {
                nt = (uint32_t)lpt;
                s = true;
            }
        } else if (hpt != NT && gravtot > hpv / 4.0f) {
            if (hpt != ST) {
                nt = (uint32_t)hpt;
                s = true;
            } else if (t == T_BMTL && (gravtot > 0.625f || (gravtot > 0.25f && tmp == 1))) {
                nt = T_BRMT;
                s = true;
            }
        }
        if (s) {
            if (nt == 0u) {
                p.ty[i] = 0u;
            } else {
                p.life[i] = 0;
                if (nt == T_BRMT) {
                    p.ctype[i] = 0u;
                    p.tmp[i] = 0;
                    p.tmp2[i] = 0;
                }
                p.ty[i] = nt;
                if (nt == T_FIRE) p.life[i] = 120 + (int32_t)(roll % 50u);
            }
        }
    }
    if (pv != pv0) {
        pv_new[i] = pv;
        atomicMax(&pv_key[ci], i + 1u);
    }
    // refresh_pmap_types for this slot, when the maps were built from the
    // types this pass started with: only a changed type can be stale, and
    // nothing in this kernel reads pmap.
    if (pmap_refresh) {
        const uint32_t tf = p.ty[i];
        if (tf != t_entry && tf != 0u) {
            const int idx = y * d.xres + x;
            const uint32_t packed = pmap_refresh[idx];
            if ((packed >> 9) == i + 1u && (packed & 0x1FFu) != tf) pmap_refresh[idx] = ((i + 1u) << 9) | (tf & 0x1FFu);
        }
    }
}

__global__ void th_transition_cells(float *__restrict__ air_pv, uint32_t *__restrict__ wall_emap,
                                    const float *__restrict__ pv_new, uint32_t *__restrict__ pv_key,
                                    uint32_t *__restrict__ spark, uint32_t cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const uint32_t key = pv_key[c];
    if (key) {
        air_pv[c] = pv_new[key - 1u];
        pv_key[c] = 0u;
    }
    if (spark[c]) {
        wall_emap[c] = 16u;
        spark[c] = 0u;
    }
}

// refresh_pmap_types_kernel (thermal_kernel.rs:1338-1387), pmap half.
__global__ __launch_bounds__(BLK) void th_refresh(const uint32_t *__restrict__ ty_a, const float *__restrict__ px,
                                                  const float *__restrict__ py, uint32_t *__restrict__ pmap, Dims d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.slots) return;
    const uint32_t ty = ty_a[i];
    if (ty == 0u || ty >= 256u) return;
    const int x = tpt_f2i(px[i]), y = tpt_f2i(py[i]);
    if (x < 0 || x >= d.xres || y < 0 || y >= d.yres) return;
    const int idx = y * d.xres + x;
    const uint32_t packed = pmap[idx];
    if ((packed >> 9) == i + 1u && (packed & 0x1FFu) != ty) pmap[idx] = ((i + 1u) << 9) | (ty & 0x1FFu);
}

}  // namespace

int tpt_thermal_step(tpt_ctx *ctx, int legacy) {
    const uint32_t *seed = ctx->counters + TPT_CTR_SEED;
    Parts p{tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE),
            tpt_buf<uint32_t>(ctx, TPT_BUF_PART_CTYPE), tpt_buf<float>(ctx, TPT_BUF_PART_X),
            tpt_buf<float>(ctx, TPT_BUF_PART_Y), tpt_buf<float>(ctx, TPT_BUF_PART_TEMP),
            tpt_buf<float>(ctx, TPT_BUF_PART_EQUIL), tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP),
            tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP2)};
    const Dims d{(int)ctx->cfg.xres, (int)ctx->cfg.yres, (int)ctx->xcells, (int)ctx->ycells, ctx->slot_bound};
    const dim3 sg((d.slots + BLK - 1) / BLK), cg((ctx->cells + BLK - 1) / BLK), b(BLK);
    const uint32_t *pmap = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP), *props = tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS);
    const float *gfx = tpt_buf<float>(ctx, TPT_BUF_GRAV_FX), *gfy = tpt_buf<float>(ctx, TPT_BUF_GRAV_FY);
    const bool fresh = ctx->maps_fresh;
    const uint32_t pixel_blocks = (ctx->pixels + BLK - 1) / BLK;
    ctx->maps_fresh = ctx->rebuild_fresh = false;
    if (!legacy && fresh) {
        // Pixel-major path over the records of the RecalcFreeParticles that
        // ran just before; the off-map list first, so its cell writes and
        // equilibria are in place for the tiles.
        const dim3 lg(256), tb(TW, 4), tg((d.xres + TW - 1) / TW, (d.yres + TH - 1) / TH);
        const unsigned long long *rec = (const unsigned long long *)ctx->pix_rec;
        float *air_hv = tpt_buf<float>(ctx, TPT_BUF_AIR_HV);
        const uint32_t aheat = ctx->cfg.aheat_enable ? 1u : 0u;
        tpt_pass_begin(ctx, PASS_CONDUCT);
        hipLaunchKernelGGL(th_conduct_list, lg, b, 0, ctx->stream, p, pmap, air_hv, props, ctx->slot_scratch_f,
                           ctx->cell_scratch_u, ctx->cell_scratch_u2, ctx->offmap, ctx->counters, d, aheat, seed,
                           (const uint32_t *)gfx, (const uint32_t *)gfy, tpt_buf<uint32_t>(ctx, TPT_BUF_AIR_PV),
                           ctx->cell_rec, ctx->cells);
        hipLaunchKernelGGL(th_conduct_px, tg, tb, 0, ctx->stream, p, rec, air_hv,
                           tpt_buf<uint32_t>(ctx, TPT_BUF_BLOCK_AIRH), props, ctx->slot_scratch_f,
                           ctx->cell_scratch_u, ctx->cell_scratch_u2, ctx->pix_heat, d, aheat, seed);
        tpt_pass_end(ctx, PASS_CONDUCT);
        tpt_pass_begin(ctx, PASS_CONDUCT_APPLY);
        // The tiles, rows of blocks for the off-map list (about 256 blocks)
        // and rows that reset the convection words (about 1024 blocks).
        const int list_rows = (int)((256u + tg.x - 1u) / tg.x), fill_rows = (int)((1024u + tg.x - 1u) / tg.x);
        const dim3 ag(tg.x, tg.y + (uint32_t)(list_rows + fill_rows));
        hipLaunchKernelGGL(th_apply_px, ag, tb, 0, ctx->stream, p, rec, (uint32_t *)ctx->pix_rec,
                           (const uint2 *)ctx->pix_heat, props, d, pmap, ctx->offmap, ctx->counters,
                           (unsigned long long *)ctx->pix_rec, (int)tg.y, list_rows, ctx->slot_scratch_u,
                           ctx->slot_scratch_u2);
        tpt_pass_end(ctx, PASS_CONDUCT_APPLY);
    } else if (!legacy) {
        tpt_pass_begin(ctx, PASS_CONDUCT);
        hipLaunchKernelGGL(th_conduct, sg, b, 0, ctx->stream, p, pmap, tpt_buf<float>(ctx, TPT_BUF_AIR_HV), props,
                           ctx->slot_scratch_f, ctx->cell_scratch_u, ctx->cell_scratch_u2, d,
                           ctx->cfg.aheat_enable ? 1u : 0u, seed);
        tpt_pass_end(ctx, PASS_CONDUCT);
        tpt_pass_begin(ctx, PASS_CONDUCT_CELL);
        hipLaunchKernelGGL(th_conduct_cells, cg, b, 0, ctx->stream, tpt_buf<float>(ctx, TPT_BUF_AIR_HV),
                           tpt_buf<uint32_t>(ctx, TPT_BUF_BLOCK_AIRH), ctx->slot_scratch_f, ctx->cell_scratch_u,
                           ctx->cell_scratch_u2, ctx->cells);
        tpt_pass_end(ctx, PASS_CONDUCT_CELL);
        tpt_pass_begin(ctx, PASS_CONDUCT_APPLY);
        hipLaunchKernelGGL(th_apply, sg, b, 0, ctx->stream, p, pmap, props, d);
        tpt_pass_end(ctx, PASS_CONDUCT_APPLY);
    }
    tpt_pass_begin(ctx, PASS_TRANSITION);
    if (legacy || !fresh)  // the fast path's th_conduct_list wrote the records
        hipLaunchKernelGGL(th_cell_pack, cg, b, 0, ctx->stream, (const uint32_t *)gfx, (const uint32_t *)gfy,
                           tpt_buf<uint32_t>(ctx, TPT_BUF_AIR_PV), ctx->cell_rec, ctx->cells);
    if (!legacy) {
        const GravCfg g{ctx->cfg.gravity_mode, ctx->cfg.custom_gravity_x, ctx->cfg.custom_gravity_y};
        if (fresh)
            hipLaunchKernelGGL(th_convect_propose_px, dim3(pixel_blocks + 256u), b, 0, ctx->stream, p,
                               (const unsigned long long *)ctx->pix_rec, props, ctx->cell_rec, ctx->offmap,
                               ctx->counters, ctx->slot_scratch_u, d, g, seed, pixel_blocks);
        else
            hipLaunchKernelGGL(th_convect_propose, sg, b, 0, ctx->stream, p, pmap, props, gfx, gfy,
                               ctx->slot_scratch_u, ctx->slot_scratch_u2, d, g, seed);
        hipLaunchKernelGGL(th_convect_touch, sg, b, 0, ctx->stream, ctx->slot_scratch_u, ctx->slot_scratch_u2,
                           d.slots);
        hipLaunchKernelGGL(th_convect_swap, sg, b, 0, ctx->stream, ctx->slot_scratch_u, ctx->slot_scratch_u2, p.temp,
                           d.slots);
    }
    hipLaunchKernelGGL(th_transition, sg, b, 0, ctx->stream, p, tpt_buf<float>(ctx, TPT_BUF_AIR_PV), ctx->cell_rec,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP), tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP), props,
                       ctx->slot_scratch_f, ctx->cell_scratch_u, ctx->cell_scratch_u2,
                       fresh ? tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP) : nullptr, d, legacy ? 1u : 0u, seed);
    tpt_pass_end(ctx, PASS_TRANSITION);
    tpt_pass_begin(ctx, PASS_TRANSITION_CELL);
    hipLaunchKernelGGL(th_transition_cells, cg, b, 0, ctx->stream, tpt_buf<float>(ctx, TPT_BUF_AIR_PV),
                       tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP), ctx->slot_scratch_f, ctx->cell_scratch_u,
                       ctx->cell_scratch_u2, ctx->cells);
    tpt_pass_end(ctx, PASS_TRANSITION_CELL);
    if (!fresh) {
        tpt_pass_begin(ctx, PASS_REFRESH_PMAP);
        hipLaunchKernelGGL(th_refresh, sg, b, 0, ctx->stream, p.ty, p.x, p.y, tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP),
                           d);
        tpt_pass_end(ctx, PASS_REFRESH_PMAP);
    }
    return tpt_check_launch("thermal pass");
}
