// The part of reaction_kernel::step_reactions_gpu
// (crates/tpt-sim/src/reaction_kernel/mod.rs:1123-1237) this library runs,
// in the driver's order:
//  * the modext clear of the gravity mass-override channel
//    (reaction_kernel/modext/mod.rs:53-99);
//  * the life pass (life_and_local_reaction_kernel, life.rs:47-205): fire,
//    plasma and TRON burn-in lifetimes, fire burning out to smoke or steam,
//    plasma to noble gas or steam, sparks settling back into their
//    conductor, the element kill at zero life, and the stasis flag.  Each
//    slot reads and writes only its own state;
//  * the combustion pass's FIRE/PLSM heating and 5x5 ignition with the
//    COAL/BCOL timer, and its WTRV branch (combustion_reaction_kernel,
//    combustion/fire.rs:26-408, 983-1013);
//  * the fluid pass's WATR, DSTW, SLTW, ICEI and SNOW branches
//    (fluid_reaction_kernel, fluid.rs:149-400).
//
// The two pixel-parallel passes race in the reference: a burner heats and
// ignites neighbours that other burners read, and a water kills a fire that
// another water reads.  They run here under one canonical rule, the one the
// oracle restates as a serial loop over pixels in ascending order:
//  * an acting particle decides from the state its pass started with (a
//    burner's own temperature, which pixels hold burners, the neighbour
//    types a steam or water rule matches, the cell pressure of the ignition
//    roll);
//  * the particle acted upon changes in the acting pixels' ascending order,
//    each action seeing the earlier ones (heat composes; a target keeps the
//    first conversion: a later rule applies only while the target still has
//    the type the rule matched);
//  * a particle's own update happens at its own pixel's turn.
// On the GPU each pass is a gather: one thread per pixel walks its window
// in ascending pixel order and composes what the actors around it do to
// its particle (replaying an actor's neighbour loop where an earlier
// neighbour can stop it).  A decide kernel reads only the pass's starting
// state, parks type and temperature changes in rx_side and lists the
// pixels it changed; an apply kernel writes the listed changes.  The life
// kernel also records, per cell, which of the rarer classes are present
// (fire, salt, rock, ...; not the common water and steam), so that a
// 64x4-pixel tile far from them returns after reading 54 flag words, and,
// per pixel, each burner owner's temperature and free neighbours, so that
// the combustion gather finds burners without random reads.  The other
// tiles stage the pixels within 2 of them in LDS (burner records for the
// combustion pass, pmap words and types for the fluid pass), which every
// window of the tile then reads.

#include "tpt_internal.h"


namespace {

constexpr int BLK = 256;
constexpr uint32_t T_WATR = 2, T_FIRE = 4, T_STNE = 5, T_METL = 11, T_SPRK = 12, T_SNOW = 13, T_ICEI = 14,
                   T_SMKE = 16, T_SAND = 17, T_SALT = 18, T_SLTW = 19, T_WTRV = 21, T_BCOL = 27, T_PLSM = 28,
                   T_DSTW = 35, T_PLNT = 45, T_PSCN = 50, T_NSCN = 51, T_ETRD = 55, T_SWCH = 60, T_COAL = 62,
                   T_INWR = 63, T_SPNG = 78, T_FRZZ = 83, T_FRZW = 84, T_RBDM = 93, T_LRBD = 94, T_NBLE = 97,
                   T_HFLM = 100, T_TRON = 151, T_ROCK = 189, T_RSST = 191, T_BASE = 193;
constexpr uint32_t PROP_LIFE_KILL = 0x8000u, FLAG_STASIS = 0x20u;
// GpuElementProps words the element passes read.
constexpr uint32_t W_FLAMMABLE = 12, W_EXPLOSIVE = 13, W_DEFAULT_TEMP = 16;
constexpr float STEAM_IGNITE_K = 285.15f, STEAM_COOL_K = 1273.0f;
// fluid.rs:99-110
constexpr uint32_t SALT_WATR = 0x2545F491u, SALT_DSTW = 0x9E3779B9u, SALT_SLTW = 0x85EBCA6Bu,
                   SALT_ICE = 0xC2B2AE35u, ROLL_DERIVE = 0x517CC1B7u;

// Per-cell class flags (rx_flags), for the classes the element rules need
// besides the common water and steam (SMKE and PLNT need none: they change
// only beside a DSTW or an SLTW, which are flagged).
constexpr uint32_t F_BURN = 1u << 0,      // FIRE or PLSM outside stasis
    F_WTRV_HOT = 1u << 1,                 // WTRV with ctype FIRE
    F_RBDM = 1u << 3,                     // RBDM or LRBD
    F_DSTW = 1u << 5, F_SLTW = 1u << 6, F_ICE = 1u << 7,  // ICEI or SNOW
    F_ICE_FRZW = 1u << 8,                 // ICEI or SNOW with ctype FRZW
    F_SALT = 1u << 9, F_FIRE = 1u << 10, F_ROCK = 1u << 11, F_FRZZ = 1u << 14;
constexpr uint32_t STAMP_MASK = 0x0FFFFFFFu;
// rx_side.x
constexpr uint32_t SIDE_TYPE = 0x80000000u, SIDE_TEMP = 0x40000000u;

// hash_u32 (reaction_kernel/shared.rs:1015-1023), MurmurHash3's finaliser.
__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    uint32_t h = x;
    h ^= h >> 16;
    h *= 2246822507u;
    h ^= h >> 13;
    h *= 3266489909u;
    h ^= h >> 16;
    return h;
}

// chance (combustion/mod.rs:280-286).
__device__ __forceinline__ bool chance(uint32_t seed, uint32_t num, uint32_t den) {
    return den != 0u && hash_u32(seed) % den < num;
}

__device__ __forceinline__ uint32_t slot_of(uint32_t v) {
    const uint32_t raw = v >> 9;
    return raw ? raw - 1u : 0u;
}

// The class flags of a particle of type ty; ctype is read only where a
// class depends on it.
__device__ __forceinline__ uint32_t class_flags(uint32_t ty, const uint32_t *ctype, uint32_t i, bool frozen) {
    switch (ty) {
    case T_FIRE: return F_FIRE | (frozen ? 0u : F_BURN);
    case T_PLSM: return frozen ? 0u : F_BURN;
    case T_WTRV: return ctype[i] == T_FIRE ? F_WTRV_HOT : 0u;
    case T_RBDM:
    case T_LRBD: return F_RBDM;
    case T_DSTW: return F_DSTW;
    case T_SLTW: return F_SLTW;
    case T_ICEI:
    case T_SNOW: return F_ICE | (ctype[i] == T_FRZW ? F_ICE_FRZW : 0u);
    case T_SALT: return F_SALT;
    case T_ROCK: return F_ROCK;
    case T_FRZZ: return F_FRZZ;
    default: return 0u;
    }
}

// Clears the class flags and advances the burner-record stamp (never 0,
// which marks a pixel never written).
__global__ void rx_flags_clear(uint32_t *__restrict__ flags, uint32_t cells, uint32_t *__restrict__ counters) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells) flags[c] = 0u;
    if (c == 0u) {
        const uint32_t s = counters[TPT_CTR_RX_STAMP] % STAMP_MASK + 1u;
        counters[TPT_CTR_RX_STAMP] = s;
    }
}

struct LifeArgs {
    uint32_t *ty;
    int32_t *life, *tmp;
    uint32_t *ctype, *flags;
    float *temp;
    const float *x, *y;
    const uint32_t *wall, *emap, *props, *counters;
    uint32_t *cell_flags;
    const uint32_t *pmap;
    uint2 *burner;
    uint32_t slots;
    int xres, yres, xc, yc;
    int stasis;  // 0: no WL_STASIS cell on the canvas, so no particle is frozen
};

// A particle's pixel (rounded position) and cell; -1 off the canvas.
__device__ __forceinline__ int life_cell(const LifeArgs &a, uint32_t i, int *xi, int *yi) {
    *xi = tpt_f2i(a.x[i] + 0.5f);
    *yi = tpt_f2i(a.y[i] + 0.5f);
    if (*xi >= 0 && *yi >= 0 && *xi < a.xres && *yi < a.yres) return (*yi / TPT_CELL) * a.xc + *xi / TPT_CELL;
    return -1;
}

// The types whose life this pass reads besides the PROP_LIFE_KILL ones.
__device__ __forceinline__ bool life_type(uint32_t ty) {
    return ty == T_FIRE || ty == T_PLSM || ty == T_HFLM || ty == T_TRON || ty == T_SPRK;
}

__global__ __launch_bounds__(BLK) void rx_life(LifeArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.slots) return;
    const uint32_t ty = a.ty[i];
    if (ty == 0u) return;
    const uint32_t seed = a.counters[TPT_CTR_SEED];
    // Position and wall are read only where they matter: the stasis test
    // (canvases with a WL_STASIS cell) and the flagged classes below.
    int xi = 0, yi = 0, cell = -1;
    bool frozen = false, located = false;
    if (a.stasis) {
        cell = life_cell(a, i, &xi, &yi);
        located = true;
        if (cell >= 0) frozen = a.wall[cell] == 18u && (a.emap[cell] & 0xFFu) < 8u;
    }
    const uint32_t flags = a.flags[i];
    if (frozen) {
        if (!(flags & FLAG_STASIS)) a.flags[i] = flags | FLAG_STASIS;
        const uint32_t f = class_flags(ty, a.ctype, i, true);
        if (f) atomicOr(a.cell_flags + cell, f);
        return;
    }
    if (flags & FLAG_STASIS) a.flags[i] = flags & ~FLAG_STASIS;
    const uint32_t props = ty < 256u ? a.props[ty * 32u + 1u] : 0u;
    if (!life_type(ty) && !(props & PROP_LIFE_KILL)) {
        // no life rule: only the class flags (life and type unchanged)
        const uint32_t f = class_flags(ty, a.ctype, i, false);
        if (!f) return;
        if (!located) cell = life_cell(a, i, &xi, &yi);
        if (cell >= 0) atomicOr(a.cell_flags + cell, f);
        return;
    }
    const int32_t life0 = a.life[i];
    int32_t life = life0;
    uint32_t new_ty = ty;
    bool dead = false;
    if (ty == T_FIRE && life == 0) life = 180 + (int32_t)(hash_u32(i ^ seed) % 80u);
    else if ((ty == T_PLSM || ty == T_HFLM) && life == 0) life = 50 + (int32_t)(hash_u32(i ^ seed ^ 0xA5A5u) % 150u);
    else if (ty == T_TRON && life == 0) life = 25 + (int32_t)(hash_u32(i ^ seed ^ 0x3C3Cu) % 25u);
    if (ty == T_FIRE) {
        if (life <= 1) {
            if ((a.tmp[i] & 3) == 3) {
                new_ty = T_WTRV;
                life = 0;
                a.ctype[i] = T_FIRE;
            } else if (a.temp[i] < 625.0f) {
                new_ty = T_SMKE;
                life = 250 + (int32_t)(hash_u32(i ^ seed ^ 0x5ACE123u) % 20u);
                a.tmp[i] = 0;
            } else {
                dead = true;
            }
        }
    } else if (ty == T_PLSM) {
        if (life <= 1) {
            const uint32_t ct = a.ctype[i];
            if (ct == T_NBLE) {
                new_ty = T_NBLE;
                life = 0;
                a.ctype[i] = 0u;
            } else if ((a.tmp[i] & 3) == 3) {
                new_ty = T_WTRV;
                life = 0;
                a.ctype[i] = T_FIRE;
            } else {
                dead = true;
            }
        }
    } else if (ty == T_SPRK) {
        if (life <= 0) {
            uint32_t ct = a.ctype[i];
            if (ct == T_WATR || ct == T_SLTW || ct == T_PSCN || ct == T_NSCN || ct == T_ETRD || ct == T_INWR)
                a.temp[i] = 293.15f;
            if (ct == 0u || ct >= 256u) ct = T_METL;
            a.ctype[i] = 0u;
            life = 4;
            if (ct == T_WATR) life = 64;
            else if (ct == T_SLTW) life = 54;
            else if (ct == T_SWCH) life = 14;
            else if (ct == T_RSST) dead = true;
            if (!dead) new_ty = ct;
        }
    }
    if (new_ty == ty && (props & PROP_LIFE_KILL) && life <= 0) dead = true;
    if (dead) {
        a.ty[i] = 0u;
        return;
    }
    if (new_ty != ty) a.ty[i] = new_ty;
    if (life != life0) a.life[i] = life;
    const uint32_t f = class_flags(new_ty, a.ctype, i, false);
    if (!f) return;
    if (!located) cell = life_cell(a, i, &xi, &yi);
    if (cell < 0) return;
    atomicOr(a.cell_flags + cell, f);
    // a burner that owns its pixel: its temperature and free neighbours
    if ((f & F_BURN) && (a.pmap[yi * a.xres + xi] >> 9) == i + 1u) {
        uint32_t surround = 0u;
        for (int k = 0; k < 9; ++k) {
            const int sx = xi + k % 3 - 1, sy = yi + k / 3 - 1;
            if (k == 4) continue;
            if (sx < 0 || sy < 0 || sx >= a.xres || sy >= a.yres || a.pmap[sy * a.xres + sx] == 0u) ++surround;
        }
        a.burner[yi * a.xres + xi] = make_uint2(a.counters[TPT_CTR_RX_STAMP] | (surround << 28), __float_as_uint(a.temp[i]));
    }
}

// The element passes' view of the canvas.
struct RxArgs {
    const uint32_t *pmap;
    uint32_t *ty;
    int32_t *life, *tmp;
    uint32_t *ctype;
    const uint32_t *pflags;
    float *temp;
    const float *vx, *vy;
    float *air_pv;
    const uint32_t *props;
    uint32_t *counters;
    uint32_t *cell_flags, *pvcount;
    uint2 *side;
    const uint2 *burner;
    uint32_t *list, *wcount;
    int xres, yres, xc, yc;
    int legacy;
};

__device__ __forceinline__ uint32_t pmap_at(const RxArgs &a, int x, int y) {
    return (x < 0 || y < 0 || x >= a.xres || y >= a.yres) ? 0u : a.pmap[y * a.xres + x];
}

__device__ __forceinline__ bool combustion_relevant(uint32_t around, uint32_t own) {
    return (around & (F_BURN | F_RBDM)) || (own & F_WTRV_HOT);
}

// Whether a fluid rule can change the pixel's particle: every rule pairs an
// actor with a neighbour within 1 pixel, and at least one of the two is of
// a flagged class.  WATR's partners (salt, rubidium, fire, salt water,
// rock) and the DSTW and SLTW actors are flagged; SMKE changes only beside
// a DSTW and PLNT only beside an SLTW, both flagged already; ice changes or
// converts frost only with salt, salt water or FRZZ around (or when it
// holds FRZW itself).
__device__ __forceinline__ bool fluid_relevant(uint32_t around, uint32_t own) {
    return (around & (F_SALT | F_RBDM | F_FIRE | F_SLTW | F_ROCK | F_DSTW)) ||
           ((around & F_FRZZ) && (around & F_ICE)) || (own & F_ICE_FRZW);
}

// The decide kernels work on 64x4-pixel tiles (one cell row, 16 cells) with
// 256 threads.  A tile first reads the class flags of the 18x3 cells around
// it into LDS: every pixel within 2 of the tile lies in one of them.  A
// tile none of whose pixels can be acted on returns there; the others stage
// what their windows read (pixels within 2 of the tile) in LDS, so that the
// windows of neighbouring threads, which overlap, are loaded once.
constexpr int TW = 64, TH = TPT_CELL, HALO = 2, SW = TW + 2 * HALO, SH = TH + 2 * HALO, STAGED = SW * SH;
constexpr int CFW = TW / TPT_CELL + 2;  // 18 cells x 3 rows
static_assert(TW * TH == BLK, "one thread per tile pixel");

struct TileGeom {
    int x0, y0;  // tile origin (px)
    int x, y;    // this thread's pixel
    int lx, ly;
    bool inb;
};

__device__ __forceinline__ TileGeom tile_geom(const RxArgs &a) {
    TileGeom g;
    const int tiles_x = (a.xres + TW - 1) / TW;
    g.x0 = (int)(blockIdx.x % (unsigned)tiles_x) * TW;
    g.y0 = (int)(blockIdx.x / (unsigned)tiles_x) * TH;
    g.lx = (int)(threadIdx.x % TW);
    g.ly = (int)(threadIdx.x / TW);
    g.x = g.x0 + g.lx;
    g.y = g.y0 + g.ly;
    g.inb = g.x < a.xres && g.y < a.yres;
    return g;
}

// Stages the class flags around the tile and returns, for this thread's
// pixel, the OR of the 3x3 cells around its cell (*own: its cell's flags).
__device__ __forceinline__ uint32_t tile_flags(const RxArgs &a, const TileGeom &g, uint32_t *s_cf, uint32_t *own) {
    const int t = (int)threadIdx.x;
    if (t < CFW * 3) {
        const int cx = g.x0 / TPT_CELL - 1 + t % CFW, cy = g.y0 / TPT_CELL - 1 + t / CFW;
        s_cf[t] = (cx >= 0 && cy >= 0 && cx < a.xc && cy < a.yc) ? a.cell_flags[cy * a.xc + cx] : 0u;
    }
    __syncthreads();
    const int c = g.lx / TPT_CELL;
    uint32_t f = 0u;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int d = 0; d < 3; ++d) f |= s_cf[r * CFW + c + d];
    *own = s_cf[CFW + c + 1];
    return f;
}

// Pixel (gx, gy) of the staged region, as an LDS index (gx, gy within 2 of
// the tile).
__device__ __forceinline__ int st_idx(const TileGeom &g, int gx, int gy) {
    return (gy - g.y0 + HALO) * SW + (gx - g.x0 + HALO);
}

// Records `q` for each lane that wants to in this wave's own segment of the
// changed-pixel list (64 entries at wave slot 4 * tile + wave) and the
// segment's length.  No atomics: a same-address counter shared by every
// wave of the grid serialises them.  A wave that records nothing leaves its
// length at 0, where the apply kernel left it.
__device__ __forceinline__ void wave_record(uint32_t *list, uint32_t *wcount, bool want, uint32_t q) {
    const uint64_t m = __ballot(want);
    if (!m) return;
    const uint32_t slot = blockIdx.x * (BLK / 64) + threadIdx.x / 64;
    const int lane = (int)__lane_id();
    if (want) list[slot * 64u + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = q;
    if (lane == __ffsll((unsigned long long)m) - 1) wcount[slot] = (uint32_t)__popcll(m);
}

// The particle a gather thread composes.
struct Subject {
    uint32_t ty;
    int32_t life, tmp;
    uint32_t ctype;
    float temp;
};

// Combustion decide: the FIRE/PLSM burners of the 5x5 window act on the
// pixel's particle in ascending pixel order (fire.rs:143-408), the particle
// runs its own WTRV branch (fire.rs:983-1013) at its own pixel's turn.
// Burners come from this step's burner records, staged per tile.
__global__ __launch_bounds__(BLK) void rx_comb_decide(RxArgs a) {
    __shared__ uint32_t s_cf[CFW * 3];
    __shared__ uint2 s_br[STAGED];
    const TileGeom g = tile_geom(a);
    uint32_t own = 0u;
    const uint32_t around = tile_flags(a, g, s_cf, &own);
    const bool rel = g.inb && combustion_relevant(around, own);
    if (!__syncthreads_or(rel ? 1 : 0)) return;
    const uint32_t stamp = a.counters[TPT_CTR_RX_STAMP];
    for (int e = (int)threadIdx.x; e < STAGED; e += BLK) {
        const int gx = g.x0 - HALO + e % SW, gy = g.y0 - HALO + e / SW;
        s_br[e] = (gx >= 0 && gy >= 0 && gx < a.xres && gy < a.yres) ? a.burner[gy * a.xres + gx] : make_uint2(0u, 0u);
    }
    __syncthreads();
    if (!rel) return;
    const int x = g.x, y = g.y, q = y * a.xres + x;
    const uint32_t v = a.pmap[q];
    if (!v) return;
    // the burners of the window, as bits of the 25 positions
    uint32_t bmask = 0u;
    if (around & F_BURN) {
#pragma unroll
        for (int k = 0; k < 25; ++k) {
            const int bx = x + k % 5 - 2, by = y + k / 5 - 2;
            if (k == 12 || bx < 0 || by < 0 || bx >= a.xres || by >= a.yres) continue;
            if ((s_br[st_idx(g, bx, by)].x & STAMP_MASK) == stamp) bmask |= 1u << k;
        }
    }
    const uint32_t id = slot_of(v);
    Subject s;
    s.ty = a.ty[id];
    if (s.ty == 0u) return;
    const bool wtrv = s.ty == T_WTRV && ((own & F_WTRV_HOT) || (around & F_RBDM));
    if (!bmask && !wtrv) return;
    const uint32_t ty0 = s.ty;
    s.life = a.life[id];
    s.tmp = a.tmp[id];
    s.ctype = a.ctype[id];
    s.temp = a.temp[id];
    const Subject s0 = s;
    const bool frozen = (a.pflags[id] & FLAG_STASIS) != 0u;
    const uint32_t seed = a.counters[TPT_CTR_SEED];
    const float fire_t = __uint_as_float(a.props[T_FIRE * 32u + W_DEFAULT_TEMP]);
    int ign_cell = -1;
    for (int k = 0; k < 25; ++k) {
        const int ddx = k % 5 - 2, ddy = k / 5 - 2;
        if (k == 12) {
            // the particle's own update: the WTRV branch
            if (s.ty != T_WTRV || frozen) continue;
            const uint32_t prand = hash_u32(seed ^ ((uint32_t)q * 2654435761u));
            const float t = s.temp;
            bool ignited = false;
            if (around & F_RBDM) {
                for (uint32_t n = 0; n < 9u && !ignited; ++n) {
                    const int dx = (int)(n % 3u) - 1, dy = (int)(n / 3u) - 1;
                    if (!dx && !dy) continue;
                    const uint32_t nv = pmap_at(a, x + dx, y + dy);
                    if (!nv) continue;
                    const uint32_t nty = a.ty[slot_of(nv)];
                    if ((nty == T_RBDM || nty == T_LRBD) && t > STEAM_IGNITE_K && chance(prand ^ n, 1u, 100u)) {
                        s.ty = T_FIRE;
                        s.life = 4;
                        s.ctype = T_WATR;
                        ignited = true;
                    }
                }
            }
            if (!ignited && t > STEAM_COOL_K && s.ctype == T_FIRE) s.temp = t - t / 1000.0f;
            continue;
        }
        if (!(bmask & (1u << k))) continue;
        const int bx = x + ddx, by = y + ddy;
        const uint2 br = s_br[st_idx(g, bx, by)];
        const float tb = __uint_as_float(br.y);
        const uint32_t surround = br.x >> 28;
        // this particle is the burner's neighbour n
        const uint32_t n = (uint32_t)((2 - ddy) * 5 + (2 - ddx));
        const uint32_t bq = (uint32_t)(by * a.xres + bx);
        const uint32_t prand = hash_u32(seed ^ (bq * 2654435761u));
        const uint32_t nroll = hash_u32(prand ^ (n * 0x9E3779B9u));
        if ((s.ty == T_COAL || s.ty == T_BCOL) && s.life > 100 && chance(nroll, 1u, 500u)) s.life = 99;
        if (tb > s.temp) {
            float h = s.temp + (tb - s.temp) * 0.05f;
            if (h > 9999.0f) h = 9999.0f;
            s.temp = h;
        }
        const int32_t nflam = s.ty < 256u ? (int32_t)a.props[s.ty * 32u + W_FLAMMABLE] : 0;
        const bool nexp = s.ty < 256u && (int32_t)a.props[s.ty * 32u + W_EXPLOSIVE] != 0;
        if (nflam <= 0) continue;
        bool ignitable = surround > 0u || nexp;
        if (s.ty == T_SPNG && s.life > 0) ignitable = false;
        if (!ignitable) continue;
        const float pv = a.air_pv[(y / TPT_CELL) * a.xc + x / TPT_CELL];
        const int32_t num = tpt_f2i((float)nflam + pv * 10.0f);
        bool hit = false;
        if (num >= 1000) hit = true;
        else if (num > 0) hit = chance(seed ^ (id * 2654435761u) ^ n, (uint32_t)num, 1000u);
        if (!hit) continue;
        s.ty = T_FIRE;
        s.life = 180 + (int32_t)(hash_u32(id ^ seed) % 80u);
        float ft = fire_t + (float)(nflam / 2);
        if (ft < 0.0f) ft = 0.0f;
        if (ft > 9999.0f) ft = 9999.0f;
        s.temp = ft;
        s.tmp = 0;
        s.ctype = 0u;
        if (nexp) ign_cell = (by / TPT_CELL) * a.xc + bx / TPT_CELL;
    }
    if (s.life != s0.life) a.life[id] = s.life;
    if (s.tmp != s0.tmp) a.tmp[id] = s.tmp;
    if (s.ctype != s0.ctype) a.ctype[id] = s.ctype;
    uint32_t side = 0u;
    if (s.ty != ty0) side |= SIDE_TYPE | s.ty;
    if (__float_as_uint(s.temp) != __float_as_uint(s0.temp)) side |= SIDE_TEMP;
    if (side) a.side[q] = make_uint2(side, __float_as_uint(s.temp));
    if (s.ty == T_FIRE && ty0 != T_FIRE) atomicOr(a.cell_flags + (y / TPT_CELL) * a.xc + x / TPT_CELL, F_FIRE);
    if (ign_cell >= 0) atomicAdd(a.pvcount + ign_cell, 1u);
    wave_record(a.list, a.wcount, side != 0u, (uint32_t)q);
}

// Applies the parked type/temperature writes of the pixels the decide
// kernel recorded (one block per tile, a lane per entry of its waves'
// segments) and empties the segments; with `with_pv` also adds each of the
// tile's cells' explosive ignitions to its pressure, 0.25 per ignition.
__global__ __launch_bounds__(BLK) void rx_apply(RxArgs a, int with_pv) {
    const uint32_t slot = blockIdx.x * (BLK / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t n = a.wcount[slot];
    if (lane < n) {
        const uint32_t q = a.list[slot * 64u + lane];
        const uint2 sd = a.side[q];
        const uint32_t id = slot_of(a.pmap[q]);
        if (sd.x & SIDE_TYPE) a.ty[id] = sd.x & 0x1FFu;
        if (sd.x & SIDE_TEMP) a.temp[id] = __uint_as_float(sd.y);
        a.side[q] = make_uint2(0u, 0u);
    }
    if (n && lane == 0) a.wcount[slot] = 0u;  // (after every lane of the wave has read it)
    if (!with_pv || threadIdx.x >= TW / TPT_CELL) return;
    const TileGeom g = tile_geom(a);
    const int cx = g.x0 / TPT_CELL + (int)threadIdx.x, cy = g.y0 / TPT_CELL;
    if (cx >= a.xc) return;
    const int c = cy * a.xc + cx;
    const uint32_t cnt = a.pvcount[c];
    if (!cnt) return;
    float pv = a.air_pv[c];
    for (uint32_t i = 0; i < cnt; ++i) pv += 0.25f;
    a.air_pv[c] = pv;
    a.pvcount[c] = 0u;
}

__device__ __forceinline__ bool is_actor(uint32_t ty) {
    return ty == T_WATR || ty == T_DSTW || ty == T_SLTW || ty == T_ICEI || ty == T_SNOW;
}

// One fluid actor's neighbour loop (fluid.rs:180-399) at pixel (px, py),
// from the pass's starting state.  With n_stop < 9 it stops before
// neighbour n_stop and reports, in *conv / *conv_ctype, what the rule does
// to that neighbour when the actor reaches it (*reached); with n_stop == 9
// it runs the whole loop and reports the actor's own results.  The window's
// pmap words and types come from the tile (s_pm, s_ty).
struct ActorOut {
    uint32_t new_ty;
    int32_t life;
    uint32_t ctype;
    float temp;
    bool reached;
    uint32_t conv, conv_ctype;  // conversion of neighbour n_stop (0xFFFFFFFF: none)
};

__device__ __forceinline__ ActorOut fluid_actor(const RxArgs &a, const TileGeom &g, const uint32_t *s_pm,
                                                const uint32_t *s_ty, uint32_t seed, int px, int py, uint32_t aid,
                                                uint32_t ty, uint32_t n_stop) {
    ActorOut o;
    o.new_ty = ty;
    o.life = a.life[aid];
    o.ctype = a.ctype[aid];
    o.temp = a.temp[aid];
    o.reached = false;
    o.conv = 0xFFFFFFFFu;
    o.conv_ctype = 0xFFFFFFFFu;
    const uint32_t idx = (uint32_t)(py * a.xres + px);
    const float speed = fabsf(a.vx[aid]) + fabsf(a.vy[aid]);
    const float own_t = o.temp;
    if ((ty == T_ICEI || ty == T_SNOW) && o.ctype == T_FRZW) {
        float t = o.temp - 1.0f;
        if (t < 0.0f) t = 0.0f;
        if (t > 9999.0f) t = 9999.0f;
        o.temp = t;
    }
    const uint32_t salt = ty == T_WATR ? SALT_WATR : ty == T_DSTW ? SALT_DSTW : ty == T_SLTW ? SALT_SLTW : SALT_ICE;
    // the window (up to n_stop); off-canvas entries are staged as empty
    const int c0 = st_idx(g, px, py);
    uint32_t wv[9], wt[9];
#pragma unroll
    for (int n = 0; n < 9; ++n) {
        const int e = c0 + (n / 3 - 1) * SW + (n % 3 - 1);
        wv[n] = (n == 4 || (uint32_t)n > n_stop) ? 0u : s_pm[e];
        wt[n] = wv[n] ? s_ty[e] : 0u;
    }
    bool stop = false;
#pragma unroll
    for (uint32_t n = 0; n < 9u; ++n) {
        if (n == 4u || !wv[n]) continue;
        if (stop) break;
        if (n == n_stop) o.reached = true;
        const uint32_t nid = slot_of(wv[n]), nty = wt[n];
        const uint32_t r = hash_u32(seed ^ (idx * salt) ^ (n * ROLL_DERIVE));
        uint32_t conv = 0xFFFFFFFFu, conv_ctype = 0xFFFFFFFFu;
        if (ty == T_WATR) {
            if (nty == T_SALT && r % 50u == 0u) {
                o.new_ty = T_SLTW;
                if (hash_u32(r ^ 0x51ED270Bu) % 3u == 0u) conv = T_SLTW;
            } else if ((nty == T_RBDM || nty == T_LRBD) && (a.legacy || own_t > 285.15f) && r % 100u == 0u) {
                o.new_ty = T_FIRE;
                o.life = 4;
                o.ctype = T_WATR;
            } else if (nty == T_FIRE && a.ctype[nid] != T_WATR) {
                conv = 0u;
                if (hash_u32(r ^ 0x2C1B3E4Du) % 30u == 0u) stop = true;
            } else if (nty == T_SLTW && r % 2000u == 0u) {
                o.new_ty = T_SLTW;
            } else if (nty == T_ROCK && speed >= 0.5f && r % 1000u == 0u) {
                conv = hash_u32(r ^ 0x7F4A7C15u) % 3u == 0u ? T_SAND : T_STNE;
            }
        } else if (ty == T_DSTW) {
            if (nty == T_SALT && r % 50u == 0u) {
                o.new_ty = T_SLTW;
                if (hash_u32(r ^ 0x51ED270Bu) % 3u == 0u) conv = T_SLTW;
            } else if (nty == T_SLTW) {
                if (r % 2000u == 0u) o.new_ty = T_SLTW;
                else if (hash_u32(r ^ 0x3AE91F27u) % 100u == 0u) o.new_ty = T_WATR;
            } else if (nty == T_WATR && r % 100u == 0u) {
                o.new_ty = T_WATR;
            } else if ((nty == T_RBDM || nty == T_LRBD) && (a.legacy || own_t > 12.0f) && r % 100u == 0u) {
                o.new_ty = T_FIRE;
                o.life = 4;
            } else if (nty == T_FIRE) {
                conv = 0u;
                if (hash_u32(r ^ 0x2C1B3E4Du) % 30u == 0u) stop = true;
            } else if (nty == T_SMKE) {
                const float nt = a.temp[nid];
                if (nt > 313.15f && nt < 333.15f && own_t > 313.15f && own_t < 333.15f && r % 100u == 0u) {
                    o.new_ty = T_BASE;
                    o.life = 1;
                    conv = 0u;
                }
            }
        } else if (ty == T_SLTW) {
            if (nty == T_SALT && r % 2000u == 0u) {
                conv = T_SLTW;
            } else if (nty == T_PLNT && r % 40u == 0u) {
                conv = 0u;
            } else if ((nty == T_RBDM || nty == T_LRBD) && (a.legacy || own_t > 285.15f) && r % 100u == 0u) {
                o.new_ty = T_FIRE;
                o.life = 4;
                o.ctype = T_WATR;
            } else if (nty == T_FIRE && a.ctype[nid] != T_WATR) {
                conv = 0u;
                if (hash_u32(r ^ 0x2C1B3E4Du) % 30u == 0u) stop = true;
            }
        } else {  // ICEI, SNOW
            const uint32_t div = ty == T_ICEI ? 200u : 333u;
            if ((nty == T_SALT || nty == T_SLTW) && (ty == T_SNOW || o.temp > 252.05f) && r % div == 0u) {
                o.new_ty = T_SLTW;
                conv = T_SLTW;
                stop = true;
            } else if (ty == T_ICEI && nty == T_FRZZ && r % 200u == 0u) {
                conv = T_ICEI;
                conv_ctype = T_FRZW;
            }
        }
        if (n == n_stop) {
            o.conv = conv;
            o.conv_ctype = conv_ctype;
            return o;
        }
    }
    // WATR, DSTW and SLTW end dead after a killing extinguish; ICEI and SNOW only stop
    if (stop && ty != T_ICEI && ty != T_SNOW) o.new_ty = 0u;
    return o;
}

// Whether an actor of type aty has a rule that converts a neighbour of
// type t (fluid.rs:180-399).
__device__ __forceinline__ bool converts(uint32_t aty, uint32_t t) {
    switch (t) {
    case T_FIRE: return aty == T_WATR || aty == T_DSTW || aty == T_SLTW;
    case T_SALT: return aty == T_WATR || aty == T_DSTW || aty == T_SLTW || aty == T_ICEI || aty == T_SNOW;
    case T_ROCK: return aty == T_WATR;
    case T_SMKE: return aty == T_DSTW;
    case T_PLNT: return aty == T_SLTW;
    case T_FRZZ: return aty == T_ICEI;
    default: return false;  // (ICEI and SNOW turning SLTW into SLTW change nothing)
    }
}

// Whether an actor of type ty, with `around` the flags within 2 pixels, can
// change itself this pass.
__device__ __forceinline__ bool actor_may_change(uint32_t ty, uint32_t around, uint32_t own) {
    switch (ty) {
    case T_WATR: return around & (F_SALT | F_RBDM | F_FIRE | F_SLTW | F_ROCK);
    case T_DSTW: return true;  // its WATR partner is not flagged
    case T_SLTW: return around & (F_SALT | F_RBDM | F_FIRE);  // (PLNT and SALT change only the neighbour)
    default: return (around & (F_SALT | F_SLTW | F_FRZZ)) || (own & F_ICE_FRZW);  // ICEI, SNOW
    }
}

// Fluid decide: the actors of the 3x3 window act on the pixel's particle in
// ascending pixel order; the particle runs its own branch at its own turn.
// After the first conversion of a target nothing later can change it.  The
// tile stages the pmap words and the types of every pixel within 2 of it:
// an actor's window, and the windows of the actors next to a target.
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(8))) void rx_fluid_decide(RxArgs a) {
    __shared__ uint32_t s_cf[CFW * 3];
    __shared__ uint32_t s_pm[STAGED], s_ty[STAGED];
    const TileGeom g = tile_geom(a);
    uint32_t own = 0u;
    const uint32_t around = tile_flags(a, g, s_cf, &own);
    const bool rel = g.inb && fluid_relevant(around, own);
    if (!__syncthreads_or(rel ? 1 : 0)) return;
    {
        uint32_t pm[(STAGED + BLK - 1) / BLK];
#pragma unroll
        for (int r = 0; r < (STAGED + BLK - 1) / BLK; ++r) {
            const int e = (int)threadIdx.x + r * BLK;
            const int gx = g.x0 - HALO + e % SW, gy = g.y0 - HALO + e / SW;
            pm[r] = (e < STAGED && gx >= 0 && gy >= 0 && gx < a.xres && gy < a.yres) ? a.pmap[gy * a.xres + gx] : 0u;
        }
#pragma unroll
        for (int r = 0; r < (STAGED + BLK - 1) / BLK; ++r) {
            const int e = (int)threadIdx.x + r * BLK;
            if (e < STAGED) {
                s_pm[e] = pm[r];
                s_ty[e] = pm[r] ? a.ty[slot_of(pm[r])] : 0u;
            }
        }
    }
    __syncthreads();
    if (!rel) return;
    const int x = g.x, y = g.y, q = y * a.xres + x;
    const int me = st_idx(g, x, y);
    const uint32_t v = s_pm[me];
    if (!v) return;
    const uint32_t id = slot_of(v), ty0 = s_ty[me];
    if (ty0 == 0u) return;
    const bool actor = is_actor(ty0) && actor_may_change(ty0, around, own) && !(a.pflags[id] & FLAG_STASIS);
    const bool target = ty0 == T_FIRE || ty0 == T_SALT || ty0 == T_ROCK || ty0 == T_SMKE || ty0 == T_PLNT ||
                        ty0 == T_FRZZ;
    if (!actor && !target) return;
    const uint32_t seed = a.counters[TPT_CTR_SEED];
    uint32_t ty = ty0;
    int32_t life = 0;
    uint32_t ctype = 0u;
    float temp = 0.0f;
    bool self_done = false, ctype_set = false;
    if (actor) {  // (actor and target types are disjoint)
        const ActorOut o = fluid_actor(a, g, s_pm, s_ty, seed, x, y, id, ty0, 9u);
        ty = o.new_ty;
        life = o.life;
        ctype = o.ctype;
        temp = o.temp;
        self_done = true;
    } else {
        // the neighbours that can convert this target, as bits of the window
        uint32_t cand = 0u;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (k == 4) continue;
            const int e = me + (k / 3 - 1) * SW + (k % 3 - 1);
            if (s_pm[e] && converts(s_ty[e], ty0)) cand |= 1u << k;
        }
        uint32_t c2 = cand;
        while (c2) {
            const int k = __ffs(c2) - 1;
            c2 &= c2 - 1u;
            const int e = me + (k / 3 - 1) * SW + (k % 3 - 1);
            if (a.pflags[slot_of(s_pm[e])] & FLAG_STASIS) cand &= ~(1u << k);
        }
        // in ascending pixel order, until the first conversion
        while (cand && ty == ty0) {
            const int k = __ffs(cand) - 1;
            cand &= cand - 1u;
            const int ddx = k % 3 - 1, ddy = k / 3 - 1;
            const int e = me + ddy * SW + ddx;
            const uint32_t aid = slot_of(s_pm[e]);
            const uint32_t n = (uint32_t)((1 - ddy) * 3 + (1 - ddx));
            const ActorOut o = fluid_actor(a, g, s_pm, s_ty, seed, x + ddx, y + ddy, aid, s_ty[e], n);
            if (!o.reached || o.conv == 0xFFFFFFFFu) continue;
            ty = o.conv;
            if (o.conv_ctype != 0xFFFFFFFFu) {
                ctype = o.conv_ctype;
                ctype_set = true;
            }
        }
    }
    if (self_done) {
        if (life != a.life[id]) a.life[id] = life;
        if (ctype != a.ctype[id]) a.ctype[id] = ctype;
    } else if (ctype_set) {
        a.ctype[id] = ctype;
    }
    uint32_t side = 0u;
    if (ty != ty0) side |= SIDE_TYPE | ty;
    if (self_done && __float_as_uint(temp) != __float_as_uint(a.temp[id])) side |= SIDE_TEMP;
    if (side) a.side[q] = make_uint2(side, __float_as_uint(temp));
    wave_record(a.list, a.wcount, side != 0u, (uint32_t)q);
}

RxArgs rx_args(tpt_ctx *ctx) {
    RxArgs a;
    a.pmap = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP);
    a.ty = tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY);
    a.life = tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE);
    a.tmp = tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP);
    a.ctype = tpt_buf<uint32_t>(ctx, TPT_BUF_PART_CTYPE);
    a.pflags = tpt_buf<uint32_t>(ctx, TPT_BUF_PART_FLAGS);
    a.temp = tpt_buf<float>(ctx, TPT_BUF_PART_TEMP);
    a.vx = tpt_buf<float>(ctx, TPT_BUF_PART_VX);
    a.vy = tpt_buf<float>(ctx, TPT_BUF_PART_VY);
    a.air_pv = tpt_buf<float>(ctx, TPT_BUF_AIR_PV);
    a.props = tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS);
    a.counters = ctx->counters;
    a.cell_flags = ctx->rx_flags;
    a.pvcount = ctx->rx_pvcount;
    a.side = ctx->rx_side;
    a.burner = ctx->rx_burner;
    a.list = ctx->rx_list;
    a.wcount = ctx->rx_wcount;
    a.xres = (int)ctx->cfg.xres;
    a.yres = (int)ctx->cfg.yres;
    a.xc = (int)ctx->xcells;
    a.yc = (int)ctx->ycells;
    a.legacy = ctx->cfg.heat_enable ? 0 : 1;
    return a;
}

}  // namespace

int tpt_reactions_step(tpt_ctx *ctx) {
    int rc = tpt_gravity_override_clear(ctx);
    if (rc) return rc;
    const uint32_t slots = ctx->slot_bound;
    hipLaunchKernelGGL(rx_flags_clear, dim3((ctx->cells + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream, ctx->rx_flags,
                       ctx->cells, ctx->counters);
    LifeArgs l{tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY),    tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE),
               tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP),    tpt_buf<uint32_t>(ctx, TPT_BUF_PART_CTYPE),
               tpt_buf<uint32_t>(ctx, TPT_BUF_PART_FLAGS), tpt_buf<float>(ctx, TPT_BUF_PART_TEMP),
               tpt_buf<float>(ctx, TPT_BUF_PART_X),        tpt_buf<float>(ctx, TPT_BUF_PART_Y),
               tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP),  tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP),
               tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS), ctx->counters,
               ctx->rx_flags,                              tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP),
               ctx->rx_burner,                             slots,
               (int)ctx->cfg.xres,                         (int)ctx->cfg.yres,
               (int)ctx->xcells,                           (int)ctx->ycells,
               ctx->has_stasis ? 1 : 0};
    tpt_pass_begin(ctx, PASS_REACTION_LIFE);
    hipLaunchKernelGGL(rx_life, dim3((slots + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream, l);
    tpt_pass_end(ctx, PASS_REACTION_LIFE);
    if ((rc = tpt_check_launch("rx_life"))) return rc;
    const RxArgs a = rx_args(ctx);
    const uint32_t tiles = ctx->rx_tiles;  // ((xres + 63) / 64) * (yres / 4)
    static_assert(TW == 64 && TH == 4, "rx_tiles counts 64x4 tiles");
    tpt_pass_begin(ctx, PASS_REACTION_COMBUSTION);
    hipLaunchKernelGGL(rx_comb_decide, dim3(tiles), dim3(BLK), 0, ctx->stream, a);
    hipLaunchKernelGGL(rx_apply, dim3(tiles), dim3(BLK), 0, ctx->stream, a, 1);
    tpt_pass_end(ctx, PASS_REACTION_COMBUSTION);
    if ((rc = tpt_check_launch("rx_combustion"))) return rc;
    tpt_pass_begin(ctx, PASS_REACTION_FLUID);
    hipLaunchKernelGGL(rx_fluid_decide, dim3(tiles), dim3(BLK), 0, ctx->stream, a);
    hipLaunchKernelGGL(rx_apply, dim3(tiles), dim3(BLK), 0, ctx->stream, a, 0);
    tpt_pass_end(ctx, PASS_REACTION_FLUID);
    // Positions are unchanged: the pixel records still cover every slot
    // (types and temperatures are read from the slot arrays downstream).
    return tpt_check_launch("rx_fluid");
}
