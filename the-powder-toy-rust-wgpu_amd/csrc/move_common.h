// Definitions shared by the movement pass (move.hip) and the water
// equalisation pass (water.hip).  The two passes live in separate
// translation units: with the water kernels in move.hip, the intent
// kernel's registers were allocated differently (its cell record spilled to
// scratch and was reloaded on every probe path, +20 us at 2048x2048).
#pragma once

#include "tpt_internal.h"

namespace tpt_move {

constexpr uint32_t LUT_CAN = 256u * 12u;

struct MoveEnv {
    const uint32_t *__restrict__ pmap;
    // pmap's type bits as bytes (types are below 256), built just before
    // the intent kernel: the planner's probes mostly need only the
    // occupant's type, and the byte map is a quarter of pmap's footprint.
    const uint8_t *__restrict__ tmap;
    // Per-cell record {air_vx bits, air_vy bits, wall_bmap, wall_emap},
    // packed once per movement pass: a particle's own cell and its probes'
    // cells cost one random line instead of one per array.
    const uint4 *__restrict__ cellrec;
    const uint32_t *__restrict__ lut;
    int xr, yr, xc;
    uint32_t edge;
    uint32_t gravity_mode;
    float cgx, cgy;
    uint32_t newton_grav;
    uint32_t *__restrict__ saw_list;
    uint32_t *__restrict__ saw_count;
    uint32_t *__restrict__ ctype;  // photon wavelength masks, narrowed by refraction
    uint32_t *__restrict__ pair_list;   // movers the e == 2 pair pass examines
    uint32_t *__restrict__ counters;
};

__device__ __forceinline__ bool wall_blocks_particle(uint32_t wall, uint32_t emap, uint32_t props) {
    if (wall == 8u || wall == 1u || wall == 9u) return true;
    if (wall == 6u) return (props & 0x2u) == 0u;
    if (wall == 10u) return (props & 0x1u) == 0u;
    if (wall == 13u) return (props & 0x8u) == 0u;
    if (wall == 15u) return (props & 0x10u) == 0u;
    if (wall == 2u) return (emap & 0xFFu) == 0u;
    if (wall == 3u) return (props & 0x4u) != 0u;
    return false;
}

__device__ __forceinline__ int wcell(const MoveEnv &e, int x, int y) { return (y / TPT_CELL) * e.xc + x / TPT_CELL; }



// The decision of eval_move_at (mod.rs:838-891) for an occupant type the
// caller supplies.
__device__ __forceinline__ uint32_t eval_move_ty(const MoveEnv &e, uint32_t props, uint32_t ty, int nx, int ny, uint32_t d_ty) {
    const int wc = wcell(e, nx, ny);
    const uint32_t wall = e.cellrec[wc].z, em = e.cellrec[wc].w;
    uint32_t can = e.lut[LUT_CAN + ty * 256u + d_ty];
    if (can == 3u) {
        if (d_ty == 23u || d_ty == 75u) can = 1u;
        else if (d_ty == 58u || d_ty == 43u || d_ty == 124u || d_ty == 60u) can = 2u;
        else can = 1u;
    }
    uint32_t result = can;
    if (wall != 0u) {
        if (wall_blocks_particle(wall, em, props)) {
            result = 0u;
        } else if (wall == 12u && (em & 0xFFu) == 0u) {
            const uint32_t dest_props = e.lut[d_ty * 12u];
            if ((props & 0x4u) == 0u && (dest_props & 0x4u) == 0u) result = 2u;
        }
    }
    return result;
}

struct Slots {
    uint32_t *ty;
    float *x, *y, *vx, *vy;
    uint32_t *flags;
    int32_t *tmp;
    uint32_t *intent;
};

// hash_u32 (reaction_kernel/shared.rs:1015-1023): MurmurHash3's fmix32.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 2246822507u;
    h ^= h >> 13;
    h *= 3266489909u;
    h ^= h >> 16;
    return h;
}

}  // namespace tpt_move
