/*
 * Oracle: RecalcFreeParticles + CheckStacking -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates crates/tpt-sim/src/buffers/spatial.rs:81-275 (kernels) and
 * spatial.rs:371-472 (driver).  The scatter runs in ascending slot order,
 * which is the serial order the reference's last-writer rule describes
 * (canonical rule R7): the highest matter slot on a pixel owns pmap unless it
 * is INVIS/FILT, which only take an empty pixel.
 */
#include "tpt_oracle.h"

#include <string.h>

#include "oracle_util.h"

/* spatial.rs:62-69 (the same avalanche as thermal_kernel.rs:289-296) */
uint32_t oracle_lowbias_hash(uint32_t seed) {
    uint32_t h = seed ^ 0x9E3779B9u;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    return h ^ (h >> 16);
}

/* spatial.rs:101-180 */
static void recalc_scatter(oracle_state *s) {
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4;
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= 256) continue;
        int x = ora_f2i(s->part_x[i] + 0.5f), y = ora_f2i(s->part_y[i] + 0.5f);
        int in_bounds = x >= 0 && y >= 0 && x < (int)xres && y < (int)yres;
        int stasis = 0;
        if (in_bounds) {
            uint32_t cell = ((uint32_t)y / 4) * xc + (uint32_t)x / 4;
            if ((s->wall_bmap[cell] & 0xFFu) == 18u && (s->wall_emap[cell] & 0xFFu) < 8u) stasis = 1;
        }
        uint32_t props = s->elem_props[ty * 32 + 1];
        int32_t life = s->part_life[i];
        if (life > 0 && (props & 0x4000u) && (!in_bounds || !stasis)) {
            int32_t next = life - 1;
            s->part_life[i] = next;
            if (next <= 0 && (props & (0x10000u | 0x8000u))) {
                s->part_ty[i] = 0;
                ty = 0;
            }
        } else if (life <= 0 && (props & 0x8000u) && (!in_bounds || !stasis)) {
            s->part_ty[i] = 0;
            ty = 0;
        }
        if (ty == 0 || !in_bounds) continue;
        uint32_t idx = (uint32_t)y * xres + (uint32_t)x;
        uint32_t packed = ((i + 1) << 9) | (ty & 0x1FFu);
        if (props & 0x10u) {
            s->photons[idx] = packed;
        } else {
            if (s->pmap[idx] == 0 || (ty != 124u && ty != 134u)) s->pmap[idx] = packed;
            if (ty != 96u && ty != 153u && ty != 159u && ty != 28u) s->pmap_count[idx] += 1;
        }
    }
}

/* spatial.rs:190-216 */
static void stacking_mark(oracle_state *s, uint32_t npart, uint32_t seed) {
    const uint32_t xres = s->xres, xc = xres / 4, n = s->xres * s->yres;
    for (uint32_t idx = 0; idx < n; ++idx) {
        uint32_t count = s->pmap_count[idx];
        if (count <= 5) continue;
        uint32_t x = idx % xres, y = idx / xres;
        uint32_t wall = s->wall_bmap[(y / 4) * xc + x / 4] & 0xFFu;
        if (wall == 12u) {
            if (count > 1500u) s->pmap_count[idx] = count + npart;
        } else {
            uint32_t roll = oracle_lowbias_hash(seed ^ (idx * 0x2545F491u)) % 1600u;
            if (count > 1500u || roll <= count + 100u) s->pmap_count[idx] = count + npart;
        }
    }
}

/* spatial.rs:224-275 */
static void stacking_apply(oracle_state *s, uint32_t npart) {
    const uint32_t xres = s->xres, yres = s->yres;
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= 256) continue;
        int x = ora_f2i(s->part_x[i] + 0.5f), y = ora_f2i(s->part_y[i] + 0.5f);
        if (x < 0 || y < 0 || x >= (int)xres || y >= (int)yres) continue;
        if (s->elem_props[ty * 32 + 1] & 0x10u) continue;
        uint32_t idx = (uint32_t)y * xres + (uint32_t)x;
        uint32_t count = s->pmap_count[idx];
        if (count < npart) continue;
        uint32_t owner = 0xFFFFFFFFu;
        if (s->pmap[idx] != 0) owner = (s->pmap[idx] >> 9) - 1;
        if (count > npart && owner == i) {
            s->part_ty[i] = 40u; /* PT_NBHL */
            s->part_temp[i] = 9999.0f;
            uint32_t strength = count - npart;
            if (strength > 51200u) strength = 51200u;
            s->part_tmp[i] = (int32_t)strength;
            s->pmap_count[idx] = npart;
            s->pmap[idx] = ((i + 1) << 9) | 40u;
        } else {
            s->part_ty[i] = 0;
        }
    }
}

void oracle_recalc_free_particles(oracle_state *s, uint32_t seed, int stacking) {
    const size_t n = (size_t)s->xres * s->yres;
    memset(s->pmap, 0, n * 4);
    memset(s->photons, 0, n * 4);
    memset(s->pmap_count, 0, n * 4);
    recalc_scatter(s);
    if (stacking) {
        uint32_t npart = s->xres * s->yres;
        stacking_mark(s, npart, seed);
        stacking_apply(s, npart);
    }
}
