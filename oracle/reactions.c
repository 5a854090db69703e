/* The part of reaction_kernel::step_reactions_gpu
 * (reaction_kernel/mod.rs:1123-1237) that the tick here runs -- TEST
 * INFRASTRUCTURE ONLY (see tpt_oracle.h): the modext clear of the gravity
 * mass-override channel (reaction_kernel/modext/mod.rs:53-99) and the life
 * pass (life_and_local_reaction_kernel, reaction_kernel/life.rs:47-205). */
#include "tpt_oracle.h"

#include <string.h>

#include "oracle_util.h"

enum {
    T_WATR = 2, T_FIRE = 4, T_METL = 11, T_SPRK = 12, T_SMKE = 16, T_SLTW = 19, T_WTRV = 21, T_PLSM = 28,
    T_PSCN = 50, T_NSCN = 51, T_ETRD = 55, T_SWCH = 60, T_INWR = 63, T_NBLE = 97, T_HFLM = 100, T_TRON = 151,
    T_RSST = 191,
};
#define PROP_LIFE_KILL 0x8000u
#define FLAG_STASIS 0x20u

/* life_and_local_reaction_kernel (life.rs:47-205): per slot, from the
 * slot's own state only.  A particle whose rounded pixel is off the canvas
 * counts as outside any stasis wall (the reference indexes the wall grid
 * without a bounds check). */
void oracle_life_pass(oracle_state *s, uint32_t seed) {
    const uint32_t xc = s->xres / 4, yc = s->yres / 4;
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        const uint32_t ty = s->part_ty[i];
        if (ty == 0) continue;
        const int xi = ora_f2i(s->part_x[i] + 0.5f), yi = ora_f2i(s->part_y[i] + 0.5f);
        int frozen = 0;
        if (xi >= 0 && yi >= 0 && xi < (int)s->xres && yi < (int)s->yres) {
            const uint32_t cell = (uint32_t)(yi / 4) * xc + (uint32_t)(xi / 4);
            frozen = cell < xc * yc && s->wall_bmap[cell] == 18u && (s->wall_emap[cell] & 0xFFu) < 8u;
        }
        if (frozen) {
            s->part_flags[i] |= FLAG_STASIS;
            continue;
        }
        s->part_flags[i] &= ~FLAG_STASIS;
        const uint32_t props = ty < 256u ? s->elem_props[ty * 32u + 1u] : 0u;
        int32_t life = s->part_life[i];
        uint32_t new_ty = ty;
        int dead = 0;
        if (ty == T_FIRE && life == 0)
            life = 180 + (int32_t)(oracle_hash_u32(i ^ seed) % 80u);
        else if ((ty == T_PLSM || ty == T_HFLM) && life == 0)
            life = 50 + (int32_t)(oracle_hash_u32(i ^ seed ^ 0xA5A5u) % 150u);
        else if (ty == T_TRON && life == 0)
            life = 25 + (int32_t)(oracle_hash_u32(i ^ seed ^ 0x3C3Cu) % 25u);
        if (ty == T_FIRE) {
            if (life <= 1) {
                if ((s->part_tmp[i] & 3) == 3) {
                    new_ty = T_WTRV;
                    life = 0;
                    s->part_ctype[i] = T_FIRE;
                } else if (s->part_temp[i] < 625.0f) {
                    new_ty = T_SMKE;
                    life = 250 + (int32_t)(oracle_hash_u32(i ^ seed ^ 0x5ACE123u) % 20u);
                    s->part_tmp[i] = 0;
                } else {
                    dead = 1;
                }
            }
        } else if (ty == T_PLSM) {
            if (life <= 1) {
                if (s->part_ctype[i] == T_NBLE) {
                    new_ty = T_NBLE;
                    life = 0;
                    s->part_ctype[i] = 0;
                } else if ((s->part_tmp[i] & 3) == 3) {
                    new_ty = T_WTRV;
                    life = 0;
                    s->part_ctype[i] = T_FIRE;
                } else {
                    dead = 1;
                }
            }
        } else if (ty == T_SPRK) {
            if (life <= 0) {
                uint32_t ctype = s->part_ctype[i];
                if (ctype == T_WATR || ctype == T_SLTW || ctype == T_PSCN || ctype == T_NSCN || ctype == T_ETRD ||
                    ctype == T_INWR)
                    s->part_temp[i] = 293.15f;
                if (ctype == 0 || ctype >= 256u) ctype = T_METL;
                s->part_ctype[i] = 0;
                life = 4;
                if (ctype == T_WATR) life = 64;
                else if (ctype == T_SLTW) life = 54;
                else if (ctype == T_SWCH) life = 14;
                else if (ctype == T_RSST) dead = 1;
                if (!dead) new_ty = ctype;
            }
        }
        if (new_ty == ty && (props & PROP_LIFE_KILL) && life <= 0) dead = 1;
        if (dead) {
            s->part_ty[i] = 0;
        } else {
            if (new_ty != ty) s->part_ty[i] = new_ty;
            s->part_life[i] = life;
        }
    }
}

void oracle_step_reactions(oracle_state *s, uint32_t seed) {
    oracle_clear_grav_override(s);
    oracle_life_pass(s, seed);
}
