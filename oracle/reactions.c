/* The part of reaction_kernel::step_reactions_gpu
 * (reaction_kernel/mod.rs:1123-1237) that the tick here runs -- TEST
 * INFRASTRUCTURE ONLY (see tpt_oracle.h): the modext clear of the gravity
 * mass-override channel (reaction_kernel/modext/mod.rs:53-99), the life
 * pass (life_and_local_reaction_kernel, reaction_kernel/life.rs:47-205),
 * the combustion pass's FIRE/PLSM and WTRV branches
 * (combustion_reaction_kernel, combustion/fire.rs:26-408, 983-1013) and the
 * fluid pass's WATR, DSTW, SLTW, ICEI and SNOW branches
 * (fluid_reaction_kernel, fluid.rs:149-400), in the driver's order.
 *
 * The two pixel-parallel passes race in the reference (a burner heats and
 * ignites neighbours other burners read; a water kills a fire another water
 * reads).  Both run here under one canonical rule, as a serial loop over the
 * pixels in ascending order:
 *   - an acting particle decides from the state the pass started with: a
 *     burner's own temperature, which pixels hold burners, the neighbour
 *     types a WTRV or water rule matches, the cell pressure the ignition
 *     roll reads;
 *   - the particle acted upon changes in the acting pixels' ascending order,
 *     each action seeing the effects of the earlier ones (heat composes; a
 *     target keeps the first conversion, since a later rule applies only
 *     while the target still has the type the rule matched);
 *   - a particle's own update happens at its own pixel's turn. */
#include "tpt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_util.h"

enum {
    T_WATR = 2, T_FIRE = 4, T_STNE = 5, T_METL = 11, T_SPRK = 12, T_SNOW = 13, T_ICEI = 14, T_SMKE = 16,
    T_SAND = 17, T_SALT = 18, T_SLTW = 19, T_WTRV = 21, T_BCOL = 27, T_PLSM = 28, T_DSTW = 35, T_PLNT = 45,
    T_PSCN = 50, T_NSCN = 51, T_ETRD = 55, T_SWCH = 60, T_COAL = 62, T_INWR = 63, T_SPNG = 78, T_FRZZ = 83,
    T_FRZW = 84, T_RBDM = 93, T_LRBD = 94, T_NBLE = 97, T_HFLM = 100, T_TRON = 151, T_ROCK = 189, T_RSST = 191,
    T_BASE = 193,
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

/* Element-table words (GpuElementProps) the passes read. */
#define W_FLAMMABLE 12u
#define W_EXPLOSIVE 13u
#define W_DEFAULT_TEMP 16u

static uint32_t elem_word(const oracle_state *s, uint32_t ty, uint32_t w) {
    return ty < 256u ? s->elem_props[ty * 32u + w] : 0u;
}

/* chance (combustion/mod.rs:280-286): hash_u32(seed) % den < num. */
static int chance(uint32_t seed, uint32_t num, uint32_t den) {
    return den != 0u && oracle_hash_u32(seed) % den < num;
}

/* pmap_at / pmap_slot (shared.rs:1069-1089). */
static uint32_t pmap_at(const oracle_state *s, int x, int y) {
    if (x < 0 || y < 0 || x >= (int)s->xres || y >= (int)s->yres) return 0u;
    return s->pmap[(uint32_t)y * s->xres + (uint32_t)x];
}
static uint32_t pmap_slot(uint32_t v) {
    const uint32_t raw = v >> 9;
    return raw ? raw - 1u : 0u;
}

/* The state a pass starts from, per pixel: the owner's type, whether it is
 * a burner (FIRE or PLSM owner outside stasis) and its temperature. */
typedef struct {
    uint32_t *ty;
    uint8_t *burner;
    float *temp;
    float *pv;
} pass_snapshot;

static void take_snapshot(const oracle_state *s, pass_snapshot *sn) {
    const size_t np = (size_t)s->xres * s->yres, nc = (size_t)(s->xres / 4) * (s->yres / 4);
    sn->ty = calloc(np, 4);
    sn->burner = calloc(np, 1);
    sn->temp = calloc(np, 4);
    sn->pv = malloc(nc * 4);
    memcpy(sn->pv, s->air_pv, nc * 4);
    for (size_t q = 0; q < np; ++q) {
        const uint32_t v = s->pmap[q];
        if (!v) continue;
        const uint32_t id = pmap_slot(v), ty = s->part_ty[id];
        sn->ty[q] = ty;
        sn->burner[q] = (ty == T_FIRE || ty == T_PLSM) && !(s->part_flags[id] & FLAG_STASIS);
        sn->temp[q] = s->part_temp[id];
    }
}

static void free_snapshot(pass_snapshot *sn) {
    free(sn->ty);
    free(sn->burner);
    free(sn->temp);
    free(sn->pv);
}

#define STEAM_IGNITE_K 285.15f
#define STEAM_COOL_K 1273.0f

/* combustion_reaction_kernel (combustion/fire.rs:26-1230): the FIRE/PLSM
 * heating and 5x5 ignition (:143-408; the COAL/BCOL timer at :188-197), and
 * the WTRV branch (:983-1013).  LAVA and the other element branches of the
 * kernel are not restated. */
void oracle_combustion_pass(oracle_state *s, uint32_t seed) {
    pass_snapshot sn;
    take_snapshot(s, &sn);
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4;
    const float fire_t = ora_bits_f(elem_word(s, T_FIRE, W_DEFAULT_TEMP));
    for (uint32_t q = 0; q < xres * yres; ++q) {
        const uint32_t v = s->pmap[q];
        if (!v) continue;
        const uint32_t id = pmap_slot(v), ty = s->part_ty[id];
        if (ty == 0 || (s->part_flags[id] & FLAG_STASIS)) continue;
        const int px = (int)(q % xres), py = (int)(q / xres);
        const uint32_t cidx = (uint32_t)(py / 4) * xc + (uint32_t)(px / 4);
        const uint32_t prand = oracle_hash_u32(seed ^ (q * 2654435761u));
        if (sn.burner[q]) {
            const float temp = sn.temp[q];
            uint32_t surround = 0;
            for (int n = 0; n < 9; ++n) {
                const int dx = n % 3 - 1, dy = n / 3 - 1;
                if ((dx || dy) && pmap_at(s, px + dx, py + dy) == 0u) ++surround;
            }
            for (uint32_t n = 0; n < 25; ++n) {
                const int dx = (int)(n % 5) - 2, dy = (int)(n / 5) - 2;
                if (!dx && !dy) continue;
                const uint32_t nv = pmap_at(s, px + dx, py + dy);
                if (!nv) continue;
                const uint32_t nid = pmap_slot(nv), nty = s->part_ty[nid];
                if (nty == 0) continue;
                const float ntemp = s->part_temp[nid];
                const uint32_t nroll = oracle_hash_u32(prand ^ (n * 0x9E3779B9u));
                if ((nty == T_COAL || nty == T_BCOL) && s->part_life[nid] > 100 && chance(nroll, 1u, 500u))
                    s->part_life[nid] = 99;
                if (temp > ntemp) {
                    float heated = ntemp + (temp - ntemp) * 0.05f;
                    if (heated > 9999.0f) heated = 9999.0f;
                    s->part_temp[nid] = heated;
                }
                const int32_t nflam = (int32_t)elem_word(s, nty, W_FLAMMABLE);
                const int nexp = (int32_t)elem_word(s, nty, W_EXPLOSIVE) != 0;
                int ignitable = nflam > 0 && (surround > 0u || nexp);
                if (nty == T_SPNG && s->part_life[nid] > 0) ignitable = 0;
                if (!ignitable) continue;
                const uint32_t ncidx = (uint32_t)((py + dy) / 4) * xc + (uint32_t)((px + dx) / 4);
                const int32_t num = ora_f2i((float)nflam + sn.pv[ncidx] * 10.0f);
                int hit = 0;
                if (num >= 1000) hit = 1;
                else if (num > 0) hit = chance(seed ^ (nid * 2654435761u) ^ n, (uint32_t)num, 1000u);
                if (!hit) continue;
                s->part_ty[nid] = T_FIRE;
                s->part_life[nid] = 180 + (int32_t)(oracle_hash_u32(nid ^ seed) % 80u);
                float t = fire_t + (float)(nflam / 2);
                if (t < 0.0f) t = 0.0f;
                if (t > 9999.0f) t = 9999.0f;
                s->part_temp[nid] = t;
                s->part_tmp[nid] = 0;
                s->part_ctype[nid] = 0;
                if (nexp) s->air_pv[cidx] += 0.25f;
            }
        } else if (ty == T_WTRV) {
            const float temp = s->part_temp[id];
            int ignited = 0;
            for (uint32_t n = 0; n < 9 && !ignited; ++n) {
                const int dx = (int)(n % 3) - 1, dy = (int)(n / 3) - 1;
                if (!dx && !dy) continue;
                if (!pmap_at(s, px + dx, py + dy)) continue;
                const uint32_t nty = sn.ty[(uint32_t)(py + dy) * xres + (uint32_t)(px + dx)];
                if ((nty == T_RBDM || nty == T_LRBD) && temp > STEAM_IGNITE_K && chance(prand ^ n, 1u, 100u)) {
                    s->part_ty[id] = T_FIRE;
                    s->part_life[id] = 4;
                    s->part_ctype[id] = T_WATR;
                    ignited = 1;
                }
            }
            if (!ignited && temp > STEAM_COOL_K && s->part_ctype[id] == T_FIRE) s->part_temp[id] = temp - temp / 1000.0f;
        }
    }
    free_snapshot(&sn);
}

/* fluid.rs:99-110 salts, :142-144 roll. */
#define SALT_WATR 0x2545F491u
#define SALT_DSTW 0x9E3779B9u
#define SALT_SLTW 0x85EBCA6Bu
#define SALT_ICE 0xC2B2AE35u
#define ROLL_DERIVE 0x517CC1B7u

static uint32_t fluid_roll(uint32_t seed, uint32_t idx, uint32_t salt, uint32_t n) {
    return oracle_hash_u32(seed ^ (idx * salt) ^ (n * ROLL_DERIVE));
}

/* A conversion of neighbour nid that the rule matched on type `matched`:
 * applied only while the neighbour still has that type (the first
 * conversion wins). */
static void convert(oracle_state *s, uint32_t nid, uint32_t matched, uint32_t to) {
    if (s->part_ty[nid] == matched) s->part_ty[nid] = to;
}

/* fluid_reaction_kernel (fluid.rs:149-400): WATR (:180-242), DSTW
 * (:243-313), SLTW (:314-355), ICEI and SNOW (:356-399).  `legacy` is the
 * reference's !heat_enable. */
void oracle_fluid_pass(oracle_state *s, uint32_t seed) {
    pass_snapshot sn;
    take_snapshot(s, &sn);
    const uint32_t xres = s->xres, yres = s->yres;
    const int legacy = !s->heat_enable;
    for (uint32_t q = 0; q < xres * yres; ++q) {
        const uint32_t v = s->pmap[q];
        if (!v) continue;
        /* the actors are the water-family and ice owners the pass started
           with: a SALT an earlier actor turned into SLTW does not act as
           one this pass */
        const uint32_t id = pmap_slot(v), ty = sn.ty[q];
        if (ty == 0 || s->part_ty[id] != ty || (s->part_flags[id] & FLAG_STASIS)) continue;
        if (ty != T_WATR && ty != T_DSTW && ty != T_SLTW && ty != T_ICEI && ty != T_SNOW) continue;
        const int px = (int)(q % xres), py = (int)(q / xres);
        uint32_t new_ty = ty;
        const float speed = fabsf(s->part_vx[id]) + fabsf(s->part_vy[id]);
        int stop = 0;
        float itemp = s->part_temp[id];
        if (ty == T_ICEI || ty == T_SNOW) {
            if (s->part_ctype[id] == T_FRZW) {
                itemp -= 1.0f;
                if (itemp < 0.0f) itemp = 0.0f;
                if (itemp > 9999.0f) itemp = 9999.0f;
                s->part_temp[id] = itemp;
            }
        }
        for (uint32_t n = 0; n < 9 && !stop; ++n) {
            const int dx = (int)(n % 3) - 1, dy = (int)(n / 3) - 1;
            if (!dx && !dy) continue;
            const uint32_t nv = pmap_at(s, px + dx, py + dy);
            if (!nv) continue;
            const uint32_t nid = pmap_slot(nv), nty = sn.ty[(uint32_t)(py + dy) * xres + (uint32_t)(px + dx)];
            if (ty == T_WATR) {
                const uint32_t r = fluid_roll(seed, q, SALT_WATR, n);
                if (nty == T_SALT && r % 50u == 0u) {
                    new_ty = T_SLTW;
                    if (oracle_hash_u32(r ^ 0x51ED270Bu) % 3u == 0u) convert(s, nid, T_SALT, T_SLTW);
                } else if ((nty == T_RBDM || nty == T_LRBD) && (legacy || s->part_temp[id] > 285.15f) &&
                           r % 100u == 0u) {
                    new_ty = T_FIRE;
                    s->part_life[id] = 4;
                    s->part_ctype[id] = T_WATR;
                } else if (nty == T_FIRE && s->part_ctype[nid] != T_WATR) {
                    convert(s, nid, T_FIRE, 0u);
                    if (oracle_hash_u32(r ^ 0x2C1B3E4Du) % 30u == 0u) stop = 1;
                } else if (nty == T_SLTW && r % 2000u == 0u) {
                    new_ty = T_SLTW;
                } else if (nty == T_ROCK && speed >= 0.5f && r % 1000u == 0u) {
                    convert(s, nid, T_ROCK, oracle_hash_u32(r ^ 0x7F4A7C15u) % 3u == 0u ? T_SAND : T_STNE);
                }
            } else if (ty == T_DSTW) {
                const uint32_t r = fluid_roll(seed, q, SALT_DSTW, n);
                if (nty == T_SALT && r % 50u == 0u) {
                    new_ty = T_SLTW;
                    if (oracle_hash_u32(r ^ 0x51ED270Bu) % 3u == 0u) convert(s, nid, T_SALT, T_SLTW);
                } else if (nty == T_SLTW) {
                    if (r % 2000u == 0u) new_ty = T_SLTW;
                    else if (oracle_hash_u32(r ^ 0x3AE91F27u) % 100u == 0u) new_ty = T_WATR;
                } else if (nty == T_WATR && r % 100u == 0u) {
                    new_ty = T_WATR;
                } else if ((nty == T_RBDM || nty == T_LRBD) && (legacy || s->part_temp[id] > 12.0f) &&
                           r % 100u == 0u) {
                    new_ty = T_FIRE;
                    s->part_life[id] = 4;
                } else if (nty == T_FIRE) {
                    convert(s, nid, T_FIRE, 0u);
                    if (oracle_hash_u32(r ^ 0x2C1B3E4Du) % 30u == 0u) stop = 1;
                } else if (nty == T_SMKE) {
                    const float nt = s->part_temp[nid], t = s->part_temp[id];
                    if (nt > 313.15f && nt < 333.15f && t > 313.15f && t < 333.15f && r % 100u == 0u) {
                        new_ty = T_BASE;
                        s->part_life[id] = 1;
                        convert(s, nid, T_SMKE, 0u);
                    }
                }
            } else if (ty == T_SLTW) {
                const uint32_t r = fluid_roll(seed, q, SALT_SLTW, n);
                if (nty == T_SALT && r % 2000u == 0u) {
                    convert(s, nid, T_SALT, T_SLTW);
                } else if (nty == T_PLNT && r % 40u == 0u) {
                    convert(s, nid, T_PLNT, 0u);
                } else if ((nty == T_RBDM || nty == T_LRBD) && (legacy || s->part_temp[id] > 285.15f) &&
                           r % 100u == 0u) {
                    new_ty = T_FIRE;
                    s->part_life[id] = 4;
                    s->part_ctype[id] = T_WATR;
                } else if (nty == T_FIRE && s->part_ctype[nid] != T_WATR) {
                    convert(s, nid, T_FIRE, 0u);
                    if (oracle_hash_u32(r ^ 0x2C1B3E4Du) % 30u == 0u) stop = 1;
                }
            } else {  /* ICEI, SNOW */
                const uint32_t r = fluid_roll(seed, q, SALT_ICE, n), div = ty == T_ICEI ? 200u : 333u;
                if ((nty == T_SALT || nty == T_SLTW) && (ty == T_SNOW || itemp > 252.05f) && r % div == 0u) {
                    new_ty = T_SLTW;
                    convert(s, nid, nty, T_SLTW);
                    stop = 1;
                } else if (ty == T_ICEI && nty == T_FRZZ && r % 200u == 0u) {
                    if (s->part_ty[nid] == T_FRZZ) {
                        s->part_ty[nid] = T_ICEI;
                        s->part_ctype[nid] = T_FRZW;
                    }
                }
            }
        }
        /* WATR, DSTW and SLTW end dead after a killing extinguish; ICEI and
           SNOW only stop */
        if (stop && ty != T_ICEI && ty != T_SNOW) new_ty = 0u;
        if (new_ty != ty) s->part_ty[id] = new_ty;
    }
    free_snapshot(&sn);
}

void oracle_step_reactions(oracle_state *s, uint32_t seed) {
    oracle_clear_grav_override(s);
    oracle_life_pass(s, seed);
    oracle_combustion_pass(s, seed);
    oracle_fluid_pass(s, seed);
}
