/*
 * Oracle: conduction, temperature apply, phase transitions and the pmap type
 * refresh -- TEST INFRASTRUCTURE ONLY (see tpt_oracle.h).
 *
 * Restates crates/tpt-sim/src/thermal_kernel.rs: conduction :459-605, apply
 * :616-708, phase transitions :718-1318, pmap refresh :1338-1387 and the
 * driver :1574-1717.
 *
 * Canonical rules for the reference's races (shared with the HIP kernels):
 *  R2  ambient exchange: every particle reads the pre-pass air_hv; the cell
 *      keeps the value written by its highest gated slot.
 *  R3  block_airh: the n gate failures of a cell apply the saturating
 *      "+1 while bit 3 is clear" n times, in one step.
 *  R5  phase transitions read the pre-pass air_pv (and wall_emap); a cell
 *      keeps the pressure of its highest slot whose pv changed; spark
 *      charges (emap = 16) land after every particle has read the map.
 *  R6  liquid convection swap: every liquid proposes one swap partner from
 *      the pre-pass temperatures and types; a proposal (i -> t) is carried
 *      out iff i is the lowest slot among all proposals that touch i or t
 *      (one round of a priority matching: the carried-out swaps are
 *      disjoint, so their order does not matter).  The rest of the
 *      transition kernel runs after all swaps.
 */
#include "tpt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_util.h"

#define COL_PROPERTIES 1
#define COL_FLAMMABLE 12
#define COL_EXPLOSIVE 13
#define COL_HEAT_CONDUCT 15
#define COL_DEFAULT_TEMP 16
#define COL_LOW_PRESSURE 18
#define COL_LOW_PRESSURE_TRANSITION 19
#define COL_HIGH_PRESSURE 20
#define COL_HIGH_PRESSURE_TRANSITION 21
#define COL_LOW_TEMP 22
#define COL_LOW_TEMP_TRANSITION 23
#define COL_HIGH_TEMP 24
#define COL_HIGH_TEMP_TRANSITION 25

enum {
    T_NONE = 0, T_WATR = 2, T_FIRE = 4, T_LAVA = 6, T_SPRK = 12, T_SNOW = 13, T_ICEI = 14, T_SAND = 17,
    T_SALT = 18, T_SLTW = 19, T_ACID = 20, T_WTRV = 21, T_PLSM = 28, T_PHOT = 29, T_ELEC = 31, T_DSTW = 35,
    T_LN2 = 38, T_BMTL = 46, T_BRMT = 47, T_HSWC = 68, T_RIME = 79, T_FOG = 80, T_PIPE = 82, T_FRZW = 84,
    T_BIZR = 85, T_BIZRG = 86, T_PLUT = 89, T_BGLA = 95, T_THRM = 99, T_CFLM = 100, T_LIFE = 105, T_DEUT = 111,
    T_FILT = 134, T_BRAY = 136, T_PQRT = 142, T_QRTZ = 141, T_GEL = 150, T_PPIP = 162, T_VIBR = 165,
    T_BVBR = 166, T_GLAS = 24, T_TUNG = 171, T_CRMC = 178, T_HEAC = 179, T_LITH = 190, T_BASE = 193,
    T_SEED = 194, T_BISM = 239
};

#define NT (-1)
#define ST 512
#define PARTICLE_HASH_SALT 2654435761u
#define TRANSITION_HASH_SALT 0x51ED2701u

/* thermal_kernel.rs:289-296 */
uint32_t oracle_gpu_hash(uint32_t seed) {
    uint32_t h = seed ^ 0x9E3779B9u;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    return h ^ (h >> 16);
}

static uint32_t P(const oracle_state *s, uint32_t t, int col) { return s->elem_props[t * 32 + col]; }
static float PF(const oracle_state *s, uint32_t t, int col) { return ora_bits_f(s->elem_props[t * 32 + col]); }

/* thermal_kernel.rs:356-366 */
static int is_heat_insulator(uint32_t ty, int32_t life, int32_t tmp, uint32_t hc) {
    return hc == 0 || (ty == T_HSWC && life != 10) || ((ty == T_PIPE || ty == T_PPIP) && (tmp & 1) == 0);
}

/* thermal_kernel.rs:373-386 */
static float heat_capacity(uint32_t ty, uint32_t ctype, uint32_t rows) {
    float hc = 1.0f;
    if (ty == T_BASE) hc = 1.5f;
    if ((ty == T_PIPE || ty == T_PPIP) && ctype > 0 && ctype < rows) hc += ctype == T_BASE ? 1.5f : 1.0f;
    return hc;
}

/* thermal_kernel.rs:391-404 */
static int conduct_pair_excluded(uint32_t t, uint32_t rt, int32_t st, int32_t rtt) {
    return (t == T_FILT && (rt == T_BRAY || rt == T_BIZR || rt == T_BIZRG)) ||
           (rt == T_FILT && (t == T_BRAY || t == T_PHOT || t == T_BIZR || t == T_BIZRG)) ||
           (t == T_ELEC && rt == T_DEUT) || (t == T_DEUT && rt == T_ELEC) ||
           (t == T_HSWC && rt == T_FILT && st == 1) || (t == T_FILT && rt == T_HSWC && rtt == 1);
}

static uint32_t neighbour_slot(const oracle_state *s, int x, int y) {
    if (x < 0 || y < 0 || x >= (int)s->xres || y >= (int)s->yres) return 0xFFFFFFFFu;
    uint32_t packed = s->pmap[(uint32_t)y * s->xres + (uint32_t)x];
    if ((packed >> 9) == 0) return 0xFFFFFFFFu;
    return (packed >> 9) - 1;
}

/* thermal_conduction_kernel (thermal_kernel.rs:459-605) with rules R2/R3.
 * The per-slot work reads only pre-pass state (temperatures, pmap, air_hv)
 * and writes the slot's own equilibrium, so it runs in parallel; each slot
 * leaves its ambient-heat write and gate failure in a per-slot record, and
 * one serial pass in ascending slot order applies them (the highest gated
 * slot's write per cell, R2; the failure count per cell, R3). */
enum { COND_NONE = 0, COND_HV = 1, COND_GATED = 2, COND_FAIL = 3 };

void oracle_thermal_conduction(oracle_state *s, uint32_t seed) {
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4, n_cells = xc * (yres / 4), rows = 256;
    const uint32_t nslots = s->slot_bound;
    const float *hv0 = s->air_hv;
    float *hv_new = malloc(sizeof(float) * n_cells);
    uint32_t *hv_key = calloc(n_cells, 4), *fails = calloc(n_cells, 4);
    uint8_t *kind = malloc(nslots ? nslots : 1);
    uint32_t *kcell = malloc(sizeof(uint32_t) * (nslots ? nslots : 1));
    float *kw = malloc(sizeof(float) * (nslots ? nslots : 1));
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint32_t i = 0; i < nslots; ++i) {
        kind[i] = COND_NONE;
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= rows) continue;
        uint32_t hcond = P(s, ty, COL_HEAT_CONDUCT);
        int x = ora_f2i(s->part_x[i]), y = ora_f2i(s->part_y[i]);
        if (x < 0 || x >= (int)xres || y < 0 || y >= (int)yres) continue;
        uint32_t ci = ((uint32_t)y / 4) * xc + (uint32_t)x / 4;
        kcell[i] = ci;
        int32_t life = s->part_life[i], tmp = s->part_tmp[i];
        uint32_t ctype = s->part_ctype[i];
        float gel = 1.0f;
        if (ty == T_GEL) gel = (float)tmp * 2.55f;
        float scaled = (float)hcond * gel;
        if (scaled < 0.0f) scaled = 0.0f;
        uint32_t threshold = ora_f2u(scaled);
        uint32_t roll = oracle_gpu_hash(seed ^ (i * PARTICLE_HASH_SALT) ^ ty) % 250u;
        int gate = !is_heat_insulator(ty, life, tmp, hcond) && roll < threshold;
        if (gate) {
            kind[i] = COND_GATED;
            float self_temp = s->part_temp[i];
            if (s->aheat_enable && (P(s, ty, COL_PROPERTIES) & 0x00200000u) == 0) {
                float dtemp = hv0[ci] - s->part_temp[i];
                float hc = heat_capacity(ty, ctype, rows);
                float alpha = 0.4f * hc;
                if (alpha > 0.04f) alpha = 0.04f;
                self_temp = ora_clampf(s->part_temp[i] + alpha * dtemp / hc, 0.0f, 9999.0f);
                kw[i] = ora_clampf(hv0[ci] - alpha * dtemp, 0.0f, 9999.0f);
                kind[i] = COND_HV;
            }
            float hc_self = heat_capacity(ty, ctype, rows);
            float c_heat = self_temp * hc_self, hc_total = hc_self;
            for (int ny = 0; ny < 3; ++ny)
                for (int nx = 0; nx < 3; ++nx) {
                    if (nx == 1 && ny == 1) continue;
                    uint32_t nid = neighbour_slot(s, x + nx - 1, y + ny - 1);
                    if (nid == 0xFFFFFFFFu || nid >= nslots) continue;
                    uint32_t nty = s->part_ty[nid];
                    if (nty == 0 || nty >= rows) continue;
                    uint32_t nc = P(s, nty, COL_HEAT_CONDUCT);
                    if (is_heat_insulator(nty, s->part_life[nid], s->part_tmp[nid], nc)) continue;
                    if (conduct_pair_excluded(ty, nty, tmp, s->part_tmp[nid])) continue;
                    float n_hc = heat_capacity(nty, s->part_ctype[nid], rows);
                    c_heat += s->part_temp[nid] * n_hc;
                    hc_total += n_hc;
                }
            s->part_equil[i] = ora_clampf(c_heat / hc_total, 0.0f, 9999.0f);
        } else {
            kind[i] = COND_FAIL;
            s->part_equil[i] = -1.0f;
        }
    }
    for (uint32_t i = 0; i < nslots; ++i) {
        if (kind[i] == COND_HV) {
            hv_key[kcell[i]] = i + 1;
            hv_new[kcell[i]] = kw[i];
        } else if (kind[i] == COND_FAIL) {
            fails[kcell[i]] += 1;
        }
    }
    for (uint32_t c = 0; c < n_cells; ++c) {
        if (hv_key[c]) s->air_hv[c] = hv_new[c];
        uint32_t v = s->block_airh[c];
        if (fails[c] && (v & 8u) == 0) {
            uint32_t room = 8u - (v & 7u);
            v += fails[c] < room ? fails[c] : room;
            s->block_airh[c] = v;
        }
    }
    free(hv_new);
    free(hv_key);
    free(fails);
    free(kind);
    free(kcell);
    free(kw);
}

/* thermal_apply_kernel (thermal_kernel.rs:616-708). */
void oracle_thermal_apply(oracle_state *s) {
    const uint32_t xres = s->xres, yres = s->yres, rows = 256;
    /* Each slot reads only its own temperature and the (unchanged)
     * equilibria, so slots are independent. */
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint32_t j = 0; j < s->slot_bound; ++j) {
        uint32_t jty = s->part_ty[j];
        if (jty == 0 || jty >= rows) continue;
        int x = ora_f2i(s->part_x[j]), y = ora_f2i(s->part_y[j]);
        if (x < 0 || x >= (int)xres || y < 0 || y >= (int)yres) continue;
        uint32_t best = j;
        float best_temp = s->part_temp[j];
        int has = 0;
        if (s->part_equil[j] >= 0.0f) {
            best_temp = s->part_equil[j];
            has = 1;
        }
        int j_ins = is_heat_insulator(jty, s->part_life[j], s->part_tmp[j], P(s, jty, COL_HEAT_CONDUCT));
        for (int ny = 0; ny < 3; ++ny)
            for (int nx = 0; nx < 3; ++nx) {
                if (nx == 1 && ny == 1) continue;
                uint32_t nid = neighbour_slot(s, x + nx - 1, y + ny - 1);
                if (nid == 0xFFFFFFFFu || nid >= s->slot_bound || nid == j) continue;
                uint32_t nty = s->part_ty[nid];
                if (nty == 0 || nty >= rows || !(has == 0 || nid > best) || !(s->part_equil[nid] >= 0.0f) || j_ins)
                    continue;
                if (is_heat_insulator(nty, s->part_life[nid], s->part_tmp[nid], P(s, nty, COL_HEAT_CONDUCT))) continue;
                if (conduct_pair_excluded(nty, jty, s->part_tmp[nid], s->part_tmp[j])) continue;
                best = nid;
                best_temp = s->part_equil[nid];
                has = 1;
            }
        s->part_temp[j] = ora_clampf(has ? best_temp : s->part_temp[j], 0.0f, 9999.0f);
    }
}

/* convection_gravity (thermal_kernel.rs:412-450). */
static void convection_gravity(const oracle_state *s, int x, int y, uint32_t ci, float *ogx, float *ogy) {
    const float pg = -2.0f;
    float gx = 0.0f, gy = 0.0f;
    if (s->gravity_mode == 0) {
        gy = pg;
    } else if (s->gravity_mode == 2) {
        float dx = (float)x - (float)s->xres * 0.5f, dy = (float)y - (float)s->yres * 0.5f;
        if (dx != 0.0f || dy != 0.0f) {
            float d = 0.01f - sqrtf(dx * dx + dy * dy);
            gx = pg * (dx / d);
            gy = pg * (dy / d);
        }
    } else if (s->gravity_mode == 3) {
        gx = pg * s->custom_gravity_x;
        gy = pg * s->custom_gravity_y;
    }
    gx += -2.0f * s->grav_fx[ci];
    gy += -2.0f * s->grav_fy[ci];
    *ogx = gx;
    *ogy = gy;
}

/* The liquid's convection partner from the pre-pass snapshot, or NONE. */
static uint32_t convect_target(const oracle_state *s, uint32_t i, uint32_t roll) {
    uint32_t t = s->part_ty[i];
    if (t == 0 || t >= 256) return 0xFFFFFFFFu;
    int x = ora_f2i(s->part_x[i]), y = ora_f2i(s->part_y[i]);
    if (x < 0 || x >= (int)s->xres || y < 0 || y >= (int)s->yres) return 0xFFFFFFFFu;
    if ((P(s, t, COL_PROPERTIES) & 0x02u) == 0) return 0xFFFFFFFFu;
    if (t == T_GEL) {
        float gel = (float)s->part_tmp[i] * 2.55f;
        uint32_t gate_roll = oracle_gpu_hash(roll ^ 0x1B873593u) % 255u;
        if (gel <= 1.0f + (float)gate_roll) return 0xFFFFFFFFu;
    }
    uint32_t ci = ((uint32_t)y / 4) * (s->xres / 4) + (uint32_t)x / 4;
    float gx, gy;
    convection_gravity(s, x, y, ci, &gx, &gy);
    float rxf = gx + (float)x, ryf = gy + (float)y;
    int rx = rxf >= 0.0f ? ora_f2i(rxf + 0.5f) : ora_f2i(rxf - 0.5f);
    int ry = ryf >= 0.0f ? ora_f2i(ryf + 0.5f) : ora_f2i(ryf - 0.5f);
    int ox = ora_clampi(rx, x - 1, x + 1), oy = ora_clampi(ry, y - 1, y + 1);
    if (ox == x && oy == y) return 0xFFFFFFFFu;
    if (ox < 0 || oy < 0 || ox >= (int)s->xres || oy >= (int)s->yres) return 0xFFFFFFFFu;
    uint32_t packed = s->pmap[(uint32_t)oy * s->xres + (uint32_t)ox];
    if (packed == 0) return 0xFFFFFFFFu;
    uint32_t nid = (packed >> 9) - 1;
    if (nid >= s->slot_bound || s->part_ty[nid] != t || !(s->part_temp[i] > s->part_temp[nid])) return 0xFFFFFFFFu;
    return nid;
}

/* Rule R6: one-round priority matching of the convection proposals. */
static void convection_swaps(oracle_state *s, uint32_t seed) {
    const uint32_t n = s->slot_bound;
    uint32_t *tgt = malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t t = s->part_ty[i];
        uint32_t roll = oracle_gpu_hash(seed ^ (i * PARTICLE_HASH_SALT) ^ t ^ TRANSITION_HASH_SALT);
        tgt[i] = convect_target(s, i, roll);
    }
    /* min proposer slot touching each particle */
    uint32_t *touch = malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) touch[i] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < n; ++i) {
        if (tgt[i] == 0xFFFFFFFFu) continue;
        if (i < touch[i]) touch[i] = i;
        if (i < touch[tgt[i]]) touch[tgt[i]] = i;
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t t = tgt[i];
        if (t == 0xFFFFFFFFu || touch[i] != i || touch[t] != i) continue;
        float hot = s->part_temp[i];
        s->part_temp[i] = s->part_temp[t];
        s->part_temp[t] = hot;
    }
    free(tgt);
    free(touch);
}

static float pres_sum(const oracle_state *s, int x, int y, const float *pv0) {
    uint32_t xc = s->xres / 4;
    uint32_t cy = (uint32_t)y / 4, cx = (uint32_t)x / 4;
    uint32_t ym = (uint32_t)(y - 2 > 0 ? y - 2 : 0) / 4, xm = (uint32_t)(x - 2 > 0 ? x - 2 : 0) / 4;
    uint32_t yp = (uint32_t)(y + 2 < (int)s->yres - 1 ? y + 2 : (int)s->yres - 1) / 4;
    uint32_t xp = (uint32_t)(x + 2 < (int)s->xres - 1 ? x + 2 : (int)s->xres - 1) / 4;
    float v = (pv0[cy * xc + cx] + pv0[ym * xc + cx] + pv0[yp * xc + cx] + pv0[cy * xc + xm] + pv0[cy * xc + xp]) * 2.0f;
    return v > 0.0f ? v : 0.0f;
}

/* phase_transition_kernel body after the convection swap (thermal_kernel.rs:822-1315). */
static void transition_one(oracle_state *s, uint32_t i, uint32_t seed, int legacy, const float *pv0,
                           const uint32_t *emap0, float *pv_out, uint32_t *pv_key, uint32_t *spark) {
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4, yc = yres / 4, rows = 256;
    uint32_t t = s->part_ty[i];
    if (t == 0 || t >= rows) return;
    int x = ora_f2i(s->part_x[i]), y = ora_f2i(s->part_y[i]);
    if (x < 0 || x >= (int)xres || y < 0 || y >= (int)yres) return;
    uint32_t ci = ((uint32_t)y / 4) * xc + (uint32_t)x / 4;
    float pv = pv0[ci];
    uint32_t roll = oracle_gpu_hash(seed ^ (i * PARTICLE_HASH_SALT) ^ t ^ TRANSITION_HASH_SALT);
    if (!legacy) {
        uint32_t props = P(s, t, COL_PROPERTIES);
        uint32_t ctype = s->part_ctype[i];
        int32_t tmp = s->part_tmp[i], tmp2 = s->part_tmp2[i];
        if (s->part_equil[i] >= 0.0f) {
            float pt = s->part_equil[i];
            if ((t == T_ICEI || t == T_SNOW) && (ctype == 0 || ctype >= rows || ctype == T_ICEI || ctype == T_SNOW)) {
                ctype = T_WATR;
                s->part_ctype[i] = ctype;
            }
            float high_temp = PF(s, t, COL_HIGH_TEMP);
            int32_t high_tr = (int32_t)P(s, t, COL_HIGH_TEMP_TRANSITION);
            float low_temp = PF(s, t, COL_LOW_TEMP);
            int32_t low_tr = (int32_t)P(s, t, COL_LOW_TEMP_TRANSITION);
            float pv_shift = 2.0f * pv;
            int boil = (props & 0x02u) && high_tr >= 0 && high_tr < 256 && (P(s, (uint32_t)high_tr, COL_PROPERTIES) & 0x08u);
            if (t == T_LN2 || t == T_SLTW) boil = 1;
            int cond = (props & 0x08u) && low_tr >= 0 && low_tr < 256 && (P(s, (uint32_t)low_tr, COL_PROPERTIES) & 0x02u);
            if (t == T_WTRV) cond = 1;
            float ctemph = boil ? pt - pv_shift : pt;
            float ctempl = cond ? pt - pv_shift : pt;
            uint32_t old_ty = t, nt = t, new_ctype = ctype;
            int32_t new_life = s->part_life[i], new_tmp = tmp, new_tmp2 = tmp2;
            int sflag = 0;
            if (high_tr != NT && ctemph >= high_temp) {
                if (high_tr != ST) {
                    if (t == T_FOG) new_ctype = 0;
                    nt = (uint32_t)high_tr;
                    sflag = 1;
                } else if (t == T_ICEI || t == T_SNOW) {
                    int melts = 0;
                    if (new_ctype > 0 && new_ctype < 256 && new_ctype != t) {
                        int32_t ct_tr = (int32_t)P(s, new_ctype, COL_LOW_TEMP_TRANSITION);
                        float ct_low = PF(s, new_ctype, COL_LOW_TEMP);
                        if (ct_tr == T_ICEI || ct_tr == T_SNOW) {
                            if (pt >= ct_low) melts = 1;
                        } else if (pt >= 273.15f) {
                            melts = 1;
                        }
                    }
                    if (melts) {
                        nt = new_ctype;
                        new_ctype = 0;
                        new_life = 0;
                        sflag = 1;
                    }
                } else if (t == T_SLTW) {
                    nt = (roll % 4u == 0) ? T_SALT : T_WTRV;
                    sflag = 1;
                } else if (t == T_BRMT) {
                    if (new_ctype == T_TUNG) {
                        if (ctemph >= PF(s, T_TUNG, COL_HIGH_TEMP)) {
                            nt = T_LAVA;
                            old_ty = T_TUNG;
                            sflag = 1;
                        }
                    } else if (ctemph >= high_temp) {
                        nt = T_LAVA;
                        sflag = 1;
                    }
                } else if (t == T_CRMC) {
                    if (ctemph >= pres_sum(s, x, y, pv0) + high_temp) {
                        nt = T_LAVA;
                        sflag = 1;
                    }
                } else if (t == T_RIME) {
                    if (new_tmp > 5) {
                        nt = T_ACID;
                        new_life = 25 + 5 * new_tmp;
                        new_tmp = 0;
                    } else {
                        nt = new_ctype == T_DSTW ? T_DSTW : T_WATR;
                        new_ctype = 0;
                    }
                    sflag = 1;
                }
            } else if (low_tr != NT && ctempl < low_temp) {
                if (low_tr != ST) {
                    nt = (uint32_t)low_tr;
                    sflag = 1;
                } else if (t == T_WTRV) {
                    nt = pt < 273.0f ? T_RIME : T_DSTW;
                    sflag = 1;
                } else if (t == T_LAVA) {
                    if (new_ctype > 0 && new_ctype < 256 && new_ctype != T_LAVA) {
                        float c_high = PF(s, new_ctype, COL_HIGH_TEMP);
                        int32_t c_high_tr = (int32_t)P(s, new_ctype, COL_HIGH_TEMP_TRANSITION);
                        int freeze = 0;
                        if (new_ctype == T_THRM) {
                            if (pt < PF(s, T_BMTL, COL_HIGH_TEMP)) freeze = 1;
                        } else if (new_ctype == T_VIBR || new_ctype == T_BVBR) {
                            if (pt < 273.15f) freeze = 1;
                        } else if (new_ctype == T_TUNG) {
                            if (pt < c_high) freeze = 1;
                        } else if (new_ctype == T_CRMC) {
                            freeze = ctemph >= pres_sum(s, x, y, pv0) + c_high ? 0 : 1;
                        } else if (c_high_tr == T_LAVA || new_ctype == T_HEAC) {
                            if (pt < c_high && new_ctype != T_BISM) freeze = 1;
                        } else if (pt < 973.0f) {
                            freeze = 1;
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
                            new_ctype = 0;
                            sflag = 1;
                        }
                    } else if (pt < 973.0f) {
                        nt = 5; /* PT_STNE */
                        sflag = 1;
                    }
                }
            }
            if (sflag) {
                if (nt == T_ICEI || nt == T_LAVA || nt == T_SNOW) new_ctype = old_ty;
                if (nt == T_RIME) new_ctype = T_DSTW;
                if (!(nt == T_ICEI && new_ctype == T_FRZW) && nt != T_ACID) new_life = 0;
                if (nt == T_FIRE) {
                    new_tmp = 0;
                    if (old_ty == T_SEED) {
                        new_ctype = 0;
                        new_tmp2 = 0;
                    }
                }
                if ((P(s, nt, COL_PROPERTIES) & 0x08u) && !(P(s, old_ty, COL_PROPERTIES) & 0x08u)) pv += 0.50f;
                if (nt == 0) {
                    s->part_ty[i] = 0;
                    t = 0;
                } else {
                    s->part_ty[i] = nt;
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
                s->part_ctype[i] = new_ctype;
                s->part_life[i] = new_life;
                s->part_tmp[i] = new_tmp;
                s->part_tmp2[i] = new_tmp2;
            }
            float pt2 = ora_clampf(pt, 0.0f, 9999.0f);
            if (t == T_LAVA) {
                int32_t ll = ora_clampi(ora_f2i((pt2 - 700.0f) / 7.0f), 0, 400);
                s->part_life[i] = ll;
                if (s->part_ctype[i] == T_THRM && s->part_tmp[i] > 0) {
                    s->part_tmp[i] -= 1;
                    pt2 = 3500.0f;
                }
                if (s->part_ctype[i] == T_PLUT && s->part_tmp[i] > 0) {
                    s->part_tmp[i] -= 1;
                    pt2 = 9999.0f;
                }
                s->part_temp[i] = pt2;
            }
        } else {
            s->part_temp[i] = ora_clampf(s->part_temp[i], 0.0f, 9999.0f);
        }
    }
    if (t != 0 && t < rows) {
        if (t == T_LIFE) s->part_temp[i] = ora_clampf(s->part_temp[i] - 50.0f, 0.0f, 9999.0f);
        uint32_t tprops = P(s, t, COL_PROPERTIES);
        if ((tprops & 0x20u) || t == T_SPRK) {
            int nx = x % 4, ny = y % 4;
            nx = nx == 0 ? x / 4 - 1 : (nx == 3 ? x / 4 + 1 : x / 4);
            ny = ny == 0 ? y / 4 - 1 : (ny == 3 ? y / 4 + 1 : y / 4);
            if (nx >= 0 && ny >= 0 && nx < (int)xc && ny < (int)yc) {
                uint32_t wi = (uint32_t)ny * xc + (uint32_t)nx;
                if (t != T_SPRK) {
                    if ((emap0[wi] & 0xFFu) == 12u && s->part_life[i] == 0 && (s->wall_bmap[wi] & 0xFFu) != 18u) {
                        s->part_ty[i] = T_SPRK;
                        s->part_life[i] = 4;
                        s->part_ctype[i] = t;
                        t = T_SPRK;
                    }
                } else {
                    uint32_t wall = s->wall_bmap[wi] & 0xFFu;
                    /* WL_DETECT, WL_EWALL, WL_ALLOWLIQUID, WL_WALLELEC, WL_ALLOWALLELEC, WL_EHOLE */
                    if (wall == 3u || wall == 2u || wall == 6u || wall == 1u || wall == 11u || wall == 12u) spark[wi] = 1;
                }
            }
        }
        int32_t explosive = (int32_t)P(s, t, COL_EXPLOSIVE);
        if ((explosive & 2) && pv > 2.5f) {
            float fire_temp = PF(s, T_FIRE, COL_DEFAULT_TEMP);
            int32_t flam = (int32_t)P(s, t, COL_FLAMMABLE);
            s->part_temp[i] = ora_clampf(fire_temp + (float)(flam / 2), 0.0f, 9999.0f);
            s->part_life[i] = 180 + (int32_t)(roll % 80u);
            s->part_ty[i] = T_FIRE;
            t = T_FIRE;
            pv += 0.25f;
        }
        int sflag = 0;
        uint32_t nt = t;
        float gx = s->grav_fx[ci], gy = s->grav_fy[ci];
        float gravtot = (gx < 0.0f ? 0.0f - gx : gx) + (gy < 0.0f ? 0.0f - gy : gy);
        int32_t hpt = (int32_t)P(s, t, COL_HIGH_PRESSURE_TRANSITION);
        float hpv = PF(s, t, COL_HIGH_PRESSURE);
        int32_t lpt = (int32_t)P(s, t, COL_LOW_PRESSURE_TRANSITION);
        float lpv = PF(s, t, COL_LOW_PRESSURE);
        int32_t tmp = s->part_tmp[i];
        if (hpt != NT && pv > hpv) {
            if (hpt != ST) {
                nt = (uint32_t)hpt;
                sflag = 1;
            } else if (t == T_BMTL && (pv > 2.5f || (pv > 1.0f && tmp == 1))) {
                nt = T_BRMT;
                sflag = 1;
            }
        } else if (lpt != NT && pv < lpv && gravtot <= lpv / 4.0f) {
            if (lpt != ST) {
                nt = (uint32_t)lpt;
                sflag = 1;
            }
        } else if (hpt != NT && gravtot > hpv / 4.0f) {
            if (hpt != ST) {
                nt = (uint32_t)hpt;
                sflag = 1;
            } else if (t == T_BMTL && (gravtot > 0.625f || (gravtot > 0.25f && tmp == 1))) {
                nt = T_BRMT;
                sflag = 1;
            }
        }
        if (sflag) {
            if (nt == 0) {
                s->part_ty[i] = 0;
            } else {
                s->part_life[i] = 0;
                if (nt == T_BRMT) {
                    s->part_ctype[i] = 0;
                    s->part_tmp[i] = 0;
                    s->part_tmp2[i] = 0;
                }
                s->part_ty[i] = nt;
                if (nt == T_FIRE) s->part_life[i] = 120 + (int32_t)(roll % 50u);
            }
        }
    }
    if (pv != pv0[ci] && i + 1 > pv_key[ci]) {
        pv_key[ci] = i + 1;
        pv_out[ci] = pv;
    }
}

/* refresh_pmap_types_kernel (thermal_kernel.rs:1338-1387), pmap half. */
void oracle_refresh_pmap_types(oracle_state *s) {
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= 256) continue;
        int x = ora_f2i(s->part_x[i]), y = ora_f2i(s->part_y[i]);
        if (x < 0 || x >= (int)s->xres || y < 0 || y >= (int)s->yres) continue;
        uint32_t pidx = (uint32_t)y * s->xres + (uint32_t)x;
        uint32_t packed = s->pmap[pidx];
        if ((packed >> 9) == i + 1 && (packed & 0x1FFu) != ty) s->pmap[pidx] = ((i + 1) << 9) | (ty & 0x1FFu);
    }
}

void oracle_phase_transitions(oracle_state *s, uint32_t seed, int legacy) {
    const uint32_t n_cells = (s->xres / 4) * (s->yres / 4);
    if (!legacy) convection_swaps(s, seed);
    float *pv0 = malloc(sizeof(float) * n_cells), *pv_out = malloc(sizeof(float) * n_cells);
    uint32_t *emap0 = malloc(sizeof(uint32_t) * n_cells), *pv_key = calloc(n_cells, 4), *spark = calloc(n_cells, 4);
    memcpy(pv0, s->air_pv, sizeof(float) * n_cells);
    memcpy(emap0, s->wall_emap, sizeof(uint32_t) * n_cells);
    for (uint32_t i = 0; i < s->slot_bound; ++i) transition_one(s, i, seed, legacy, pv0, emap0, pv_out, pv_key, spark);
    for (uint32_t c = 0; c < n_cells; ++c) {
        if (pv_key[c]) s->air_pv[c] = pv_out[c];
        if (spark[c]) s->wall_emap[c] = 16u;
    }
    free(pv0); free(pv_out); free(emap0); free(pv_key); free(spark);
}

/* step_thermal_and_transitions_gpu (thermal_kernel.rs:1574-1717). */
void oracle_step_thermal(oracle_state *s, uint32_t seed, int legacy) {
    if (!legacy) {
        oracle_thermal_conduction(s, seed);
        oracle_thermal_apply(s);
    }
    oracle_phase_transitions(s, seed, legacy);
    oracle_refresh_pmap_types(s);
}
