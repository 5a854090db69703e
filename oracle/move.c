/*
 * Oracle: lock-free movement pass -- TEST INFRASTRUCTURE ONLY (see tpt_oracle.h).
 *
 * Restates crates/tpt-sim/src/move_kernel/: the driver (mod.rs:1176-1227),
 * air_deposit_kernel (mod.rs:297-346), the eval/do_move helpers
 * (mod.rs:838-1151), velocity_and_intent_kernel and its stages
 * (velocity.rs:98-262, 649-1157), the Falldown >= 1 chain (liquid.rs:137-963),
 * commit_move_kernel (commit.rs:76-535) and the map rebuild
 * (commit.rs:541-603).
 *
 * Canonical rules (shared with the HIP kernels):
 *  R1  air deposit: all eligible particles read the pre-pass air velocity;
 *      the cell keeps the write of its highest eligible slot.
 *  R7  map rebuild: the highest matter slot owns a pixel, INVIS/FILT only
 *      take an empty one (the lowest such slot), energy: highest slot.
 *  SAW the WOOD -> SAWD conversion of a rejected probe (mod.rs:213-232) is
 *      applied after the intent pass (the reference writes it mid-pass).
 *  R10 blocked-photon hits are applied after the commit loop, in ascending
 *      photon slot order per hit particle.
 *  DET WL_DETECT charging: the union of the runs through every triggered
 *      cell, computed from the pre-pass emap.
 *  R11/R12 the e == 2 pair pass and its creations (see pair_resolve).
 * Also restated: photon refraction at glass (velocity.rs:262-647), the
 * e == 2 pair pass (reaction_kernel/energy_entry.rs:90-393) with the drain of
 * its photons, WL_DETECT charging and its emap closure (mod.rs:473-828).
 */
#include "tpt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_util.h"

#define NO_INTENT 0xFFFFFFFFu
#define DIE_INTENT 0xFFFFFFFEu
#define INTENT_HALF_MIRROR 0x40000000u
#define INTENT_PIXEL_MASK 0x3FFFFFFFu
#define DO_BLOCKED 0u
#define DO_OK 1u
#define DO_DIE 2u
#define ISTP 2.0f
#define MAX_VELOCITY 1.0e4f
#define MAX_PLAN_STEPS 5001u
#define EDGE_VOID 0u
#define EDGE_LOOP 2u

typedef struct {
    const oracle_state *s;
    int xr, yr;
    uint32_t xc, edge;
    uint8_t *saw; /* deferred WOOD -> SAWD (per slot) */
} mctx;

/* mod.rs:1147-1151 */
uint32_t oracle_pcg_hash(uint32_t seed) {
    uint32_t state = seed * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

static uint32_t can_move_at(const oracle_state *s, uint32_t mty, uint32_t dty) {
    return s->move_lut[256 * 12 + mty * 256 + dty];
}

/* mod.rs:1079-1104 */
static int wall_blocks_particle(uint32_t wall, uint32_t emap, uint32_t props) {
    if (wall == 8 || wall == 1 || wall == 9) return 1;
    if (wall == 6) return (props & 0x2u) == 0;
    if (wall == 10) return (props & 0x1u) == 0;
    if (wall == 13) return (props & 0x8u) == 0;
    if (wall == 15) return (props & 0x10u) == 0;
    if (wall == 2) return (emap & 0xFFu) == 0;
    if (wall == 3) return (props & 0x4u) != 0;
    return 0;
}

/* mod.rs:1114-1131 */
static int wall_kills_particle(uint32_t wall, uint32_t emap, uint32_t props) {
    if (wall == 8 || wall == 1 || wall == 9 || wall == 7) return 1;
    if (wall == 6) return (props & 0x2u) == 0;
    if (wall == 10) return (props & 0x1u) == 0;
    if (wall == 13) return (props & 0x8u) == 0;
    if (wall == 15) return (props & 0x10u) == 0;
    if (wall == 2) return (emap & 0xFFu) == 0;
    return 0;
}

/* mod.rs:838-891 */
static uint32_t eval_move_at(const mctx *m, uint32_t props, uint32_t ty, int nx, int ny) {
    const oracle_state *s = m->s;
    if (!(nx >= 0 && nx < m->xr && ny >= 0 && ny < m->yr)) return 0;
    uint32_t d_ty = s->pmap[(uint32_t)(ny * m->xr + nx)] & 0x1FFu;
    uint32_t wcell = ((uint32_t)ny / 4) * m->xc + (uint32_t)nx / 4;
    uint32_t wall = s->wall_bmap[wcell], emap = s->wall_emap[wcell];
    uint32_t can = can_move_at(s, ty, d_ty);
    if (can == 3) {
        if (d_ty == 23 || d_ty == 75) can = 1;                                     /* VOID, PVOD */
        else if (d_ty == 58 || d_ty == 43 || d_ty == 124 || d_ty == 60) can = 2;    /* LCRY GPMP INVIS SWCH */
        else can = 1;
    }
    uint32_t result = can;
    if (wall != 0) {
        if (wall_blocks_particle(wall, emap, props)) {
            result = 0;
        } else if (wall == 12 && (emap & 0xFFu) == 0) {
            uint32_t dest_props = s->move_lut[d_ty * 12];
            if ((props & 0x4u) == 0 && (dest_props & 0x4u) == 0) result = 2;
        }
    }
    return result;
}

/* mod.rs:959-967 */
static float remainder_p(float x, float y) {
    float out = x - floorf(x / y) * y;
    if (out < 0.0f) out += y;
    else if (out >= y) out -= y;
    return out;
}

/* mod.rs:976-984 */
static int wrap_loop_pixel(int n, uint32_t res) {
    int lo = 4, span = (int)res - 8;
    if (span > 0 && (n < lo || n >= lo + span)) return lo + ora_f2i(remainder_p((float)(n - lo), (float)span));
    return n;
}

/* mod.rs:1009-1056 */
static uint32_t do_move_at(const mctx *m, uint32_t props, uint32_t ty, int nx, int ny, int *ox, int *oy) {
    int mx = nx, my = ny;
    if (m->edge == EDGE_LOOP) {
        mx = wrap_loop_pixel(nx, (uint32_t)m->xr);
        my = wrap_loop_pixel(ny, (uint32_t)m->yr);
    }
    uint32_t status = DO_BLOCKED;
    if (mx < 0 || my < 0 || mx >= m->xr || my >= m->yr) {
        if (m->edge != EDGE_LOOP) status = DO_DIE;
    } else if (eval_move_at(m, props, ty, mx, my) != 0) {
        int in_band = mx >= 4 && my >= 4 && mx < m->xr - 4 && my < m->yr - 4;
        status = DO_OK;
        if (!in_band && m->edge != EDGE_LOOP) {
            status = DO_BLOCKED;
            if (m->edge == EDGE_VOID) status = DO_DIE;
        }
    }
    *ox = mx;
    *oy = my;
    return status;
}

/* mod.rs:213-232, deferred (rule SAW) */
static void saw_wood_at(mctx *m, int nx, int ny, float vx, float vy) {
    const oracle_state *s = m->s;
    if (nx >= 0 && ny >= 0 && nx < m->xr && ny < m->yr) {
        uint32_t v = s->pmap[(uint32_t)(ny * m->xr + nx)];
        if (v >= 512u && (v & 0x1FFu) == 15u && vx * vx + vy * vy > 25.0f) m->saw[(v >> 9) - 1] = 1;
    }
}

static int round_i32(float v) { return ora_f2i(v + 0.5f); }

/* liquid.rs:80-118 */
static int plan_sample_blocked(const mctx *m, uint32_t props, uint32_t ty, int nx, int ny, int closed_start) {
    const oracle_state *s = m->s;
    if (nx < 0 || ny < 0 || nx >= m->xr || ny >= m->yr) return 1;
    uint32_t packed = s->pmap[(uint32_t)(ny * m->xr + nx)];
    uint32_t d_ty = packed & 0x1FFu;
    uint32_t raw_can = can_move_at(s, ty, d_ty);
    uint32_t eval = eval_move_at(m, props, ty, nx, ny);
    uint32_t wcell = ((uint32_t)ny / 4) * m->xc + (uint32_t)nx / 4;
    uint32_t wall = s->wall_bmap[wcell];
    int closed_now = wall == 12 && (s->wall_emap[wcell] & 0xFFu) == 0;
    return eval == 0 || (raw_can == 3 && eval == 1) || (ty == 29 && packed != 0) || wall == 7 || closed_now != closed_start;
}

typedef struct {
    int fin_x, fin_y, clear_x, clear_y;
    float clear_xf, clear_yf;
} plan_t;

static float wrap_band(float f, int res) {
    int ok = f >= 4.0f - 0.5f && f < (float)res - 4.0f - 0.5f;
    if (!ok) f = remainder_p(f - 4.0f + 0.5f, (float)res - 8.0f) + 4.0f - 0.5f;
    return f;
}

/* liquid.rs:137-261 */
static plan_t plan_move(const mctx *m, float vx, float vy, float px, float py, int xi, int yi, uint32_t props, uint32_t ty) {
    const oracle_state *s = m->s;
    float mv = fabsf(vx);
    if (fabsf(vy) > mv) mv = fabsf(vy);
    plan_t p;
    p.clear_xf = px;
    p.clear_yf = py;
    float fin_xf = px + vx, fin_yf = py + vy;
    p.fin_x = round_i32(fin_xf);
    p.fin_y = round_i32(fin_yf);
    p.clear_x = xi;
    p.clear_y = yi;
    if (mv >= ISTP) {
        uint32_t sc = ((uint32_t)yi / 4) * m->xc + (uint32_t)xi / 4;
        int start_closed = s->wall_bmap[sc] == 12 && (s->wall_emap[sc] & 0xFFu) == 0;
        float step_xf = vx * ISTP / mv, step_yf = vy * ISTP / mv;
        float walk_xf = px, walk_yf = py, remaining = mv;
        int planned = 0;
        uint32_t guard = 0;
        while (!planned && guard < MAX_PLAN_STEPS) {
            guard++;
            remaining -= ISTP;
            walk_xf += step_xf;
            walk_yf += step_yf;
            if (m->edge == EDGE_LOOP) {
                walk_xf = wrap_band(walk_xf, m->xr);
                walk_yf = wrap_band(walk_yf, m->yr);
            }
            int nfx = round_i32(walk_xf), nfy = round_i32(walk_yf);
            if (remaining <= 0.0f) {
                fin_xf = px + vx;
                fin_yf = py + vy;
                if (m->edge == EDGE_LOOP) {
                    fin_xf = wrap_band(fin_xf, m->xr);
                    fin_yf = wrap_band(fin_yf, m->yr);
                }
                p.fin_x = round_i32(fin_xf);
                p.fin_y = round_i32(fin_yf);
                p.clear_xf = fin_xf - step_xf;
                p.clear_yf = fin_yf - step_yf;
                p.clear_x = round_i32(p.clear_xf);
                p.clear_y = round_i32(p.clear_yf);
                planned = 1;
            } else if (plan_sample_blocked(m, props, ty, nfx, nfy, start_closed)) {
                p.fin_x = nfx;
                p.fin_y = nfy;
                p.clear_xf = walk_xf - step_xf;
                p.clear_yf = walk_yf - step_yf;
                p.clear_x = round_i32(p.clear_xf);
                p.clear_y = round_i32(p.clear_yf);
                planned = 1;
            }
        }
    }
    return p;
}

/* velocity.rs:133-170 */
static void base_gravity(const oracle_state *s, float gravity, float px, float py, float *gx, float *gy) {
    *gx = 0.0f;
    *gy = 0.0f;
    if (s->gravity_mode == 0) {
        *gy = gravity;
    } else if (s->gravity_mode == 2) {
        float dx = px - (float)(s->xres / 2), dy = py - (float)(s->yres / 2);
        if (dx != 0.0f || dy != 0.0f) {
            float d = 0.01f - sqrtf(dx * dx + dy * dy);
            *gx = gravity * (dx / d);
            *gy = gravity * (dy / d);
        }
    } else if (s->gravity_mode == 3) {
        *gx = gravity * s->custom_gravity_x;
        *gy = gravity * s->custom_gravity_y;
    }
}

/* velocity.rs:98-111 */
static int step_axis(float v, uint32_t h) {
    float a = v < 0.0f ? -v : v;
    float r = (float)(h & 0xFFFFu) * (1.0f / 65535.0f);
    if (a >= 1.0f || r < a) {
        if (v > 0.0f) return 1;
        if (v < 0.0f) return -1;
    }
    return 0;
}

/* liquid.rs:264-286 */
static void liquid_neighbourhood(const mctx *m, uint32_t ty, int xi, int yi, uint32_t *nt, uint32_t *space) {
    *nt = 0;
    *space = 0;
    for (int oy = 0; oy < 3; ++oy)
        for (int ox = 0; ox < 3; ++ox) {
            if (ox == 1 && oy == 1) continue;
            int nx = xi + ox - 1, ny = yi + oy - 1;
            uint32_t nty = 0;
            if (nx >= 0 && nx < m->xr && ny >= 0 && ny < m->yr) nty = m->s->pmap[(uint32_t)(ny * m->xr + nx)] & 0x1FFu;
            if (nty == 0) *space += 1;
            if (nty != ty) *nt += 1;
        }
}

static uint32_t pix(const mctx *m, int x, int y) { return (uint32_t)(y * m->xr + x); }

/* liquid.rs:298-350 */
static uint32_t diagonal_liquid_move(const mctx *m, float vx, float vy, float cxf, float cyf, int rdir, uint32_t props, uint32_t ty) {
    float ddx = vx - vy * (float)rdir;
    float ddy = vy + vx * (float)rdir;
    float dmv = fabsf(ddx);
    if (fabsf(ddy) > dmv) dmv = fabsf(ddy);
    if (!(dmv > 0.0001f)) return NO_INTENT;
    ddx /= dmv;
    ddy /= dmv;
    int ox, oy;
    uint32_t d1 = do_move_at(m, props, ty, round_i32(cxf + ddx), round_i32(cyf + ddy), &ox, &oy);
    if (d1 == DO_DIE) return DIE_INTENT;
    if (d1 == DO_OK) return pix(m, ox, oy);
    uint32_t d2 = do_move_at(m, props, ty, round_i32(cxf + ddy * (float)rdir), round_i32(cyf - ddx * (float)rdir), &ox, &oy);
    if (d2 == DO_DIE) return DIE_INTENT;
    if (d2 == DO_OK) return pix(m, ox, oy);
    return NO_INTENT;
}

/* liquid.rs:630-733 */
static uint32_t scan_liquid_vertical(const mctx *m, int xi, int yi, int fin_y, int clear_x, int clear_y, int rdir, int rt,
                                     uint32_t ty, uint32_t props, int *fall) {
    const oracle_state *s = m->s;
    int scan_x = xi, scan_y = yi, hit = 0, die = 0, stop = 0, steps = 0;
    int j = clear_x + rdir;
    *fall = 0;
    while (!stop && steps < rt) {
        if (j < 0 || j >= m->xr || j < clear_x - rt || j >= clear_x + rt) {
            stop = 1;
        } else {
            if (fin_y >= 0 && fin_y < m->yr) {
                uint32_t d_ty = s->pmap[pix(m, j, fin_y)] & 0x1FFu;
                uint32_t wcell = ((uint32_t)fin_y / 4) * m->xc + (uint32_t)j / 4;
                if (d_ty != ty || s->wall_bmap[wcell] != 0) {
                    int ox, oy;
                    uint32_t d = do_move_at(m, props, ty, j, fin_y, &ox, &oy);
                    if (d == DO_DIE) {
                        die = 1;
                        stop = 1;
                    } else if (d == DO_OK) {
                        hit = 1;
                        *fall = 1;
                        scan_x = ox;
                        scan_y = oy;
                        stop = 1;
                    }
                }
            }
            if (!stop && fin_y != clear_y && clear_y >= 0 && clear_y < m->yr) {
                uint32_t d_ty = s->pmap[pix(m, j, clear_y)] & 0x1FFu;
                uint32_t wcell = ((uint32_t)clear_y / 4) * m->xc + (uint32_t)j / 4;
                if (d_ty != ty || s->wall_bmap[wcell] != 0) {
                    int ox, oy;
                    uint32_t d = do_move_at(m, props, ty, j, clear_y, &ox, &oy);
                    if (d == DO_DIE) {
                        die = 1;
                        stop = 1;
                    } else if (d == DO_OK) {
                        hit = 1;
                        *fall = 1;
                        scan_x = ox;
                        scan_y = oy;
                        stop = 1;
                    }
                }
            }
            if (!stop && clear_y >= 0 && clear_y < m->yr) {
                uint32_t d_ty = s->pmap[pix(m, j, clear_y)] & 0x1FFu;
                uint32_t wall = s->wall_bmap[((uint32_t)clear_y / 4) * m->xc + (uint32_t)j / 4];
                if (d_ty != ty || (wall != 0 && wall != 4)) stop = 1;
            }
            if (!stop) {
                j += rdir;
                steps++;
            }
        }
    }
    if (die) return DIE_INTENT;
    if (hit) return pix(m, scan_x, scan_y);
    return NO_INTENT;
}

/* liquid.rs:735-787 */
static uint32_t scroll_liquid_down(const mctx *m, int scan_x, int scan_y, float vy, int rt, uint32_t ty, uint32_t props) {
    const oracle_state *s = m->s;
    uint32_t target = pix(m, scan_x, scan_y);
    int vr = vy > 0.0f ? 1 : -1;
    int jj = scan_y + vr, stop = 0, steps = 0;
    while (!stop && steps < rt) {
        if (jj < 0 || jj >= m->yr || jj < scan_y - rt || jj >= scan_y + rt) {
            stop = 1;
        } else {
            uint32_t d_ty = s->pmap[pix(m, scan_x, jj)] & 0x1FFu;
            uint32_t wall = s->wall_bmap[((uint32_t)jj / 4) * m->xc + (uint32_t)scan_x / 4];
            if (d_ty != ty || wall != 0) {
                int ox, oy;
                uint32_t d = do_move_at(m, props, ty, scan_x, jj, &ox, &oy);
                if (d == DO_DIE) {
                    target = DIE_INTENT;
                    stop = 1;
                } else if (d == DO_OK) {
                    target = pix(m, ox, oy);
                    stop = 1;
                } else if (d_ty != ty || wall != 4) {
                    stop = 1;
                }
            }
            if (!stop) {
                jj += vr;
                steps++;
            }
        }
    }
    return target;
}

/* liquid.rs:801-963 */
static uint32_t scan_liquid_perpendicular(const mctx *m, float cxf, float cyf, int cx, int cy, int rdir, int rt, float gravity,
                                          uint32_t props, uint32_t ty) {
    const oracle_state *s = m->s;
    uint32_t target = NO_INTENT;
    float nxf = cxf, nyf = cyf, pgx = 0.0f, pgy = 0.0f;
    int nx = cx, ny = cy, moved = 0, j = 0, stop = 0;
    while (!stop && j < rt) {
        float gx, gy;
        base_gravity(s, gravity, (float)nx, (float)ny, &gx, &gy);
        float gmv = fabsf(gx);
        if (fabsf(gy) > gmv) gmv = fabsf(gy);
        if (gmv < 0.0001f) {
            stop = 1;
        } else {
            gx /= gmv;
            gy /= gmv;
            if (j > 0) {
                nxf += (float)rdir * (gy * 2.0f - pgy);
                nyf += -(float)rdir * (gx * 2.0f - pgx);
            } else {
                nxf += (float)rdir * gy;
                nyf += -(float)rdir * gx;
            }
            pgx = gx;
            pgy = gy;
            nx = round_i32(nxf);
            ny = round_i32(nyf);
            if (nx < 0 || ny < 0 || nx >= m->xr || ny >= m->yr) {
                stop = 1;
            } else {
                uint32_t d_ty = s->pmap[pix(m, nx, ny)] & 0x1FFu;
                uint32_t wall = s->wall_bmap[((uint32_t)ny / 4) * m->xc + (uint32_t)nx / 4];
                if (d_ty != ty || wall != 0) {
                    int ox, oy;
                    uint32_t d = do_move_at(m, props, ty, nx, ny, &ox, &oy);
                    if (d == DO_DIE) {
                        target = DIE_INTENT;
                        stop = 1;
                    } else if (d == DO_OK) {
                        target = pix(m, ox, oy);
                        moved = 1;
                        stop = 1;
                    } else if (d_ty != ty || wall != 4) {
                        stop = 1;
                    }
                }
            }
        }
        if (!stop) j++;
    }
    if (moved) {
        int j2 = 0, stop2 = 0;
        while (!stop2 && j2 < rt) {
            float gx, gy;
            base_gravity(s, gravity, (float)nx, (float)ny, &gx, &gy);
            float gmv = fabsf(gx);
            if (fabsf(gy) > gmv) gmv = fabsf(gy);
            if (gmv < 0.0001f) {
                stop2 = 1;
            } else {
                gx /= gmv;
                gy /= gmv;
                nxf += gx;
                nyf += gy;
                int nnx = round_i32(nxf), nny = round_i32(nyf);
                if (nnx < 0 || nny < 0 || nnx >= m->xr || nny >= m->yr) {
                    stop2 = 1;
                } else {
                    uint32_t d_ty = s->pmap[pix(m, nnx, nny)] & 0x1FFu;
                    uint32_t wall = s->wall_bmap[((uint32_t)nny / 4) * m->xc + (uint32_t)nnx / 4];
                    if (d_ty != ty || wall != 0) {
                        int ox, oy;
                        uint32_t d = do_move_at(m, props, ty, nnx, nny, &ox, &oy);
                        if (d == DO_DIE) {
                            target = DIE_INTENT;
                            stop2 = 1;
                        } else if (d == DO_OK) {
                            target = pix(m, ox, oy);
                            stop2 = 1;
                        } else if (d_ty != ty || wall != 4) {
                            stop2 = 1;
                        }
                    }
                }
            }
            if (!stop2) j2++;
        }
    }
    return target;
}

/* liquid.rs:388-617 */
static uint32_t fall_down_destination(mctx *m, float *pvx, float *pvy, float px, float py, int xi, int yi, int sx, int sy,
                                      uint32_t props, uint32_t ty, uint32_t falldown, uint32_t flags, float collision,
                                      uint32_t h2, float gravity, int32_t gel_tmp) {
    const oracle_state *s = m->s;
    float vx = *pvx, vy = *pvy;
    plan_t p = plan_move(m, vx, vy, px, py, xi, yi, props, ty);
    if (p.fin_x == xi && p.fin_y == yi && (sx != 0 || sy != 0)) {
        p.fin_x = xi + sx;
        p.fin_y = yi + sy;
        p.clear_x = xi;
        p.clear_y = yi;
        p.clear_xf = px;
        p.clear_yf = py;
    }
    uint32_t target = NO_INTENT;
    if (p.fin_x != xi || p.fin_y != yi) {
        int ox, oy;
        uint32_t d0 = do_move_at(m, props, ty, p.fin_x, p.fin_y, &ox, &oy);
        if (d0 == DO_DIE) target = DIE_INTENT;
        else if (d0 == DO_OK) target = pix(m, ox, oy);
        else saw_wood_at(m, ox, oy, vx, vy);
        if (target == NO_INTENT && p.fin_x != xi) {
            uint32_t d1 = do_move_at(m, props, ty, p.fin_x, p.clear_y, &ox, &oy);
            if (d1 == DO_DIE) target = DIE_INTENT;
            else if (d1 == DO_OK) target = pix(m, ox, oy);
            else saw_wood_at(m, ox, oy, vx, vy);
        }
        if (target == NO_INTENT && p.fin_y != yi) {
            uint32_t d2 = do_move_at(m, props, ty, p.clear_x, p.fin_y, &ox, &oy);
            if (d2 == DO_DIE) target = DIE_INTENT;
            else if (d2 == DO_OK) target = pix(m, ox, oy);
            else saw_wood_at(m, ox, oy, vx, vy);
        }
        if (target == NO_INTENT) {
            uint32_t nt, space;
            liquid_neighbourhood(m, ty, xi, yi, &nt, &space);
            int rdir = (h2 & 2u) == 0 ? 1 : -1;
            int clear_moved = p.clear_x != xi || p.clear_y != yi;
            int moving = fabsf(vx) > 0.01f || fabsf(vy) > 0.01f;
            if ((clear_moved || nt > 0 || space > 0) && moving)
                target = diagonal_liquid_move(m, vx, vy, p.clear_xf, p.clear_yf, rdir, props, ty);
            if (target == NO_INTENT) {
                if (falldown > 1) {
                    if (s->newton_grav == 0 && s->gravity_mode == 0 && vy > fabsf(vx)) {
                        int rt = ((flags & 0x01u) && nt == 0) ? 10 : 30;
                        if (ty == 150) rt = ora_f2i((float)gel_tmp * 0.20f + 5.0f); /* GEL */
                        int fall;
                        uint32_t st = scan_liquid_vertical(m, xi, yi, p.fin_y, p.clear_x, p.clear_y, rdir, rt, ty, props, &fall);
                        if (st != NO_INTENT) {
                            if (fall) target = scroll_liquid_down(m, (int)(st % (uint32_t)m->xr), (int)(st / (uint32_t)m->xr), vy, rt, ty, props);
                            else target = st;
                        }
                    } else {
                        float gx, gy;
                        base_gravity(s, gravity, (float)xi, (float)yi, &gx, &gy);
                        float gnorm = fabsf(gx);
                        if (fabsf(gy) > gnorm) gnorm = fabsf(gy);
                        int aligned = 0;
                        if (gnorm > 0.0001f) {
                            float ngx = gx / gnorm, ngy = gy / gnorm;
                            if (fabsf(ngx * vx + ngy * vy) > fabsf(ngy * vx - ngx * vy)) aligned = 1;
                        }
                        if (aligned) {
                            int rt = ((flags & 0x01u) && nt == 0) ? 10 : 30;
                            target = scan_liquid_perpendicular(m, p.clear_xf, p.clear_yf, p.clear_x, p.clear_y, rdir, rt, gravity, props, ty);
                        }
                    }
                }
                if (target == NO_INTENT && clear_moved) {
                    uint32_t d = do_move_at(m, props, ty, p.clear_x, p.clear_y, &ox, &oy);
                    if (d == DO_DIE) target = DIE_INTENT;
                    else if (d == DO_OK) target = pix(m, ox, oy);
                }
            }
            vx *= collision;
            vy *= collision;
        }
    }
    *pvx = vx;
    *pvy = vy;
    return target;
}

/* Photon refraction at glass: velocity.rs:262-647 (Simulation.cpp:1626-1750,
 * 886-922, 2966-3005). Only GLAS and BGLA bend light. */
static int refract_blocking(const mctx *m, int x, int y) {
    if (x < 0 || y < 0 || x >= m->xr || y >= m->yr) return 0;
    uint32_t t = m->s->pmap[y * m->xr + x] & 0x1FFu;
    return t == 24u || t == 95u;
}

static int refract_boundary(const mctx *m, int x, int y) {
    return refract_blocking(m, x, y) &&
           !(refract_blocking(m, x, y - 1) && refract_blocking(m, x, y + 1) && refract_blocking(m, x - 1, y) &&
             refract_blocking(m, x + 1, y));
}

static const int ring_dx[8] = {1, 1, 0, -1, -1, -1, 0, 1};
static const int ring_dy[8] = {0, 1, 1, 1, 0, -1, -1, -1};
static const uint32_t ring_mask[8] = {0x83u, 0x07u, 0x0Eu, 0x1Cu, 0x38u, 0x70u, 0xE0u, 0xC1u};

/* find_next_boundary (velocity.rs:343-389): the next boundary step of a walk
 * around the glass; returns 1 and moves (x, y, em) when one exists. */
static int next_boundary(const mctx *m, int *x, int *y, uint32_t dm, int *em, int reverse) {
    if (!(*x > 0 && *x < m->xr - 1 && *y > 0 && *y < m->yr - 1)) return 0;
    if (*em != -1) dm &= ring_mask[*em];
    uint32_t mask = 0;
    for (int i = 0; i < 8; ++i)
        if ((dm & (1u << i)) && refract_blocking(m, *x + ring_dx[i], *y + ring_dy[i])) mask |= 1u << i;
    for (int i = 0; i < 8; ++i) {
        int n = (i + (reverse ? 1 : 7)) & 7;
        if ((mask & (1u << i)) && !(mask & (1u << n))) {
            *x += ring_dx[i];
            *y += ring_dy[i];
            *em = i;
            return 1;
        }
    }
    return 0;
}

/* get_normal (velocity.rs:401-456) */
static int surface_normal(const mctx *m, int x, int y, float dx, float dy, float *nx, float *ny) {
    if (!((dx != 0.0f || dy != 0.0f) && refract_boundary(m, x, y))) return 0;
    int lx = x, ly = y, rx = x, ry = y, lm = -1, rm = -1, lv = 1, rv = 1, j = 0;
    for (int k = 0; k < 10; ++k) {
        if (lv) lv = next_boundary(m, &lx, &ly, 0xFFu, &lm, 1);
        if (rv) rv = next_boundary(m, &rx, &ry, 0xFFu, &rm, 0);
        j += lv + rv;
        if (!lv && !rv) break;
    }
    if (!(j >= 3 && (lx != rx || ly != ry))) return 0;
    float ex = (float)(rx - lx), ey = (float)(ry - ly);
    float inv = 1.0f / sqrtf(ex * ex + ey * ey);
    *nx = ey * inv;
    *ny = -ex * inv;
    return 1;
}

/* get_normal_interp (velocity.rs:462-516) */
static int surface_normal_along(const mctx *m, float x0, float y0, float dx, float dy, float *nx, float *ny) {
    dx = dx / 16.0f;
    dy = dy / 16.0f;
    for (int k = 0; k < 20; ++k) {
        int x = ora_f2i(x0 + 0.5f), y = ora_f2i(y0 + 0.5f);
        if (x < 0 || y < 0 || x >= m->xr || y >= m->yr) return 0;
        if (refract_boundary(m, x, y)) return surface_normal(m, x, y, dx, dy, nx, ny);
        x0 += dx;
        y0 += dy;
    }
    return 0;
}

/* get_wavelength_bin (velocity.rs:523-560); -1: no wavelength left. */
static int wavelength_bin(uint32_t wm, uint32_t seed, uint32_t *out) {
    uint32_t masked = wm & 0x3FFFFFFFu;
    *out = wm;
    if (!masked) return -1;
    uint32_t lo = 30, hi = 0;
    for (uint32_t i = 0; i < 30; ++i)
        if (masked & (1u << i)) {
            if (i < lo) lo = i;
            if (i > hi) hi = i;
        }
    if (hi - lo < 5u) return (int)(hi + lo);
    uint32_t r = oracle_pcg_hash(seed ^ (lo * 2654435769u));
    uint32_t i = ((r >> 1) % (hi - lo - 4u)) + lo;
    if (r & 1u) {
        *out = wm & (0x1Fu << i);
        return (int)((i + 2u) * 2u);
    }
    *out = wm & (0xFu << i);
    return (int)((i + 2u) * 2u - 1u);
}

/* refract_photon (velocity.rs:577-637): 0 = passes, 1 = dies, 2 = total
 * internal reflection (stays this tick). */
static int refract_photon(mctx *m, float px, float py, int xi, int yi, float *vx, float *vy, int fx, int fy,
                          uint32_t slot, uint32_t seed) {
    const oracle_state *s = m->s;
    uint32_t rt = s->pmap[fy * m->xr + fx] & 0x1FFu, lt = s->pmap[yi * m->xr + xi] & 0x1FFu;
    int r_glass = rt == 24u || rt == 95u, l_glass = lt == 24u || lt == 95u;
    if (r_glass == l_glass) return 0;
    float nrx, nry;
    if (!surface_normal_along(m, px, py, *vx, *vy, &nrx, &nry)) return 1;
    uint32_t narrowed;
    int bin = wavelength_bin(s->part_ctype[slot], seed, &narrowed);
    if (bin == -1) return 1;
    s->part_ctype[slot] = narrowed;
    float nn = 1.9f - 0.07f * ((float)bin - 30.0f) / 30.0f;
    nn *= nn;
    if (r_glass) { /* entering the glass */
        nrx = -nrx;
        nry = -nry;
        nn = 1.0f / nn;
    }
    float ct1 = *vx * nrx + *vy * nry;
    float ct2 = 1.0f - (nn * nn) * (1.0f - (ct1 * ct1));
    if (ct2 < 0.0f) {
        *vx -= 2.0f * ct1 * nrx;
        *vy -= 2.0f * ct1 * nry;
        return 2;
    }
    float ct2s = sqrtf(ct2) - nn * ct1;
    *vx = nn * *vx + ct2s * nrx;
    *vy = nn * *vy + ct2s * nry;
    return 0;
}

/* velocity.rs:755-928 */
static uint32_t gas_destination(mctx *m, float *pvx, float *pvy, float px, float py, int xi, int yi, int sx, int sy,
                                uint32_t props, uint32_t ty, uint32_t slot, float collision, uint32_t seed) {
    const oracle_state *s = m->s;
    float vx = *pvx, vy = *pvy;
    plan_t p = plan_move(m, vx, vy, px, py, xi, yi, props, ty);
    int fin_x = p.fin_x, fin_y = p.fin_y;
    if (fin_x == xi && fin_y == yi && (sx != 0 || sy != 0)) {
        fin_x = xi + sx;
        fin_y = yi + sy;
    }
    uint32_t target = NO_INTENT;
    int blocked = 0, died = 0, tir = 0;
    if (fin_x != xi || fin_y != yi) {
        int ox, oy;
        uint32_t d = do_move_at(m, props, ty, fin_x, fin_y, &ox, &oy);
        if (d == DO_OK) {
            if (ty == 29) {
                int r = refract_photon(m, px, py, xi, yi, &vx, &vy, ox, oy, slot, seed);
                died = r == 1;
                tir = r == 2;
            }
            if (!died && !tir) target = pix(m, ox, oy);
        } else if (d == DO_DIE) {
            died = 1;
        } else {
            blocked = 1;
            if (ty == 29) {
                uint32_t dest_px = pix(m, ox, oy);
                uint32_t dest_ty = s->pmap[dest_px] & 0x1FFu;
                uint32_t src_ty = s->pmap[pix(m, xi, yi)] & 0x1FFu;
                uint32_t roll = oracle_pcg_hash(seed ^ (slot * 0x9E3779B9u) ^ (dest_px * 0x85EBCA6Bu));
                if ((dest_ty == 46 && (roll & 1u) == 0) || src_ty == 46) {
                    blocked = 0;
                    target = INTENT_HALF_MIRROR | dest_px;
                } else {
                    target = 0x80000000u | dest_px;
                    saw_wood_at(m, ox, oy, vx, vy);
                }
            } else {
                saw_wood_at(m, ox, oy, vx, vy);
            }
        }
    }
    if (died) {
        target = DIE_INTENT;
    } else if (blocked) {
        if (ty == 29) {
            if (fin_x != xi) vx *= collision;
            if (fin_y != yi) vy *= collision;
        } else {
            int bx = fin_x, by = fin_y;
            if (bx > xi + 2) bx = xi + 2;
            if (bx < xi - 2) bx = xi - 2;
            if (by > yi + 2) by = yi + 2;
            if (by < yi - 2) by = yi - 2;
            int rx = 2 * xi - bx, ry = 2 * yi - by, ox, oy;
            uint32_t d1 = do_move_at(m, props, ty, rx, by, &ox, &oy);
            if (d1 != DO_BLOCKED) {
                if (d1 == DO_DIE) target = DIE_INTENT;
                else if (ox != xi || oy != yi) target = pix(m, ox, oy);
                vx *= collision;
            } else {
                saw_wood_at(m, ox, oy, vx, vy);
                uint32_t d2 = do_move_at(m, props, ty, bx, ry, &ox, &oy);
                if (d2 != DO_BLOCKED) {
                    if (d2 == DO_DIE) target = DIE_INTENT;
                    else if (ox != xi || oy != yi) target = pix(m, ox, oy);
                    vy *= collision;
                } else {
                    saw_wood_at(m, ox, oy, vx, vy);
                    vx *= collision;
                    vy *= collision;
                }
            }
        }
    }
    *pvx = vx;
    *pvy = vy;
    return target;
}

/* air_deposit_kernel (mod.rs:297-346), rule R1 */
void oracle_air_deposit(oracle_state *s) {
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4, n = xc * (yres / 4);
    uint32_t *key = calloc(n, 4);
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= 256) continue;
        uint32_t flags = s->part_flags[i];
        if ((flags & 0x10u) || (flags & 0x02u)) continue;
        int xi = ora_f2i(s->part_x[i] + 0.5f), yi = ora_f2i(s->part_y[i] + 0.5f);
        if (xi < 4 || yi < 4 || xi >= (int)xres - 4 || yi >= (int)yres - 4) continue;
        uint32_t cell = ((uint32_t)yi / 4) * xc + (uint32_t)xi / 4;
        if (s->wall_bmap[cell] == 18u && (s->wall_emap[cell] & 0xFFu) < 8u) continue;
        key[cell] = i + 1;
    }
    for (uint32_t c = 0; c < n; ++c) {
        if (!key[c]) continue;
        uint32_t i = key[c] - 1, base = s->part_ty[i] * 12 + 4;
        float drag = ora_bits_f(s->move_lut[base + 1]), loss = ora_bits_f(s->move_lut[base + 2]);
        s->air_vx[c] = s->air_vx[c] * loss + drag * s->part_vx[i];
        s->air_vy[c] = s->air_vy[c] * loss + drag * s->part_vy[i];
    }
    free(key);
}

/* velocity_and_intent_kernel (velocity.rs:951-1157) for all slots. */
void oracle_velocity_intent(oracle_state *s, uint32_t seed) {
    mctx m = {s, (int)s->xres, (int)s->yres, s->xres / 4, s->edge_mode, calloc(s->capacity, 1)};
    const uint32_t n_pix = s->xres * s->yres;
    memset(s->cell_winner, 0, sizeof(uint32_t) * n_pix);
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        s->intent_target[i] = NO_INTENT;
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= 256) continue;
        uint32_t flags = s->part_flags[i];
        if ((flags & 0x10u) || (flags & 0x02u)) continue;
        float px = s->part_x[i], py = s->part_y[i];
        int xi = ora_f2i(px + 0.5f), yi = ora_f2i(py + 0.5f);
        if (xi < 0 || xi >= m.xr || yi < 0 || yi >= m.yr) {
            s->part_ty[i] = 0;
            continue;
        }
        uint32_t cell = ((uint32_t)yi / 4) * m.xc + (uint32_t)xi / 4;
        uint32_t meta = ty * 12, props = s->move_lut[meta], falldown = s->move_lut[meta + 1], rank = s->move_lut[meta + 2];
        float advection = ora_bits_f(s->move_lut[meta + 4]), loss = ora_bits_f(s->move_lut[meta + 7]);
        float collision = ora_bits_f(s->move_lut[meta + 8]), gravity = ora_bits_f(s->move_lut[meta + 9]);
        float diffusion = ora_bits_f(s->move_lut[meta + 10]);
        /* pre_move_checks (velocity.rs:235-260) */
        uint32_t wall = s->wall_bmap[cell], emap = s->wall_emap[cell];
        int in_margin = !(xi < 4 || yi < 4 || xi >= m.xr - 4 || yi >= m.yr - 4);
        int stickman = ty == 59 || ty == 137 || ty == 159;
        if (!in_margin || (!stickman && wall_kills_particle(wall, emap, props))) {
            s->part_ty[i] = 0;
            continue;
        }
        if (wall == 18u && (emap & 0xFFu) < 8u) continue;
        float gx, gy;
        base_gravity(s, gravity, (float)xi, (float)yi, &gx, &gy);
        float ax = s->air_vx[cell], ay = s->air_vy[cell];
        uint32_t h1 = oracle_pcg_hash(seed ^ (i * 2654435769u)), h2 = oracle_pcg_hash(h1);
        float loss_here = s->newton_grav ? 1.0f : loss;
        float rdx = (float)(h1 & 0xFFFFu) * (2.0f / 65535.0f) - 1.0f;
        float rdy = (float)(h2 & 0xFFFFu) * (2.0f / 65535.0f) - 1.0f;
        float vx = s->part_vx[i] * loss_here + advection * ax + gx + diffusion * rdx;
        float vy = s->part_vy[i] * loss_here + advection * ay + gy + diffusion * rdy;
        /* clamp_plan_velocity (velocity.rs:205-219) */
        float mv = fabsf(vx);
        if (fabsf(vy) > mv) mv = fabsf(vy);
        if (mv > MAX_VELOCITY) {
            vx = vx * (MAX_VELOCITY / mv);
            vy = vy * (MAX_VELOCITY / mv);
        }
        int sx = step_axis(vx, h1), sy = step_axis(vy, h2);
        uint32_t target;
        if (falldown == 0)
            target = gas_destination(&m, &vx, &vy, px, py, xi, yi, sx, sy, props, ty, i, collision, seed);
        else
            target = fall_down_destination(&m, &vx, &vy, px, py, xi, yi, sx, sy, props, ty, falldown, flags, collision, h2,
                                           gravity, s->part_tmp[i]);
        if (target == DIE_INTENT) {
            s->part_ty[i] = 0;
            continue;
        }
        if (target != NO_INTENT) {
            s->intent_target[i] = target;
            if ((target & 0x80000000u) == 0) {
                uint32_t p = target & INTENT_PIXEL_MASK;
                uint32_t key = ((rank & 0x1FFu) << 23) | ((i + 1) & 0x7FFFFFu);
                if (key > s->cell_winner[p]) s->cell_winner[p] = key;
            }
        }
        s->part_vx[i] = vx;
        s->part_vy[i] = vy;
    }
    for (uint32_t i = 0; i < s->capacity; ++i)
        if (m.saw[i]) s->part_ty[i] = 180u; /* PT_SAWD */
    free(m.saw);
}

/* commit.rs:76-146 */
static int void_admits(uint32_t ty, uint32_t vctype, int32_t vtmp) {
    if (vctype == 0) return 1;
    int matches = vctype == ty, inv = (vtmp & 1) != 0;
    return inv ? !matches : matches;
}

static uint32_t resolve_can(oracle_state *s, uint32_t can, uint32_t ty, uint32_t dty, uint32_t other, float pv) {
    if (can != 3) return can;
    if (dty == 58 && ty == 29) return s->part_life[other] > 5 ? 2 : 0;
    if (dty == 43 && ty == 29) return s->part_life[other] < 10 ? 2 : 0;
    if (dty == 124) {
        float res = 4.0f;
        if (s->part_tmp[other] > 0) res = (float)s->part_tmp[other];
        return (pv < -res || pv > res) ? 2 : 0;
    }
    if (dty == 75) return (s->part_life[other] == 10 && void_admits(ty, s->part_ctype[other], s->part_tmp[other])) ? 1 : 0;
    if (dty == 23) return void_admits(ty, s->part_ctype[other], s->part_tmp[other]) ? 1 : 0;
    if (dty == 60) return ty == 151 ? (s->part_life[other] >= 10 ? 2 : 0) : 1;
    return 1;
}

static int is_heat_insulator_c(uint32_t ty, int32_t life, int32_t tmp, uint32_t hc) {
    return hc == 0 || (ty == 68 && life != 10) || ((ty == 82 || ty == 162) && (tmp & 1) == 0);
}

static uint32_t pmap_at(const oracle_state *s, int x, int y) {
    if (x < 0 || y < 0 || x >= (int)s->xres || y >= (int)s->yres) return 0;
    return s->pmap[(uint32_t)y * s->xres + (uint32_t)x];
}

/* commit_move_kernel (commit.rs:181-535) for all slots. */
/* A blocked photon's hit on the pixel it could not enter (commit.rs:263-320):
 * wavelength filter, PSCN/NSCN sparking, temperature exchange. The reference
 * runs these read-modify-writes of the hit particle from concurrent threads;
 * the canonical order (R10) applies every hit after the other commits of the
 * pass, in ascending photon slot order. */
static void photon_hit(oracle_state *s, uint32_t i, uint32_t target) {
    const uint32_t xres = s->xres;
    uint32_t hit = target & 0x7FFFFFFFu, r = s->pmap[hit];
    if (r >= 512u) {
        uint32_t rty = r & 0x1FFu, rslot = (r >> 9) - 1;
        uint32_t mask = s->photon_mask[rty];
        if (rty == 190) {
            int bin = (int)(s->part_ctype[rslot] / 4u);
            if (bin < 0) bin = 0;
            if (bin > 25) bin = 25;
            mask = 0x1Fu << (uint32_t)bin;
        } else if (rty == 194) {
            uint32_t colour = (s->part_ctype[rslot] >> 6) & 0x3Fu;
            if ((colour & 48u) == 0) mask |= 1u << 25;
            if ((colour & 12u) == 0) mask |= 1u << 15;
            if ((colour & 3u) == 0) mask |= 1u << 5;
        }
        s->part_ctype[i] &= mask;
        if (rty == 50 && s->part_life[rslot] == 0) {
            int hx = (int)(hit % xres), hy = (int)(hit / xres);
            if ((pmap_at(s, hx - 1, hy) & 0x1FFu) == 51 || (pmap_at(s, hx + 1, hy) & 0x1FFu) == 51 ||
                (pmap_at(s, hx, hy - 1) & 0x1FFu) == 51 || (pmap_at(s, hx, hy + 1) & 0x1FFu) == 51) {
                s->part_ctype[rslot] = 50;
                s->part_ty[rslot] = 12;
                s->part_life[rslot] = 4;
            }
        }
        if (s->heat_enable) {
            if (rty == 62 || rty == 27) s->part_temp[rslot] = s->part_temp[i];
            uint32_t hc = s->move_lut[rty * 12 + 3];
            if (!is_heat_insulator_c(rty, s->part_life[rslot], s->part_tmp[rslot], hc) && rty != 134) {
                float avg = (s->part_temp[rslot] + s->part_temp[i]) * 0.5f;
                if (avg > 9999.0f) avg = 9999.0f;
                if (avg < 0.0f) avg = 0.0f;
                s->part_temp[i] = avg;
                s->part_temp[rslot] = avg;
            }
        }
    }
}

void oracle_commit(oracle_state *s) {
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4;
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= 256) continue;
        uint32_t raw = s->intent_target[i];
        if (raw == NO_INTENT) {
            if (s->move_lut[ty * 12 + 1] >= 2 && (s->part_flags[i] & 0x10u) == 0) s->part_flags[i] |= 0x01u;
            continue;
        }
        int half_mirror = (raw & INTENT_HALF_MIRROR) != 0;
        uint32_t target = raw & INTENT_PIXEL_MASK;
        if (raw & 0x80000000u) continue; /* photon hits: after the loop (R10) */
        if (target >= xres * yres) continue;
        s->part_flags[i] &= 0xFFFFFFFEu;
        uint32_t key = s->cell_winner[target];
        if ((key & 0x7FFFFFu) != i + 1) continue;
        if (half_mirror) {
            s->part_x[i] = (float)(target % xres);
            s->part_y[i] = (float)(target / xres);
            continue;
        }
        int nx = (int)(target % xres), ny = (int)(target / xres);
        uint32_t wcell = ((uint32_t)ny / 4) * xc + (uint32_t)nx / 4, wall = s->wall_bmap[wcell];
        uint32_t dest = s->pmap[target], dty = dest & 0x1FFu, docc = dest >> 9;
        uint32_t mprops = s->move_lut[ty * 12], dprops = s->move_lut[dty * 12];
        uint32_t can = 0, other = i;
        if (docc != 0) {
            other = docc - 1;
            can = resolve_can(s, can_move_at(s, ty, dty), ty, dty, other, s->air_pv[wcell]);
        }
        int ehole_exit = 0;
        if (wall == 12 && (s->wall_emap[wcell] & 0xFFu) == 0) {
            int cx = ora_f2i(s->part_x[i] + 0.5f) / 4, cy = ora_f2i(s->part_y[i] + 0.5f) / 4;
            uint32_t here = (uint32_t)cy * xc + (uint32_t)cx;
            if (s->wall_bmap[here] == 12 && (s->wall_emap[here] & 0xFFu) == 0) ehole_exit = 1;
        }
        if (ehole_exit) {
        } else if (docc != 0 && can == 0) {
        } else if (docc != 0 && can == 1 && other != i) {
            if (dty == 23 || dty == 75) {
                s->part_ty[i] = 0;
            } else if (dty == 91 || dty == 40) {
                float t = s->part_temp[other] + s->part_temp[i] / 2.0f;
                s->part_temp[other] = ora_clampf(t, 0.0f, 9999.0f);
                s->part_ty[i] = 0;
            } else if (dty == 92 || dty == 41) {
                if (ty == 122) {
                    float t = s->part_temp[other] - (9999.0f - s->part_temp[i]) / 2.0f;
                    s->part_temp[other] = ora_clampf(t, 0.0f, 9999.0f);
                    s->part_ty[i] = 0;
                }
            } else if (dty == 111) {
                if (ty == 31) {
                    if (s->part_life[other] < 6000) s->part_life[other] += 1;
                    s->part_temp[other] = 0.0f;
                    s->part_ty[i] = 0;
                }
            } else if (dty == 165 || dty == 166) {
                if (mprops & 0x10u) {
                    s->part_tmp[other] += 20;
                    s->part_ty[i] = 0;
                }
            } else if (ty == 30 && (dprops & 0x100u)) {
                s->part_ty[i] = 0;
            } else if (ty == 30) {
                int own_x = ora_f2i(s->part_x[i] + 0.5f), own_y = ora_f2i(s->part_y[i] + 0.5f);
                uint32_t own_cell = ((uint32_t)own_y / 4) * xc + (uint32_t)own_x / 4;
                uint32_t own = s->pmap[(uint32_t)own_y * xres + (uint32_t)own_x];
                int drag = own != 0 && (s->move_lut[(own & 0x1FFu) * 12] & 0x80u) != 0;
                if (drag && s->wall_bmap[own_cell] != 15u) {
                    uint32_t under = (own >> 9) - 1;
                    s->part_x[under] = (float)nx;
                    s->part_y[under] = (float)ny;
                    s->part_x[other] = (float)own_x;
                    s->part_y[other] = (float)own_y;
                }
                s->part_x[i] = (float)nx;
                s->part_y[i] = (float)ny;
            } else if (ty == 42 && s->part_life[i] > 0) {
            } else {
                uint32_t oraw = s->intent_target[other], otgt = oraw & INTENT_PIXEL_MASK;
                int other_moves = 0;
                if (oraw != NO_INTENT && (oraw & 0x80000000u) == 0 && otgt < xres * yres &&
                    (s->cell_winner[otgt] & 0x7FFFFFu) == other + 1)
                    other_moves = 1;
                if (!other_moves) {
                    float ox = s->part_x[other], oy = s->part_y[other];
                    s->part_x[other] = s->part_x[i];
                    s->part_y[other] = s->part_y[i];
                    s->part_x[i] = ox;
                    s->part_y[i] = oy;
                }
            }
        } else {
            s->part_x[i] = (float)nx;
            s->part_y[i] = (float)ny;
        }
    }
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t raw = s->intent_target[i];
        if (s->part_ty[i] == 29 && raw != NO_INTENT && (raw & 0x80000000u))
            photon_hit(s, i, raw & INTENT_PIXEL_MASK);
    }
}

/* clear_spatial_maps + scatter_spatial_maps (commit.rs:541-603), rule R7 */
void oracle_rebuild_maps(oracle_state *s) {
    const uint32_t xres = s->xres, yres = s->yres;
    memset(s->pmap, 0, sizeof(uint32_t) * xres * yres);
    memset(s->photons, 0, sizeof(uint32_t) * xres * yres);
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= 256) continue;
        int xi = ora_f2i(s->part_x[i] + 0.5f), yi = ora_f2i(s->part_y[i] + 0.5f);
        if (xi < 0 || xi >= (int)xres || yi < 0 || yi >= (int)yres) continue;
        uint32_t idx = (uint32_t)yi * xres + (uint32_t)xi, packed = ((i + 1) << 9) | (ty & 0x1FFu);
        if (s->elem_props[ty * 32 + 1] & 0x10u) s->photons[idx] = packed;
        else if (s->pmap[idx] == 0 || (ty != 124 && ty != 134)) s->pmap[idx] = packed;
    }
}

/* WL_DETECT charging and the set_emap closure (mod.rs:348-828,
 * Simulation.cpp:928-983, 2330-2331). Emap values live in the low 6 bits;
 * bits 7/6 mark the left/right ends of a span filled in this pass. */
#define EMAP_CHARGED 16u
#define EMAP_LEFT 0x80u
#define EMAP_RIGHT 0x40u

static int is_wire(uint32_t w) { return w == 3 || w == 2 || w == 6 || w == 1 || w == 5 || w == 12 || w == 18; }
static int wire_off(const oracle_state *s, uint32_t c) { return is_wire(s->wall_bmap[c]) && (s->wall_emap[c] & 0x3Fu) < 8u; }
static int wire_at(const oracle_state *s, int x, int y, int xc, int yc) {
    return x >= 0 && y >= 0 && x < xc && y < yc && is_wire(s->wall_bmap[y * xc + x]);
}

/* Charges the uncharged wire run through (x, y) and marks its ends. */
static void fill_run(oracle_state *s, uint32_t row, uint32_t x1, uint32_t x2) {
    for (uint32_t x = x1; x <= x2; ++x) s->wall_emap[row + x] = EMAP_CHARGED;
    s->wall_emap[row + x1] |= EMAP_LEFT;
    s->wall_emap[row + x2] |= EMAP_RIGHT;
}

/* ladder shortcuts of set_emap (mod.rs:412-456, Simulation.cpp:955-985):
 * dir -1 looks up (cap row below the target), +1 looks down. */
static int ladder(const oracle_state *s, int singleton, int x, int target, int dir, int xc, int yc) {
    int cap = target - dir;
    return singleton && wire_at(s, x - 1, cap, xc, yc) && wire_at(s, x, cap, xc, yc) && wire_at(s, x + 1, cap, xc, yc) &&
           !wire_at(s, x - 1, target, xc, yc) && wire_at(s, x, target, xc, yc) && !wire_at(s, x + 1, target, xc, yc);
}

/* detect_charge_kernel (mod.rs:474-545) in slot order. */
void oracle_detect_charge(oracle_state *s) {
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4;
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= 256 || (s->part_flags[i] & 0x10u)) continue;
        int xi = ora_f2i(s->part_x[i] + 0.5f), yi = ora_f2i(s->part_y[i] + 0.5f);
        if (xi < 4 || yi < 4 || xi >= (int)xres - 4 || yi >= (int)yres - 4) continue;
        uint32_t cx = (uint32_t)xi / 4, row = ((uint32_t)yi / 4) * xc, cell = row + cx;
        if (s->wall_bmap[cell] != 3u || (s->wall_emap[cell] & 0x3Fu) >= 8u) continue;
        uint32_t x1 = cx, x2 = cx;
        while (x1 > 0 && wire_off(s, row + x1 - 1)) --x1;
        while (x2 + 1 < xc && wire_off(s, row + x2 + 1)) ++x2;
        fill_run(s, row, x1, x2);
    }
}

/* emap_gate_kernel + emap_closure_kernel + strip (mod.rs:556-828). */
void oracle_emap_closure(oracle_state *s) {
    const int xc = (int)(s->xres / 4), yc = (int)(s->yres / 4);
    const uint32_t cells = (uint32_t)(xc * yc);
    int any_off = 0, any_charged = 0;
    for (uint32_t c = 0; c < cells; ++c) {
        any_off |= wire_off(s, c);
        any_charged |= (s->wall_emap[c] & 0x3Fu) == EMAP_CHARGED;
    }
    if (any_off && any_charged) {
        int changed = 1;
        for (int pass = 0; changed && pass < yc + 2; ++pass) {
            changed = 0;
            for (int y = 0; y < yc; ++y) {
                uint32_t row = (uint32_t)(y * xc);
                for (int x = 0; x < xc;) {
                    if (!wire_off(s, row + (uint32_t)x)) {
                        ++x;
                        continue;
                    }
                    int x1 = x, x2 = x;
                    while (x2 + 1 < xc && wire_off(s, row + (uint32_t)x2 + 1)) ++x2;
                    int act = 0;
                    /* child above: a charged span one row below */
                    for (int xx = x1; xx <= x2 && !act && y + 1 < yc; ++xx) {
                        uint32_t se = s->wall_emap[(uint32_t)((y + 1) * xc + xx)];
                        int sy = y + 1, single = (se & (EMAP_LEFT | EMAP_RIGHT)) == (EMAP_LEFT | EMAP_RIGHT);
                        if ((se & 0x3Fu) == EMAP_CHARGED && !ladder(s, single, xx, sy - 2, -1, xc, yc) &&
                            ((se & (EMAP_LEFT | EMAP_RIGHT)) || y + 1 == yc - 1 || wire_at(s, xx - 1, y, xc, yc) ||
                             wire_at(s, xx + 1, y, xc, yc) || wire_at(s, xx - 1, sy + 1, xc, yc) ||
                             !wire_at(s, xx, sy + 1, xc, yc) || wire_at(s, xx + 1, sy + 1, xc, yc)))
                            act = 1;
                    }
                    /* child below: a charged span one row above */
                    for (int xx = x1; xx <= x2 && !act && y > 0; ++xx) {
                        uint32_t se = s->wall_emap[(uint32_t)((y - 1) * xc + xx)];
                        int sy = y - 1, single = (se & (EMAP_LEFT | EMAP_RIGHT)) == (EMAP_LEFT | EMAP_RIGHT);
                        if ((se & 0x3Fu) == EMAP_CHARGED && !ladder(s, single, xx, y + 1, 1, xc, yc) &&
                            ((se & (EMAP_LEFT | EMAP_RIGHT)) || wire_at(s, xx - 1, y, xc, yc) ||
                             wire_at(s, xx + 1, y, xc, yc) || wire_at(s, xx - 1, sy - 1, xc, yc) ||
                             !wire_at(s, xx, sy - 1, xc, yc) || wire_at(s, xx + 1, sy - 1, xc, yc)))
                            act = 1;
                    }
                    /* ladder jumps from a singleton two rows away */
                    for (int xx = x1; xx <= x2 && !act; ++xx) {
                        int up = y + 2 < yc && (s->wall_emap[(uint32_t)((y + 2) * xc + xx)] & (EMAP_LEFT | EMAP_RIGHT)) ==
                                                   (EMAP_LEFT | EMAP_RIGHT);
                        int down = y >= 2 && (s->wall_emap[(uint32_t)((y - 2) * xc + xx)] & (EMAP_LEFT | EMAP_RIGHT)) ==
                                                 (EMAP_LEFT | EMAP_RIGHT);
                        if (ladder(s, up, xx, y, -1, xc, yc) || ladder(s, down, xx, y, 1, xc, yc)) act = 1;
                    }
                    if (act) {
                        fill_run(s, row, (uint32_t)x1, (uint32_t)x2);
                        changed = 1;
                    }
                    x = x2 + 1;
                }
            }
        }
    }
    for (uint32_t c = 0; c < cells; ++c) s->wall_emap[c] &= ~(EMAP_LEFT | EMAP_RIGHT);
}

/* ---- e == 2 pair pass (reaction_kernel/energy_entry.rs:90-393) ----------
 *
 * Canonical form, shared with the HIP kernels:
 *  R11 movers are resolved in ascending slot order against the types the
 *      intent pass left (the reference's transient PAIR_CLAIMED type of an
 *      occupant is not modelled: each occupant has at most one mover);
 *  R12 creation records are allocated in ascending requester slot order, the
 *      k-th record taking the k-th free slot of the device window
 *      [1024, 17408) in ascending order; at most 16384 records. */

#define PT_GLAS 24u
#define PT_PHOT 29u
#define PT_NEUT 30u
#define PT_ELEC 31u
#define PT_H2 32u
#define PT_GRVT 39u
#define PT_GPMP 43u
#define PT_GLOW 66u
#define PT_BIZR 85u
#define PT_BIZRG 86u
#define PT_BIZRS 87u
#define PT_BGLA 95u
#define PT_INVIS 124u
#define PT_FILT 134u
#define PT_C5 139u
#define PT_PROT 173u
#define SPAWN_START 1024u
#define SPAWN_END 17408u
#define CREATE_COUNT 16384u

/* hash_u32 (reaction_kernel/shared.rs:1015-1023). */
uint32_t oracle_hash_u32(uint32_t x) {
    uint32_t h = x;
    h ^= h >> 16;
    h *= 2246822507u;
    h ^= h >> 13;
    h *= 3266489909u;
    h ^= h >> 16;
    return h;
}

/* Element_FILT_getWavelengths (reaction_kernel/rays/emitter.rs:10-26). */
static uint32_t filt_wavelengths(uint32_t ctype, float temp) {
    if ((ctype & 0x3FFFFFFFu) != 0u) return ctype & 0x3FFFFFFFu;
    int bin = ora_f2i((temp - 273.0f) * 0.025f);
    if (bin < 0) bin = 0;
    if (bin > 25) bin = 25;
    return 0x1Fu << (uint32_t)bin;
}

/* Element_FILT_interactWavelengths (reaction_kernel/rays/emitter.rs:28-80);
 * shift counts keep their low five bits, as u32 shifts in WGSL do. */
uint32_t oracle_interact_wavelengths(int32_t filt_tmp, uint32_t filt_ctype, float filt_temp, uint32_t orig_wl) {
    const uint32_t mask = 0x3FFFFFFFu;
    uint32_t filt_wl = filt_wavelengths(filt_ctype, filt_temp);
    if (filt_tmp == 0) return filt_wl;
    if (filt_tmp == 1) return orig_wl & filt_wl;
    if (filt_tmp == 2) return orig_wl | filt_wl;
    if (filt_tmp == 3) return orig_wl & ~filt_wl;
    if (filt_tmp == 4 || filt_tmp == 5) {
        int shift = ora_f2i((filt_temp - 273.0f) * 0.025f);
        if (shift <= 0) shift = 1;
        uint32_t sh = (uint32_t)shift & 31u;
        return (filt_tmp == 4 ? orig_wl << sh : orig_wl >> sh) & mask;
    }
    if (filt_tmp == 6) return orig_wl;
    if (filt_tmp == 7) return orig_wl ^ filt_wl;
    if (filt_tmp == 8) return ~orig_wl & mask;
    if (filt_tmp == 10) {
        uint32_t lsb = filt_wl & (~filt_wl + 1u);
        return (orig_wl * lsb) & mask;
    }
    if (filt_tmp == 11) {
        uint32_t lsb = filt_wl & (~filt_wl + 1u);
        return lsb != 0u ? (orig_wl / lsb) & mask : orig_wl;
    }
    return filt_wl;
}

typedef struct {
    uint32_t slot, x, y, ctype;
    float temp, vx, vy;
} create_rec;

typedef struct {
    create_rec *rec;
    uint32_t n, cap;
    uint32_t *free_slots; /* ascending */
    uint32_t n_free;
} pair_queue;

static void queue_photon(pair_queue *q, uint32_t slot, int x, int y, uint32_t ctype, float temp, float vx, float vy) {
    if (q->n >= q->cap) return; /* requesters come in ascending slot order: the lowest ones are kept */
    create_rec r = {slot, (uint32_t)x, (uint32_t)y, ctype, temp, vx, vy};
    q->rec[q->n++] = r;
}

/* begin_pending_creates' free-slot collection (collect_gpu_free_slots_kernel,
 * reaction_kernel/shared.rs:1197-1222): empty, not portal-held window slots. */
static void collect_free(const oracle_state *s, pair_queue *q) {
    q->n_free = 0;
    for (uint32_t slot = SPAWN_START; slot < SPAWN_END && slot < s->capacity; ++slot)
        if (s->part_ty[slot] == 0u && (s->part_flags[slot] & 0x04u) == 0u) q->free_slots[q->n_free++] = slot;
}

/* pair_resolve_kernel (energy_entry.rs:90-393) for every slot, in order. */
static void pair_resolve(oracle_state *s, uint32_t frame_seed, pair_queue *q) {
    const int xres = (int)s->xres, yres = (int)s->yres;
    const uint32_t xc = s->xres / 4, pixels = s->xres * s->yres;
    const uint32_t *def_elec = s->spawn_defaults + PT_ELEC * 8u;
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0u || ty >= 256u) continue;
        if (!(ty == PT_PHOT || ty == PT_NEUT || ty == PT_ELEC || ty == PT_PROT || ty == PT_BIZR || ty == PT_BIZRG))
            continue;
        uint32_t raw = s->intent_target[i];
        if (raw == NO_INTENT || (raw & 0x80000000u) != 0u) continue;
        uint32_t dest = raw & INTENT_PIXEL_MASK;
        if (dest >= pixels) continue;
        float xf = s->part_x[i], yf = s->part_y[i];
        int xi = ora_f2i(xf + 0.5f), yi = ora_f2i(yf + 0.5f);
        uint32_t own = (uint32_t)(yi * xres + xi);
        if (own == dest) continue;
        uint32_t m = s->pmap[dest];
        if (m < 512u) continue;
        uint32_t rty = m & 0x1FFu, r = (m >> 9) - 1u;
        if (r >= s->slot_bound) continue;
        /* energy_entry.rs:171-175: a 27-bit mask, which also takes in the low
         * four bits of the winner's weight rank. */
        if ((s->cell_winner[dest] & 0x07FFFFFFu) != ((i + 1u) & 0x07FFFFFFu)) continue;
        /* The occupant claim (a CAS from the map's type) fails on a slot the
         * intent pass killed or a WOOD it sawed. */
        if (s->part_ty[r] != rty) continue;
        uint32_t seed = oracle_hash_u32(frame_seed ^ oracle_hash_u32(i));
        uint32_t final_ty = rty;
        if (ty == PT_PHOT) {
            if (rty == PT_GLOW) { /* Simulation.cpp:1263-1272 */
                if (s->part_life[r] == 0 && seed % 30u == 0u) {
                    s->part_life[r] = 120;
                    int lr = -1;
                    if (oracle_hash_u32(seed ^ 0x68E31DA9u) % 2u == 0u) lr = 1;
                    float fx = xf - (float)lr * 0.3f * s->part_vy[i];
                    float fy = yf + (float)lr * 0.3f * s->part_vx[i];
                    int gx = ora_f2i(fx + 0.5f), gy = ora_f2i(fy + 0.5f);
                    if (gx >= 0 && gx < xres && gy >= 0 && gy < yres) {
                        uint32_t g = s->pmap[gy * xres + gx];
                        uint32_t gslot = (g >> 9) - 1u;
                        if (g >= 512u && (g & 0x1FFu) == PT_GLOW && gslot < s->slot_bound) {
                            int bin = ora_f2i((s->part_temp[gslot] - 273.0f) * 0.25f);
                            if (bin < 0) bin = 0;
                            if (bin > 25) bin = 25;
                            queue_photon(q, i, gx, gy, 0x1Fu << (uint32_t)bin, s->part_temp[gslot], s->part_vx[i],
                                         s->part_vy[i]);
                        }
                    }
                }
            } else if (rty == PT_FILT) { /* :1273-1275 */
                s->part_ctype[i] = oracle_interact_wavelengths(s->part_tmp[r], s->part_ctype[r], s->part_temp[r],
                                                               s->part_ctype[i]);
            } else if (rty == PT_C5) {
                uint32_t rc = s->part_ctype[r], ic = s->part_ctype[i];
                if (s->part_life[r] > 0 && (rc & ic & 0xFFFFFFC0u) != 0u) { /* :1277-1303 */
                    int32_t rt = s->part_tmp[r];
                    float cvx = (float)((int32_t)((uint32_t)rt << 16) >> 16) / 255.0f;
                    float cvy = (float)(rt >> 16) / 255.0f;
                    float ovx = s->part_vx[i], ovy = s->part_vy[i];
                    if (ovx + cvx == 0.0f && ovy + cvy == 0.0f) {
                        s->part_vx[i] = cvy;
                        s->part_vy[i] = -cvx;
                    } else {
                        s->part_ctype[i] = (rc & ic) >> 6;
                        s->part_vx[i] = ovx + cvx;
                        s->part_vy[i] = ovy + cvy;
                        float speed = s->part_vx[i] * s->part_vx[i] + s->part_vy[i] * s->part_vy[i];
                        float scale = sqrtf((ovx * ovx + ovy * ovy) / speed);
                        s->part_vx[i] *= scale;
                        s->part_vy[i] *= scale;
                    }
                    s->part_life[r] = 0;
                    s->part_ctype[r] = 0u;
                } else if (rc == 0u && (ic & 0xFFFFFFC0u) != 0u) { /* :1304-1314 */
                    s->part_life[r] = 1;
                    s->part_ctype[r] = ic;
                    s->part_tmp[r] = (int32_t)((0xFFFFu & (uint32_t)ora_f2i(s->part_vx[i] * 255.0f)) |
                                               (0xFFFF0000u & (uint32_t)ora_f2i(s->part_vy[i] * 16711680.0f)));
                    s->part_tmp2[r] = (int32_t)((0xFFFFu & (uint32_t)ora_f2i((xf - (float)xi) * 255.0f)) |
                                                (0xFFFF0000u & (uint32_t)ora_f2i((yf - (float)yi) * 16711680.0f)));
                    s->part_ty[i] = 0u;
                }
            } else if (rty == PT_INVIS) { /* :1315-1326 */
                float resistance = 4.0f;
                if (s->part_tmp[r] > 0) resistance = (float)s->part_tmp[r];
                uint32_t px = dest % s->xres, py = dest / s->xres;
                float pv = s->air_pv[(py / 4u) * xc + px / 4u];
                if (pv >= -resistance && pv <= resistance) {
                    s->part_ty[i] = PT_NEUT;
                    s->part_ctype[i] = 0u;
                }
            } else if (rty == PT_BIZR || rty == PT_BIZRG || rty == PT_BIZRS) { /* :1326-1332 */
                s->part_ty[i] = PT_ELEC;
                s->part_ctype[i] = 0u;
            } else if (rty == PT_H2) { /* :1333-1343 */
                if ((s->part_tmp[i] & 1) == 0) {
                    s->part_ty[i] = PT_PROT;
                    s->part_ctype[i] = 0u;
                    s->part_tmp2[i] = 1;
                    final_ty = PT_ELEC;
                    s->part_x[r] = (float)xi;
                    s->part_y[r] = (float)yi;
                    s->intent_target[r] = NO_INTENT;
                    /* create_part(ID(r), x, y, PT_ELEC): DefaultProperties, then
                     * ELEC's Create hook (ELEC.cpp:159-165). */
                    s->part_temp[r] = ora_bits_f(def_elec[0]);
                    s->part_life[r] = 680;
                    s->part_ctype[r] = def_elec[2];
                    s->part_tmp[r] = (int32_t)def_elec[3];
                    s->part_tmp2[r] = (int32_t)def_elec[4];
                    uint32_t roll = oracle_hash_u32(frame_seed ^ ((r + 1u) << 7) ^ (PT_ELEC << 3));
                    float a = (float)(roll % 360u) / 360.0f * 6.28318530717958647692f;
                    s->part_vx[r] = 2.0f * (float)cos((double)a);
                    s->part_vy[r] = 2.0f * (float)sin((double)a);
                }
            } else if (rty == PT_GPMP && s->part_life[r] == 0) { /* :1344-1351 */
                s->part_ty[i] = PT_GRVT;
                s->part_tmp[i] = ora_f2i(s->part_temp[r] - 273.15f);
            }
        } else if (ty == PT_NEUT) { /* :1353-1358 */
            if ((rty == PT_GLAS || rty == PT_BGLA) && seed % 10u == 0u) {
                uint32_t g = s->pmap[own];
                uint32_t gty = g & 0x1FFu, gslot = (g >> 9) - 1u;
                if (g >= 512u && (gty == PT_GLAS || gty == PT_BGLA) && gslot < s->slot_bound) {
                    float vx = s->part_vx[i], vy = s->part_vy[i];
                    if (vx * vx + vy * vy >= 1.44f * 1.44f) {
                        /* cherenkov_photon (reaction_kernel/neut/mod.rs:335-369) */
                        uint32_t roll = oracle_hash_u32(seed ^ 0x9E3779B9u);
                        int lr = (int)(roll % 2u) * 2 - 1;
                        float rvx = vx - (float)lr * 2.5f * vy;
                        float rvy = vy + (float)lr * 2.5f * vx;
                        float speed = sqrtf(rvx * rvx + rvy * rvy);
                        float scale = 1.269f / speed;
                        queue_photon(q, i, xi, yi, 0x00000F80u, s->part_temp[gslot], rvx * scale, rvy * scale);
                    }
                }
            }
        } else if (ty == PT_ELEC) { /* :1362-1368 */
            if (rty == PT_GLOW) {
                s->part_ty[i] = PT_PHOT;
                s->part_ctype[i] = 0x3FFFFFFFu;
            }
        } else if (ty == PT_PROT) { /* :1370-1374 */
            if (rty == PT_INVIS) s->part_ty[i] = PT_NEUT;
        } else if (rty == PT_FILT) { /* BIZR / BIZRG + FILT, :1375-1379 */
            s->part_ctype[i] = oracle_interact_wavelengths(s->part_tmp[r], s->part_ctype[r], s->part_temp[r],
                                                           s->part_ctype[i]);
        }
        s->part_ty[r] = final_ty;
    }
}

/* drain_pending_creates for the pair pass's photons (create_alloc_kernel +
 * create_fields_kernel, reaction_kernel/shared.rs:516-912) with the record
 * mask both creators use (CR_NO_OCCUPANCY | CR_SET_CTYPE | CR_SET_TEMP |
 * CR_SET_VX | CR_SET_VY): no occupancy test; life and tmp..tmp4 from PHOT's
 * DefaultProperties, dcolour and flags cleared.  The map words the
 * reference writes are rebuilt by the scatter that follows. */
static void drain_creates(oracle_state *s, const pair_queue *q) {
    const uint32_t *def = s->spawn_defaults + PT_PHOT * 8u;
    uint32_t cursor = 0;
    for (uint32_t k = 0; k < q->n; ++k) {
        const create_rec *r = &q->rec[k];
        if (r->x >= s->xres || r->y >= s->yres) continue;
        uint32_t a = cursor++;
        if (a >= q->n_free) continue;
        uint32_t slot = q->free_slots[a];
        s->part_ty[slot] = PT_PHOT;
        s->part_x[slot] = (float)r->x;
        s->part_y[slot] = (float)r->y;
        s->part_flags[slot] = 0u;
        s->part_life[slot] = (int32_t)def[1];
        s->part_ctype[slot] = r->ctype;
        s->part_temp[slot] = r->temp;
        s->part_tmp[slot] = (int32_t)def[3];
        s->part_tmp2[slot] = (int32_t)def[4];
        s->part_tmp3[slot] = (int32_t)def[5];
        s->part_tmp4[slot] = (int32_t)def[6];
        s->part_dcolour[slot] = 0u;
        s->part_vx[slot] = r->vx;
        s->part_vy[slot] = r->vy;
    }
}

/* step_movement_gpu (mod.rs:1176-1227). */
void oracle_step_movement(oracle_state *s, uint32_t seed) {
    /* gpu_spawn_used (mod.rs:1212-1216): the bound covers the window. */
    uint32_t spawn_end = s->capacity < SPAWN_END ? s->capacity : SPAWN_END;
    if (s->slot_bound < spawn_end) s->slot_bound = spawn_end;
    oracle_air_deposit(s);
    oracle_detect_charge(s);
    oracle_emap_closure(s);
    oracle_velocity_intent(s, seed);
    pair_queue q = {calloc(CREATE_COUNT, sizeof(create_rec)), 0, CREATE_COUNT,
                    calloc(SPAWN_END - SPAWN_START, sizeof(uint32_t)), 0};
    collect_free(s, &q);
    pair_resolve(s, seed, &q);
    oracle_commit(s);
    drain_creates(s, &q);
    free(q.rec);
    free(q.free_slots);
    oracle_rebuild_maps(s);
}

/* ---- water equalisation (water_equal_gpu/, lib.rs:418-420) ----------------
 *
 * Canonical rule (shared with the HIP kernels):
 *  WEQ the per-column target that several rows write (row.rs:135-139) keeps
 *      the write of the highest-numbered row, as processing the rows in
 *      increasing y does. */

static int is_liquid_word(const oracle_state *s, uint32_t p) {
    if (p == 0u) return 0;
    uint32_t ty = p & 0x1FFu;
    return ty < 256u && s->elem_props[ty * 32u + 3u] == 2u;
}

/* hash_u32-based 1-in-200 gate (intent.rs:46-49). */
static int flood_chance(uint32_t frame_seed, uint32_t id) {
    return oracle_hash_u32(frame_seed ^ (id * 0x9E3779B9u)) % 200u == 0u;
}

void oracle_water_equal(oracle_state *s, uint32_t seed) {
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4, P = xres * yres;
    uint32_t *col_surf = malloc(sizeof(uint32_t) * xres), *col_target = calloc(xres, sizeof(uint32_t));
    uint32_t *winner = calloc(P, sizeof(uint32_t));
    mctx m = {s, (int)xres, (int)yres, xc, s->edge_mode, NULL};
    /* Pass 1: column surfaces (column.rs:20-82). */
    for (uint32_t x = 0; x < xres; ++x) {
        int found = 0;
        uint32_t min_y = yres, max_y = 0;
        for (uint32_t y = 0; y < yres; ++y)
            if (is_liquid_word(s, s->pmap[y * xres + x])) {
                if (!found) {
                    found = 1;
                    min_y = y;
                }
                max_y = y;
            }
        if (found) {
            int free_above = 0;
            if (min_y > 0 && s->pmap[(min_y - 1) * xres + x] == 0u) {
                uint32_t wc = ((min_y - 1) / 4) * xc + x / 4;
                if (!wall_blocks_particle(s->wall_bmap[wc], s->wall_emap[wc], 0x2u)) free_above = 1;
            }
            col_surf[x] = (min_y << 16) | (free_above ? 1u << 15 : 0u) | (max_y & 0x7FFFu);
        } else {
            col_surf[x] = 0xFFFFFFFFu;
        }
    }
    /* Pass 2: spans of liquid per row, in increasing y (row.rs:22-150). */
    for (uint32_t y = 0; y < yres; ++y) {
        uint32_t x = 0;
        while (x < xres) {
            if (!is_liquid_word(s, s->pmap[y * xres + x])) {
                ++x;
                continue;
            }
            uint32_t x_start = x, x_end = x;
            for (uint32_t nx = x + 1; nx < xres; ++nx) {
                uint32_t wc = (y / 4) * xc + nx / 4;
                if (is_liquid_word(s, s->pmap[y * xres + nx]) &&
                    !wall_blocks_particle(s->wall_bmap[wc], s->wall_emap[wc], 0x2u))
                    x_end = nx;
                else
                    break;
            }
            uint32_t min_top = yres, max_free = 0, dmin = xres, dmax = 0;
            for (uint32_t k = x_start; k <= x_end; ++k) {
                uint32_t sf = col_surf[k];
                if (sf == 0xFFFFFFFFu) continue;
                uint32_t top = sf >> 16, bot = sf & 0x7FFFu;
                int fr = (sf & (1u << 15)) != 0;
                if (y >= top && y <= bot) {
                    if (top < min_top) min_top = top;
                    if (fr) {
                        if (top > max_free) {
                            max_free = top;
                            dmin = dmax = k;
                        } else if (top == max_free) {
                            if (k < dmin) dmin = k;
                            if (k > dmax) dmax = k;
                        }
                    }
                }
            }
            if (max_free > min_top + 1u) {
                uint32_t dest_row = max_free - 1u, span_len = dmax - dmin + 1u;
                for (uint32_t k = x_start; k <= x_end; ++k) {
                    uint32_t sf = col_surf[k];
                    if (sf == 0xFFFFFFFFu) continue;
                    if ((sf >> 16) < dest_row) {
                        uint32_t dest_x = dmin + (k + seed) % span_len, dest = dest_row * xres + dest_x;
                        col_target[k] = dest + 1u;
                        winner[dest] = 0xFFFFFFFFu;
                    }
                }
            }
            x = x_end + 1u;
        }
    }
    /* Pass 3: intents, the lowest slot taking a target (intent.rs:50-129). */
    for (uint32_t id = 0; id < s->slot_bound; ++id) {
        uint32_t ty = s->part_ty[id];
        if (ty == 0u || ty >= 256u || s->elem_props[ty * 32u + 3u] != 2u) continue;
        if (s->part_flags[id] & (0x10u | 0x02u)) continue;
        if (!flood_chance(seed, id)) continue;
        int xi = ora_f2i(s->part_x[id] + 0.5f), yi = ora_f2i(s->part_y[id] + 0.5f);
        if (xi < 0 || xi >= (int)xres || yi < 0 || yi >= (int)yres) continue;
        if ((s->pmap[(uint32_t)yi * xres + (uint32_t)xi] >> 9) != id + 1u) continue;
        uint32_t packed = col_target[xi];
        if (packed == 0u) continue;
        uint32_t target = packed - 1u, dest_y = target / xres, dest_x = target % xres;
        int can_enter = eval_move_at(&m, s->elem_props[ty * 32u + 1u], ty, (int)dest_x, (int)dest_y) != 0u;
        if ((uint32_t)yi < dest_y && s->pmap[target] == 0u && can_enter && id + 1u < winner[target])
            winner[target] = id + 1u;
    }
    /* Pass 4: commits (commit.rs:15-63). */
    for (uint32_t id = 0; id < s->slot_bound; ++id) {
        uint32_t ty = s->part_ty[id];
        if (ty == 0u || ty >= 256u) continue;
        int xi = ora_f2i(s->part_x[id] + 0.5f), yi = ora_f2i(s->part_y[id] + 0.5f);
        if (xi < 0 || xi >= (int)xres || yi < 0 || yi >= (int)yres) continue;
        uint32_t old_idx = (uint32_t)yi * xres + (uint32_t)xi, packed = col_target[xi];
        if (packed == 0u) continue;
        uint32_t target = packed - 1u, dest_y = target / xres;
        if ((uint32_t)yi < dest_y && winner[target] == id + 1u) {
            if ((s->pmap[old_idx] >> 9) == id + 1u) s->pmap[old_idx] = 0u;
            s->pmap[target] = ((id + 1u) << 9) | (ty & 0x1FFu);
            s->part_x[id] = (float)(target % xres);
            s->part_y[id] = (float)dest_y;
        }
    }
    free(col_surf);
    free(col_target);
    free(winner);
}

/* SimulationEngine::step (lib.rs:303-459) over the passes this build covers. */
void oracle_step(oracle_state *s, uint64_t frame_count) {
    uint32_t v = (uint32_t)frame_count * 0x9E3779B9u;
    uint32_t seed = ((v << 13) | (v >> 19)) ^ 0x85EBCA6Bu;
    oracle_step_air(s);
    oracle_step_gravity(s);
    oracle_recalc_free_particles(s, seed, (frame_count % 10u) == 0);
    oracle_step_thermal(s, seed, s->heat_enable ? 0 : 1);
    oracle_step_reactions(s, seed);  /* the element pass (lib.rs:379), as far as it is restated */
    oracle_step_movement(s, seed);
    if (s->water_equal_test) oracle_water_equal(s, seed);
}
