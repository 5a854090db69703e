// Lock-free movement for gfx950: move_kernel::step_movement_gpu
// (crates/tpt-sim/src/move_kernel/mod.rs:1176-1227).
//
//   mv_air_deposit(+_cells)  particle -> air drag (mod.rs:297-346); the racy
//                            per-cell store becomes "highest eligible slot
//                            writes" via atomicMax on slot+1 and a cell pass.
//   mv_air_deposit_list/_px  the same rule over the pixel records of the
//                            RecalcFreeParticles that ran before: one thread
//                            per cell takes its pixels' owners from the
//                            highest slot down; the off-map list first.
//   mv_prepare               cell_winner = 0 (mod.rs:268-273), the planner's
//                            byte type map and per-cell record, in one launch.
//   mv_intent                velocity integration, PlanMove and the per-
//                            Falldown destination chain, then the claim
//                            atomicMax(cell_winner[pixel],
//                            (weight_rank << 23) | (slot + 1))
//                            (velocity.rs:951-1157, liquid.rs:137-963).
//   mv_saw_pair              WOOD -> SAWD of rejected probes, after the pass,
//                            and the e == 2 pair pass (reaction_kernel/
//                            energy_entry.rs) over the movers mv_intent listed.
//   mv_commit                winner check, consume/swap/move, FLAG_STAGNANT
//                            (commit.rs:181-535).
//   map rebuild              pmap/photons rebuild (commit.rs:541-603): the
//                            highest matter slot owns a pixel, INVIS/FILT only
//                            take an empty one; built by the map builder of
//                            RecalcFreeParticles (spatial.hip) without life
//                            decay, so that the next recalc can update it.
//
//   mv_detect_* / mv_emap_*  WL_DETECT charging and the set_emap closure
//                            (mod.rs:348-828), on canvases with wire walls.
//
//   mv_after_commit          the deferred blocked-photon hits, then the drain
//                            of the photons the pair pass creates into the
//                            device window (reaction_kernel/shared.rs).

#include "move_common.h"

namespace {

using namespace tpt_move;

constexpr int BLK = 256;
constexpr uint32_t NO_INTENT = 0xFFFFFFFFu;
constexpr uint32_t DIE_INTENT = 0xFFFFFFFEu;
constexpr uint32_t HALF_MIRROR = 0x40000000u;
constexpr uint32_t PIXEL_MASK = 0x3FFFFFFFu;
constexpr uint32_t DO_BLOCKED = 0u, DO_OK = 1u, DO_DIE = 2u;
constexpr float ISTP = 2.0f;
constexpr float MAX_VELOCITY = 1.0e4f;
constexpr uint32_t MAX_PLAN_STEPS = 5001u;
constexpr uint32_t EDGE_VOID = 0u, EDGE_LOOP = 2u;

// Element ids the e == 2 pair pass names (tpt-core/src/constants.rs:63-424).
constexpr uint32_t PT_GLAS = 24u, PT_PHOT = 29u, PT_NEUT = 30u, PT_ELEC = 31u, PT_H2 = 32u, PT_GRVT = 39u,
                   PT_GPMP = 43u, PT_GLOW = 66u, PT_BIZR = 85u, PT_BIZRG = 86u, PT_BIZRS = 87u, PT_BGLA = 95u,
                   PT_INVIS = 124u, PT_FILT = 134u, PT_C5 = 139u, PT_PROT = 173u;

__device__ __forceinline__ bool pair_tracked(uint32_t ty) {
    return ty == PT_PHOT || ty == PT_NEUT || ty == PT_ELEC || ty == PT_PROT || ty == PT_BIZR || ty == PT_BIZRG;
}

__device__ __forceinline__ uint32_t pcg_hash(uint32_t seed) {
    const uint32_t state = seed * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__device__ __forceinline__ int round_i32(float v) { return tpt_f2i(v + 0.5f); }

__device__ __forceinline__ bool wall_kills_particle(uint32_t wall, uint32_t emap, uint32_t props) {
    if (wall == 8u || wall == 1u || wall == 9u || wall == 7u) return true;
    if (wall == 6u) return (props & 0x2u) == 0u;
    if (wall == 10u) return (props & 0x1u) == 0u;
    if (wall == 13u) return (props & 0x8u) == 0u;
    if (wall == 15u) return (props & 0x10u) == 0u;
    if (wall == 2u) return (emap & 0xFFu) == 0u;
    return false;
}

// eval_move_at (mod.rs:838-891).
__device__ uint32_t eval_move_at(const MoveEnv &e, uint32_t props, uint32_t ty, int nx, int ny) {
    if (!(nx >= 0 && nx < e.xr && ny >= 0 && ny < e.yr)) return 0u;
    return eval_move_ty(e, props, ty, nx, ny, (uint32_t)e.tmap[ny * e.xr + nx]);
}

__device__ __forceinline__ float remainder_p(float x, float y) {
    float out = x - floorf(x / y) * y;
    if (out < 0.0f) out += y;
    else if (out >= y) out -= y;
    return out;
}

__device__ __forceinline__ int wrap_loop_pixel(int n, int res) {
    const int lo = 4, span = res - 8;
    if (span > 0 && (n < lo || n >= lo + span)) return lo + tpt_f2i(remainder_p((float)(n - lo), (float)span));
    return n;
}

// do_move_at (mod.rs:1009-1056).
__device__ uint32_t do_move_at(const MoveEnv &e, uint32_t props, uint32_t ty, int nx, int ny, int &ox, int &oy) {
    int mx = nx, my = ny;
    if (e.edge == EDGE_LOOP) {
        mx = wrap_loop_pixel(nx, e.xr);
        my = wrap_loop_pixel(ny, e.yr);
    }
    uint32_t status = DO_BLOCKED;
    if (mx < 0 || my < 0 || mx >= e.xr || my >= e.yr) {
        if (e.edge != EDGE_LOOP) status = DO_DIE;
    } else if (eval_move_at(e, props, ty, mx, my) != 0u) {
        const bool in_band = mx >= 4 && my >= 4 && mx < e.xr - 4 && my < e.yr - 4;
        status = DO_OK;
        if (!in_band && e.edge != EDGE_LOOP) {
            status = DO_BLOCKED;
            if (e.edge == EDGE_VOID) status = DO_DIE;
        }
    }
    ox = mx;
    oy = my;
    return status;
}

__device__ void saw_wood_at(const MoveEnv &e, int nx, int ny, float vx, float vy) {
    if (nx >= 0 && ny >= 0 && nx < e.xr && ny < e.yr) {
        if (e.tmap[ny * e.xr + nx] != 15u) return;
        const uint32_t v = e.pmap[ny * e.xr + nx];
        if (v >= 512u && (v & 0x1FFu) == 15u && vx * vx + vy * vy > 25.0f)
            e.saw_list[atomicAdd(e.saw_count, 1u)] = (v >> 9) - 1u;
    }
}

__device__ __forceinline__ uint32_t pix(const MoveEnv &e, int x, int y) { return (uint32_t)(y * e.xr + x); }

// plan_sample_blocked (liquid.rs:80-118).
__device__ bool plan_sample_blocked(const MoveEnv &e, uint32_t props, uint32_t ty, int nx, int ny, bool closed_start) {
    if (nx < 0 || ny < 0 || nx >= e.xr || ny >= e.yr) return true;
    const uint32_t d_ty = e.tmap[ny * e.xr + nx];
    const uint32_t raw_can = e.lut[LUT_CAN + ty * 256u + d_ty];
    const uint32_t ev = eval_move_at(e, props, ty, nx, ny);
    const int wc = wcell(e, nx, ny);
    const uint32_t wall = e.cellrec[wc].z;
    const bool closed_now = wall == 12u && (e.cellrec[wc].w & 0xFFu) == 0u;
    return ev == 0u || (raw_can == 3u && ev == 1u) || (ty == 29u && d_ty != 0u) || wall == 7u ||
           closed_now != closed_start;
}

struct Plan {
    int fin_x, fin_y, clear_x, clear_y;
    float clear_xf, clear_yf;
};

__device__ __forceinline__ float wrap_band(float f, int res) {
    const bool ok = f >= 4.0f - 0.5f && f < (float)res - 4.0f - 0.5f;
    if (!ok) f = remainder_p(f - 4.0f + 0.5f, (float)res - 8.0f) + 4.0f - 0.5f;
    return f;
}

// plan_move (liquid.rs:137-261).
__device__ Plan plan_move(const MoveEnv &e, float vx, float vy, float px, float py, int xi, int yi, uint32_t props,
                          uint32_t ty) {
    float mv = fabsf(vx);
    if (fabsf(vy) > mv) mv = fabsf(vy);
    Plan p;
    p.clear_xf = px;
    p.clear_yf = py;
    float fin_xf = px + vx, fin_yf = py + vy;
    p.fin_x = round_i32(fin_xf);
    p.fin_y = round_i32(fin_yf);
    p.clear_x = xi;
    p.clear_y = yi;
    if (mv >= ISTP) {
        const int sc = wcell(e, xi, yi);
        const bool start_closed = e.cellrec[sc].z == 12u && (e.cellrec[sc].w & 0xFFu) == 0u;
        const float step_xf = vx * ISTP / mv, step_yf = vy * ISTP / mv;
        float walk_xf = px, walk_yf = py, remaining = mv;
        for (uint32_t guard = 0; guard < MAX_PLAN_STEPS; ++guard) {
            remaining -= ISTP;
            walk_xf += step_xf;
            walk_yf += step_yf;
            if (e.edge == EDGE_LOOP) {
                walk_xf = wrap_band(walk_xf, e.xr);
                walk_yf = wrap_band(walk_yf, e.yr);
            }
            const int nfx = round_i32(walk_xf), nfy = round_i32(walk_yf);
            if (remaining <= 0.0f) {
                fin_xf = px + vx;
                fin_yf = py + vy;
                if (e.edge == EDGE_LOOP) {
                    fin_xf = wrap_band(fin_xf, e.xr);
                    fin_yf = wrap_band(fin_yf, e.yr);
                }
                p.fin_x = round_i32(fin_xf);
                p.fin_y = round_i32(fin_yf);
                p.clear_xf = fin_xf - step_xf;
                p.clear_yf = fin_yf - step_yf;
                p.clear_x = round_i32(p.clear_xf);
                p.clear_y = round_i32(p.clear_yf);
                break;
            }
            if (plan_sample_blocked(e, props, ty, nfx, nfy, start_closed)) {
                p.fin_x = nfx;
                p.fin_y = nfy;
                p.clear_xf = walk_xf - step_xf;
                p.clear_yf = walk_yf - step_yf;
                p.clear_x = round_i32(p.clear_xf);
                p.clear_y = round_i32(p.clear_yf);
                break;
            }
        }
    }
    return p;
}

// base_gravity (velocity.rs:133-170).
__device__ __forceinline__ void base_gravity(const MoveEnv &e, float gravity, float px, float py, float &gx,
                                             float &gy) {
    gx = 0.0f;
    gy = 0.0f;
    if (e.gravity_mode == 0u) {
        gy = gravity;
    } else if (e.gravity_mode == 2u) {
        const float dx = px - (float)(e.xr / 2), dy = py - (float)(e.yr / 2);
        if (dx != 0.0f || dy != 0.0f) {
            const float d = 0.01f - sqrtf(dx * dx + dy * dy);
            gx = gravity * (dx / d);
            gy = gravity * (dy / d);
        }
    } else if (e.gravity_mode == 3u) {
        gx = gravity * e.cgx;
        gy = gravity * e.cgy;
    }
}

__device__ __forceinline__ int step_axis(float v, uint32_t h) {
    const float a = v < 0.0f ? -v : v;
    const float r = (float)(h & 0xFFFFu) * (1.0f / 65535.0f);
    if (a >= 1.0f || r < a) {
        if (v > 0.0f) return 1;
        if (v < 0.0f) return -1;
    }
    return 0;
}

__device__ uint32_t diagonal_liquid_move(const MoveEnv &e, float vx, float vy, float cxf, float cyf, int rdir,
                                         uint32_t props, uint32_t ty) {
    float ddx = vx - vy * (float)rdir;
    float ddy = vy + vx * (float)rdir;
    float dmv = fabsf(ddx);
    if (fabsf(ddy) > dmv) dmv = fabsf(ddy);
    if (!(dmv > 0.0001f)) return NO_INTENT;
    ddx /= dmv;
    ddy /= dmv;
    int ox, oy;
    const uint32_t d1 = do_move_at(e, props, ty, round_i32(cxf + ddx), round_i32(cyf + ddy), ox, oy);
    if (d1 == DO_DIE) return DIE_INTENT;
    if (d1 == DO_OK) return pix(e, ox, oy);
    const uint32_t d2 =
        do_move_at(e, props, ty, round_i32(cxf + ddy * (float)rdir), round_i32(cyf - ddx * (float)rdir), ox, oy);
    if (d2 == DO_DIE) return DIE_INTENT;
    if (d2 == DO_OK) return pix(e, ox, oy);
    return NO_INTENT;
}

// scan_liquid_vertical (liquid.rs:630-733).
__device__ uint32_t scan_liquid_vertical(const MoveEnv &e, int xi, int yi, int fin_y, int clear_x, int clear_y,
                                         int rdir, int rt, uint32_t ty, uint32_t props, bool &fall) {
    int scan_x = xi, scan_y = yi, steps = 0;
    bool hit = false, die = false, stop = false;
    int j = clear_x + rdir;
    fall = false;
    while (!stop && steps < rt) {
        if (j < 0 || j >= e.xr || j < clear_x - rt || j >= clear_x + rt) {
            stop = true;
        } else {
            if (fin_y >= 0 && fin_y < e.yr) {
                const uint32_t d_ty = (uint32_t)e.tmap[pix(e, j, fin_y)];
                if (d_ty != ty || e.cellrec[wcell(e, j, fin_y)].z != 0u) {
                    int ox, oy;
                    const uint32_t d = do_move_at(e, props, ty, j, fin_y, ox, oy);
                    if (d == DO_DIE) {
                        die = true;
                        stop = true;
                    } else if (d == DO_OK) {
                        hit = true;
                        fall = true;
                        scan_x = ox;
                        scan_y = oy;
                        stop = true;
                    }
                }
            }
            if (!stop && fin_y != clear_y && clear_y >= 0 && clear_y < e.yr) {
                const uint32_t d_ty = (uint32_t)e.tmap[pix(e, j, clear_y)];
                if (d_ty != ty || e.cellrec[wcell(e, j, clear_y)].z != 0u) {
                    int ox, oy;
                    const uint32_t d = do_move_at(e, props, ty, j, clear_y, ox, oy);
                    if (d == DO_DIE) {
                        die = true;
                        stop = true;
                    } else if (d == DO_OK) {
                        hit = true;
                        fall = true;
                        scan_x = ox;
                        scan_y = oy;
                        stop = true;
                    }
                }
            }
            if (!stop && clear_y >= 0 && clear_y < e.yr) {
                const uint32_t d_ty = (uint32_t)e.tmap[pix(e, j, clear_y)];
                const uint32_t wall = e.cellrec[wcell(e, j, clear_y)].z;
                if (d_ty != ty || (wall != 0u && wall != 4u)) stop = true;
            }
            if (!stop) {
                j += rdir;
                steps++;
            }
        }
    }
    if (die) return DIE_INTENT;
    if (hit) return pix(e, scan_x, scan_y);
    return NO_INTENT;
}

// scroll_liquid_down (liquid.rs:735-787).
__device__ uint32_t scroll_liquid_down(const MoveEnv &e, int scan_x, int scan_y, float vy, int rt, uint32_t ty,
                                       uint32_t props) {
    uint32_t target = pix(e, scan_x, scan_y);
    const int vr = vy > 0.0f ? 1 : -1;
    int jj = scan_y + vr, steps = 0;
    bool stop = false;
    while (!stop && steps < rt) {
        if (jj < 0 || jj >= e.yr || jj < scan_y - rt || jj >= scan_y + rt) {
            stop = true;
        } else {
            const uint32_t d_ty = (uint32_t)e.tmap[pix(e, scan_x, jj)];
            const uint32_t wall = e.cellrec[wcell(e, scan_x, jj)].z;
            if (d_ty != ty || wall != 0u) {
                int ox, oy;
                const uint32_t d = do_move_at(e, props, ty, scan_x, jj, ox, oy);
                if (d == DO_DIE) {
                    target = DIE_INTENT;
                    stop = true;
                } else if (d == DO_OK) {
                    target = pix(e, ox, oy);
                    stop = true;
                } else if (d_ty != ty || wall != 4u) {
                    stop = true;
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

// scan_liquid_perpendicular (liquid.rs:801-963).
__device__ uint32_t scan_liquid_perpendicular(const MoveEnv &e, float cxf, float cyf, int cx, int cy, int rdir,
                                              int rt, float gravity, uint32_t props, uint32_t ty) {
    uint32_t target = NO_INTENT;
    float nxf = cxf, nyf = cyf, pgx = 0.0f, pgy = 0.0f;
    int nx = cx, ny = cy, j = 0;
    bool moved = false, stop = false;
    while (!stop && j < rt) {
        float gx, gy;
        base_gravity(e, gravity, (float)nx, (float)ny, gx, gy);
        float gmv = fabsf(gx);
        if (fabsf(gy) > gmv) gmv = fabsf(gy);
        if (gmv < 0.0001f) {
            stop = true;
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
            if (nx < 0 || ny < 0 || nx >= e.xr || ny >= e.yr) {
                stop = true;
            } else {
                const uint32_t d_ty = (uint32_t)e.tmap[pix(e, nx, ny)];
                const uint32_t wall = e.cellrec[wcell(e, nx, ny)].z;
                if (d_ty != ty || wall != 0u) {
                    int ox, oy;
                    const uint32_t d = do_move_at(e, props, ty, nx, ny, ox, oy);
                    if (d == DO_DIE) {
                        target = DIE_INTENT;
                        stop = true;
                    } else if (d == DO_OK) {
                        target = pix(e, ox, oy);
                        moved = true;
                        stop = true;
                    } else if (d_ty != ty || wall != 4u) {
                        stop = true;
                    }
                }
            }
        }
        if (!stop) j++;
    }
    if (moved) {
        int j2 = 0;
        bool stop2 = false;
        while (!stop2 && j2 < rt) {
            float gx, gy;
            base_gravity(e, gravity, (float)nx, (float)ny, gx, gy);
            float gmv = fabsf(gx);
            if (fabsf(gy) > gmv) gmv = fabsf(gy);
            if (gmv < 0.0001f) {
                stop2 = true;
            } else {
                gx /= gmv;
                gy /= gmv;
                nxf += gx;
                nyf += gy;
                const int nnx = round_i32(nxf), nny = round_i32(nyf);
                if (nnx < 0 || nny < 0 || nnx >= e.xr || nny >= e.yr) {
                    stop2 = true;
                } else {
                    const uint32_t d_ty = (uint32_t)e.tmap[pix(e, nnx, nny)];
                    const uint32_t wall = e.cellrec[wcell(e, nnx, nny)].z;
                    if (d_ty != ty || wall != 0u) {
                        int ox, oy;
                        const uint32_t d = do_move_at(e, props, ty, nnx, nny, ox, oy);
                        if (d == DO_DIE) {
                            target = DIE_INTENT;
                            stop2 = true;
                        } else if (d == DO_OK) {
                            target = pix(e, ox, oy);
                            stop2 = true;
                        } else if (d_ty != ty || wall != 4u) {
                            stop2 = true;
                        }
                    }
                }
            }
            if (!stop2) j2++;
        }
    }
    return target;
}

// fall_down_destination (liquid.rs:388-617).
__device__ uint32_t fall_down_destination(const MoveEnv &e, float &vx, float &vy, float px, float py, int xi, int yi,
                                          int sx, int sy, uint32_t props, uint32_t ty, uint32_t falldown,
                                          uint32_t flags, float collision, uint32_t h2, float gravity,
                                          int32_t gel_tmp) {
    Plan p = plan_move(e, vx, vy, px, py, xi, yi, props, ty);
    if (p.fin_x == xi && p.fin_y == yi && (sx != 0 || sy != 0)) {
        p.fin_x = xi + sx;
        p.fin_y = yi + sy;
        p.clear_x = xi;
        p.clear_y = yi;
        p.clear_xf = px;
        p.clear_yf = py;
    }
    uint32_t target = NO_INTENT;
    if (p.fin_x == xi && p.fin_y == yi) return target;
    int ox, oy;
    const uint32_t d0 = do_move_at(e, props, ty, p.fin_x, p.fin_y, ox, oy);
    if (d0 == DO_DIE) target = DIE_INTENT;
    else if (d0 == DO_OK) target = pix(e, ox, oy);
    else saw_wood_at(e, ox, oy, vx, vy);
    if (target == NO_INTENT && p.fin_x != xi) {
        const uint32_t d1 = do_move_at(e, props, ty, p.fin_x, p.clear_y, ox, oy);
        if (d1 == DO_DIE) target = DIE_INTENT;
        else if (d1 == DO_OK) target = pix(e, ox, oy);
        else saw_wood_at(e, ox, oy, vx, vy);
    }
    if (target == NO_INTENT && p.fin_y != yi) {
        const uint32_t d2 = do_move_at(e, props, ty, p.clear_x, p.fin_y, ox, oy);
        if (d2 == DO_DIE) target = DIE_INTENT;
        else if (d2 == DO_OK) target = pix(e, ox, oy);
        else saw_wood_at(e, ox, oy, vx, vy);
    }
    if (target != NO_INTENT) return target;
    // liquid_neighbourhood (liquid.rs:264-286)
    uint32_t nt = 0, space = 0;
    for (int oy2 = 0; oy2 < 3; ++oy2)
        for (int ox2 = 0; ox2 < 3; ++ox2) {
            if (ox2 == 1 && oy2 == 1) continue;
            const int nx = xi + ox2 - 1, ny = yi + oy2 - 1;
            uint32_t nty = 0u;
            if (nx >= 0 && nx < e.xr && ny >= 0 && ny < e.yr) nty = (uint32_t)e.tmap[ny * e.xr + nx];
            if (nty == 0u) space++;
            if (nty != ty) nt++;
        }
    const int rdir = (h2 & 2u) == 0u ? 1 : -1;
    const bool clear_moved = p.clear_x != xi || p.clear_y != yi;
    const bool moving = fabsf(vx) > 0.01f || fabsf(vy) > 0.01f;
    if ((clear_moved || nt > 0u || space > 0u) && moving)
        target = diagonal_liquid_move(e, vx, vy, p.clear_xf, p.clear_yf, rdir, props, ty);
    if (target == NO_INTENT) {
        if (falldown > 1u) {
            if (e.newton_grav == 0u && e.gravity_mode == 0u && vy > fabsf(vx)) {
                int rt = ((flags & 0x01u) && nt == 0u) ? 10 : 30;
                if (ty == 150u) rt = tpt_f2i((float)gel_tmp * 0.20f + 5.0f);
                bool fall;
                const uint32_t st =
                    scan_liquid_vertical(e, xi, yi, p.fin_y, p.clear_x, p.clear_y, rdir, rt, ty, props, fall);
                if (st != NO_INTENT) {
                    if (fall)
                        target = scroll_liquid_down(e, (int)(st % (uint32_t)e.xr), (int)(st / (uint32_t)e.xr), vy, rt,
                                                    ty, props);
                    else
                        target = st;
                }
            } else {
                float gx, gy;
                base_gravity(e, gravity, (float)xi, (float)yi, gx, gy);
                float gnorm = fabsf(gx);
                if (fabsf(gy) > gnorm) gnorm = fabsf(gy);
                bool aligned = false;
                if (gnorm > 0.0001f) {
                    const float ngx = gx / gnorm, ngy = gy / gnorm;
                    aligned = fabsf(ngx * vx + ngy * vy) > fabsf(ngy * vx - ngx * vy);
                }
                if (aligned) {
                    const int rt = ((flags & 0x01u) && nt == 0u) ? 10 : 30;
                    target = scan_liquid_perpendicular(e, p.clear_xf, p.clear_yf, p.clear_x, p.clear_y, rdir, rt,
                                                       gravity, props, ty);
                }
            }
        }
        if (target == NO_INTENT && clear_moved) {
            const uint32_t d = do_move_at(e, props, ty, p.clear_x, p.clear_y, ox, oy);
            if (d == DO_DIE) target = DIE_INTENT;
            else if (d == DO_OK) target = pix(e, ox, oy);
        }
    }
    vx *= collision;
    vy *= collision;
    return target;
}

// gas_destination (velocity.rs:755-928) without the glass refraction.
// Photon refraction at glass (velocity.rs:262-647; Simulation.cpp:1626-1750,
// 886-922, 2966-3005).  Only GLAS (24) and BGLA (95) bend light.
__device__ __forceinline__ bool refract_blocking(const MoveEnv &e, int x, int y) {
    if (x < 0 || y < 0 || x >= e.xr || y >= e.yr) return false;
    const uint32_t t = e.tmap[y * e.xr + x];
    return t == 24u || t == 95u;
}

__device__ bool refract_boundary(const MoveEnv &e, int x, int y) {
    return refract_blocking(e, x, y) && !(refract_blocking(e, x, y - 1) && refract_blocking(e, x, y + 1) &&
                                          refract_blocking(e, x - 1, y) && refract_blocking(e, x + 1, y));
}

// The 8-neighbour ring (x right, y down, counter-clockwise from +x) and the
// directions each step may continue in.
__device__ __forceinline__ int ring_dx(int i) { return (i == 0 || i == 1 || i == 7) ? 1 : ((i >= 3 && i <= 5) ? -1 : 0); }
__device__ __forceinline__ int ring_dy(int i) { return (i >= 1 && i <= 3) ? 1 : ((i >= 5) ? -1 : 0); }
__device__ __forceinline__ uint32_t ring_mask(int i) {
    const uint32_t m = 0x07u << ((i + 7) & 7);  // bits i-1, i, i+1 (mod 8)
    return (m | (m >> 8)) & 0xFFu;
}

// find_next_boundary (velocity.rs:343-389).
__device__ bool next_boundary(const MoveEnv &e, int &x, int &y, uint32_t dm, int &em, bool reverse) {
    if (!(x > 0 && x < e.xr - 1 && y > 0 && y < e.yr - 1)) return false;
    if (em != -1) dm &= ring_mask(em);
    uint32_t mask = 0u;
    for (int i = 0; i < 8; ++i)
        if ((dm & (1u << i)) && refract_blocking(e, x + ring_dx(i), y + ring_dy(i))) mask |= 1u << i;
    for (int i = 0; i < 8; ++i) {
        const int n = (i + (reverse ? 1 : 7)) & 7;
        if ((mask & (1u << i)) && !(mask & (1u << n))) {
            x += ring_dx(i);
            y += ring_dy(i);
            em = i;
            return true;
        }
    }
    return false;
}

// get_normal (velocity.rs:401-456).
__device__ bool surface_normal(const MoveEnv &e, int x, int y, float dx, float dy, float &nx, float &ny) {
    if (!((dx != 0.0f || dy != 0.0f) && refract_boundary(e, x, y))) return false;
    int lx = x, ly = y, rx = x, ry = y, lm = -1, rm = -1, j = 0;
    bool lv = true, rv = true;
    for (int k = 0; k < 10; ++k) {
        if (lv) lv = next_boundary(e, lx, ly, 0xFFu, lm, true);
        if (rv) rv = next_boundary(e, rx, ry, 0xFFu, rm, false);
        j += (lv ? 1 : 0) + (rv ? 1 : 0);
        if (!lv && !rv) break;
    }
    if (!(j >= 3 && (lx != rx || ly != ry))) return false;
    const float ex = (float)(rx - lx), ey = (float)(ry - ly);
    const float inv = 1.0f / sqrtf(ex * ex + ey * ey);
    nx = ey * inv;
    ny = -ex * inv;
    return true;
}

// get_normal_interp (velocity.rs:462-516).
__device__ bool surface_normal_along(const MoveEnv &e, float x0, float y0, float dx, float dy, float &nx, float &ny) {
    dx = dx / 16.0f;
    dy = dy / 16.0f;
    for (int k = 0; k < 20; ++k) {
        const int x = tpt_f2i(x0 + 0.5f), y = tpt_f2i(y0 + 0.5f);
        if (x < 0 || y < 0 || x >= e.xr || y >= e.yr) return false;
        if (refract_boundary(e, x, y)) return surface_normal(e, x, y, dx, dy, nx, ny);
        x0 += dx;
        y0 += dy;
    }
    return false;
}

// get_wavelength_bin (velocity.rs:523-560); -1: no wavelength left.
__device__ int wavelength_bin(uint32_t wm, uint32_t seed, uint32_t &out) {
    const uint32_t masked = wm & 0x3FFFFFFFu;
    out = wm;
    if (!masked) return -1;
    const uint32_t lo = (uint32_t)__ffs(masked) - 1u, hi = 31u - (uint32_t)__clz(masked);
    if (hi - lo < 5u) return (int)(hi + lo);
    const uint32_t r = pcg_hash(seed ^ (lo * 2654435769u));
    const uint32_t i = ((r >> 1) % (hi - lo - 4u)) + lo;
    if (r & 1u) {
        out = wm & (0x1Fu << i);
        return (int)((i + 2u) * 2u);
    }
    out = wm & (0xFu << i);
    return (int)((i + 2u) * 2u - 1u);
}

// refract_photon (velocity.rs:577-637): 0 passes, 1 dies, 2 total internal
// reflection (stays this tick).
__device__ int refract_photon(const MoveEnv &e, float px, float py, int xi, int yi, float &vx, float &vy, int fx,
                              int fy, uint32_t slot, uint32_t seed) {
    const uint32_t rt = e.tmap[fy * e.xr + fx], lt = e.tmap[yi * e.xr + xi];
    const bool r_glass = rt == 24u || rt == 95u, l_glass = lt == 24u || lt == 95u;
    if (r_glass == l_glass) return 0;
    float nrx, nry;
    if (!surface_normal_along(e, px, py, vx, vy, nrx, nry)) return 1;
    uint32_t narrowed;
    const int bin = wavelength_bin(e.ctype[slot], seed, narrowed);
    if (bin == -1) return 1;
    e.ctype[slot] = narrowed;
    float nn = 1.9f - 0.07f * ((float)bin - 30.0f) / 30.0f;
    nn *= nn;
    if (r_glass) {  // entering the glass
        nrx = -nrx;
        nry = -nry;
        nn = 1.0f / nn;
    }
    const float ct1 = vx * nrx + vy * nry;
    const float ct2 = 1.0f - (nn * nn) * (1.0f - (ct1 * ct1));
    if (ct2 < 0.0f) {
        vx -= 2.0f * ct1 * nrx;
        vy -= 2.0f * ct1 * nry;
        return 2;
    }
    const float ct2s = sqrtf(ct2) - nn * ct1;
    vx = nn * vx + ct2s * nrx;
    vy = nn * vy + ct2s * nry;
    return 0;
}

__device__ uint32_t gas_destination(const MoveEnv &e, float &vx, float &vy, float px, float py, int xi, int yi, int sx,
                                    int sy, uint32_t props, uint32_t ty, uint32_t slot, float collision,
                                    uint32_t seed) {
    const Plan p = plan_move(e, vx, vy, px, py, xi, yi, props, ty);
    int fin_x = p.fin_x, fin_y = p.fin_y;
    if (fin_x == xi && fin_y == yi && (sx != 0 || sy != 0)) {
        fin_x = xi + sx;
        fin_y = yi + sy;
    }
    uint32_t target = NO_INTENT;
    bool blocked = false, died = false, tir = false;
    if (fin_x != xi || fin_y != yi) {
        int ox, oy;
        const uint32_t d = do_move_at(e, props, ty, fin_x, fin_y, ox, oy);
        if (d == DO_OK) {
            if (ty == 29u) {
                const int r = refract_photon(e, px, py, xi, yi, vx, vy, ox, oy, slot, seed);
                died = r == 1;
                tir = r == 2;
            }
            if (!died && !tir) target = pix(e, ox, oy);
        } else if (d == DO_DIE) {
            died = true;
        } else {
            blocked = true;
            if (ty == 29u) {
                const uint32_t dest_px = pix(e, ox, oy);
                const uint32_t dest_ty = (uint32_t)e.tmap[dest_px];
                const uint32_t src_ty = (uint32_t)e.tmap[pix(e, xi, yi)];
                const uint32_t roll = pcg_hash(seed ^ (slot * 0x9E3779B9u) ^ (dest_px * 0x85EBCA6Bu));
                if ((dest_ty == 46u && (roll & 1u) == 0u) || src_ty == 46u) {
                    blocked = false;
                    target = HALF_MIRROR | dest_px;
                } else {
                    target = 0x80000000u | dest_px;
                    saw_wood_at(e, ox, oy, vx, vy);
                }
            } else {
                saw_wood_at(e, ox, oy, vx, vy);
            }
        }
    }
    if (died) {
        target = DIE_INTENT;
    } else if (blocked) {
        // Each axis is scaled once, after the probes (which read the
        // unscaled velocity), by `collision` or by 1 (exact).  Scaling in
        // the branches let the compiler merge the stores through a selected
        // address, which put vx/vy in scratch memory for the whole kernel.
        float kx = 1.0f, ky = 1.0f;
        if (ty == 29u) {
            if (fin_x != xi) kx = collision;
            if (fin_y != yi) ky = collision;
        } else {
            const int bx = min(max(fin_x, xi - 2), xi + 2), by = min(max(fin_y, yi - 2), yi + 2);
            const int rx = 2 * xi - bx, ry = 2 * yi - by;
            int ox, oy;
            const uint32_t d1 = do_move_at(e, props, ty, rx, by, ox, oy);
            if (d1 != DO_BLOCKED) {
                if (d1 == DO_DIE) target = DIE_INTENT;
                else if (ox != xi || oy != yi) target = pix(e, ox, oy);
                kx = collision;
            } else {
                saw_wood_at(e, ox, oy, vx, vy);
                const uint32_t d2 = do_move_at(e, props, ty, bx, ry, ox, oy);
                if (d2 != DO_BLOCKED) {
                    if (d2 == DO_DIE) target = DIE_INTENT;
                    else if (ox != xi || oy != yi) target = pix(e, ox, oy);
                    ky = collision;
                } else {
                    saw_wood_at(e, ox, oy, vx, vy);
                    kx = collision;
                    ky = collision;
                }
            }
        }
        vx *= kx;
        vy *= ky;
    }
    return target;
}

__global__ __launch_bounds__(BLK) void mv_air_deposit(Slots s, const uint32_t *__restrict__ wall,
                                                      const uint32_t *__restrict__ emap, uint32_t *__restrict__ key,
                                                      uint32_t slots, int xres, int yres, int xc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint32_t ty = s.ty[i];
    if (ty == 0u || ty >= 256u) return;
    const uint32_t flags = s.flags[i];
    if ((flags & 0x10u) || (flags & 0x02u)) return;
    const int xi = tpt_f2i(s.x[i] + 0.5f), yi = tpt_f2i(s.y[i] + 0.5f);
    if (xi < 4 || yi < 4 || xi >= xres - 4 || yi >= yres - 4) return;
    const int cell = (yi / TPT_CELL) * xc + xi / TPT_CELL;
    if (wall[cell] == 18u && (emap[cell] & 0xFFu) < 8u) return;
    atomicMax(&key[cell], i + 1u);
}

__global__ void mv_air_deposit_cells(Slots s, const uint32_t *__restrict__ lut, float *__restrict__ air_vx,
                                     float *__restrict__ air_vy, uint32_t *__restrict__ key, uint32_t cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const uint32_t k = key[c];
    if (!k) return;
    const uint32_t i = k - 1u, base = s.ty[i] * 12u + 4u;
    const float drag = __uint_as_float(lut[base + 1u]), loss = __uint_as_float(lut[base + 2u]);
    air_vx[c] = air_vx[c] * loss + drag * s.vx[i];
    air_vy[c] = air_vy[c] * loss + drag * s.vy[i];
    key[c] = 0u;
}

// The same reduction over the pixel records: one thread per cell looks at
// the owners of its 16 pixels from the highest slot down and stops at the
// first eligible one; slots that own no pixel (the off-map list) were put
// into key[] by mv_air_deposit_list first.  A record's pixel is the rounded
// position the eligibility test uses, and cells within one cell of the
// border lie inside the 4-pixel margin entirely, as do stasis cells.
__global__ __launch_bounds__(BLK) void mv_air_deposit_list(Slots s, const uint32_t *__restrict__ wall,
                                                           const uint32_t *__restrict__ emap,
                                                           uint32_t *__restrict__ key,
                                                           const uint32_t *__restrict__ list,
                                                           const uint32_t *__restrict__ counters, uint32_t slots,
                                                           int xres, int yres, int xc) {
    const uint32_t n = counters[TPT_CTR_OFFMAP];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t i = list[k];
        if (i >= slots) continue;
        const uint32_t ty = s.ty[i];
        if (ty == 0u || ty >= 256u) continue;
        const uint32_t flags = s.flags[i];
        if ((flags & 0x10u) || (flags & 0x02u)) continue;
        const int xi = tpt_f2i(s.x[i] + 0.5f), yi = tpt_f2i(s.y[i] + 0.5f);
        if (xi < 4 || yi < 4 || xi >= xres - 4 || yi >= yres - 4) continue;
        const int cell = (yi / TPT_CELL) * xc + xi / TPT_CELL;
        if (wall[cell] == 18u && (emap[cell] & 0xFFu) < 8u) continue;
        atomicMax(&key[cell], i + 1u);
    }
}

__global__ __launch_bounds__(BLK) void mv_air_deposit_px(Slots s, const ulonglong2 *__restrict__ rec,
                                                         const uint32_t *__restrict__ wall,
                                                         const uint32_t *__restrict__ emap,
                                                         const uint32_t *__restrict__ lut, float *__restrict__ air_vx,
                                                         float *__restrict__ air_vy, uint32_t *__restrict__ key,
                                                         uint32_t slots, int xres, int xc, int yc) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= xc * yc) return;
    const int cx = c % xc, cy = c / xc;
    uint32_t best = key[c];
    if (best) key[c] = 0u;
    if (cx >= 1 && cy >= 1 && cx <= xc - 2 && cy <= yc - 2 && !(wall[c] == 18u && (emap[c] & 0xFFu) < 8u)) {
        uint32_t cand[16];
        for (int r = 0; r < 4; ++r) {
            const size_t p2 = ((size_t)(cy * 4 + r) * xres + cx * 4) >> 1;
            const ulonglong2 a = rec[p2], b2 = rec[p2 + 1];
            cand[r * 4 + 0] = (uint32_t)(a.x >> 41);
            cand[r * 4 + 1] = (uint32_t)(a.y >> 41);
            cand[r * 4 + 2] = (uint32_t)(b2.x >> 41);
            cand[r * 4 + 3] = (uint32_t)(b2.y >> 41);
        }
        // cand = slot + 1 of each pixel's owner (0: none).
        for (;;) {
            int m = -1;
            uint32_t mv = best;
            for (int q = 0; q < 16; ++q)
                if (cand[q] > mv) {
                    mv = cand[q];
                    m = q;
                }
            if (m < 0) break;
            cand[m] = 0u;
            const uint32_t i = mv - 1u;
            if (i >= slots) continue;
            const uint32_t ty = s.ty[i];
            if (ty == 0u || ty >= 256u) continue;
            if (s.flags[i] & 0x12u) continue;
            best = mv;
            break;
        }
    }
    if (!best) return;
    const uint32_t i = best - 1u, base = s.ty[i] * 12u + 4u;
    const float drag = __uint_as_float(lut[base + 1u]), loss = __uint_as_float(lut[base + 2u]);
    air_vx[c] = air_vx[c] * loss + drag * s.vx[i];
    air_vy[c] = air_vy[c] * loss + drag * s.vy[i];
}

// WL_DETECT charging and the set_emap closure (mod.rs:348-828,
// Simulation.cpp:928-983, 2330-2331).  Emap values are the low 6 bits;
// bits 7/6 mark the left/right ends of a span filled this tick.
//
// A particle in an uncharged detector charges the whole uncharged wire run
// through its cell.  The reference fills runs from concurrent threads that
// read and write emap at once; here every trigger computes its run from the
// untouched emap (mv_detect_span) and a second kernel fills the runs
// (mv_detect_fill): each cell gets one value whichever trigger wrote it, the
// union of the triggered runs, as the serial slot order gives.
constexpr uint32_t EMAP_CHARGED = 16u, EMAP_LEFT = 0x80u, EMAP_RIGHT = 0x40u, EMAP_ENDS = 0xC0u;

__device__ __forceinline__ bool is_wire(uint32_t w) {
    return w == 3u || w == 2u || w == 6u || w == 1u || w == 5u || w == 12u || w == 18u;
}
__device__ __forceinline__ bool wire_off(const uint32_t *bmap, const uint32_t *emap, uint32_t c) {
    return is_wire(bmap[c]) && (emap[c] & 0x3Fu) < 8u;
}
__device__ __forceinline__ bool wire_at(const uint32_t *bmap, int x, int y, int xc, int yc) {
    return x >= 0 && y >= 0 && x < xc && y < yc && is_wire(bmap[y * xc + x]);
}
// set_emap's ladder shortcuts (mod.rs:412-456): dir -1 looks up (the cap row
// is below the target row), +1 looks down.
__device__ bool ladder(const uint32_t *bmap, bool single, int x, int target, int dir, int xc, int yc) {
    const int cap = target - dir;
    return single && wire_at(bmap, x - 1, cap, xc, yc) && wire_at(bmap, x, cap, xc, yc) &&
           wire_at(bmap, x + 1, cap, xc, yc) && !wire_at(bmap, x - 1, target, xc, yc) &&
           wire_at(bmap, x, target, xc, yc) && !wire_at(bmap, x + 1, target, xc, yc);
}

__global__ __launch_bounds__(BLK) void mv_detect_span(Slots s, const uint32_t *__restrict__ bmap,
                                                      const uint32_t *__restrict__ emap, uint32_t *__restrict__ span,
                                                      uint32_t slots, int xres, int yres, int xc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint32_t ty = s.ty[i];
    if (ty == 0u || ty >= 256u || (s.flags[i] & 0x10u)) return;
    const int xi = tpt_f2i(s.x[i] + 0.5f), yi = tpt_f2i(s.y[i] + 0.5f);
    if (xi < 4 || yi < 4 || xi >= xres - 4 || yi >= yres - 4) return;
    const uint32_t cx = (uint32_t)xi / 4u, row = ((uint32_t)yi / 4u) * (uint32_t)xc, cell = row + cx;
    if (bmap[cell] != 3u || (emap[cell] & 0x3Fu) >= 8u) return;
    uint32_t x1 = cx, x2 = cx;
    while (x1 > 0u && wire_off(bmap, emap, row + x1 - 1u)) --x1;
    while (x2 + 1u < (uint32_t)xc && wire_off(bmap, emap, row + x2 + 1u)) ++x2;
    span[cell] = ((x1 + 1u) << 16) | (x2 + 1u);  // every trigger of the run stores the same value
}

__global__ void mv_detect_fill(uint32_t *__restrict__ emap, uint32_t *__restrict__ span, int xc, uint32_t cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const uint32_t v = span[c];
    if (!v) return;
    span[c] = 0u;
    const uint32_t row = (c / (uint32_t)xc) * (uint32_t)xc, x1 = (v >> 16) - 1u, x2 = (v & 0xFFFFu) - 1u;
    for (uint32_t x = x1; x <= x2; ++x)
        emap[row + x] = EMAP_CHARGED | (x == x1 ? EMAP_LEFT : 0u) | (x == x2 ? EMAP_RIGHT : 0u);
}

// emap_gate_kernel (mod.rs:556-575): bit 0 an uncharged wire, bit 1 a cell
// charged this tick.
__global__ void mv_emap_gate(const uint32_t *__restrict__ bmap, const uint32_t *__restrict__ emap,
                             uint32_t *__restrict__ gate, uint32_t cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bits = 0u;
    if (c < cells) bits = (wire_off(bmap, emap, c) ? 1u : 0u) | ((emap[c] & 0x3Fu) == EMAP_CHARGED ? 2u : 0u);
    const uint64_t b0 = __ballot(bits & 1u), b1 = __ballot(bits & 2u);
    if (__lane_id() == 0 && (b0 || b1)) atomicOr(gate, (b0 ? 1u : 0u) | (b1 ? 2u : 0u));
}

// emap_closure_kernel (mod.rs:594-810): the recursive half of set_emap as
// repeated top-to-bottom sweeps until one changes nothing, at most ycells + 2
// of them; it returns at once unless the gate saw both an uncharged and a
// charged wire.
//
// The reference runs the sweep on one thread.  Whether an uncharged run
// fills depends only on the emap of the rows above and below it (and the
// static walls), never on its own row, so the runs of one row are
// independent: one workgroup takes a row at a time with a thread per cell,
// and rows go in the reference's order, so every row sees the rows above it
// as this sweep left them and the rows below as the previous sweep did.  The
// result is the serial sweep's, also when the pass bound stops it early.
// A cell's run start comes from a block max-scan; the run's activation is an
// LDS OR keyed by that start (double-buffered by row parity, so the next
// row can clear its buffer without another barrier).
constexpr int EMAP_THREADS = 1024;
static_assert(TPT_MAX_DIM / TPT_CELL <= (unsigned)EMAP_THREADS, "one thread per cell of a row");

__device__ __forceinline__ int block_scan_max(int v, int *wave_tot) {
    const int lane = (int)(threadIdx.x & 63u), w = (int)(threadIdx.x >> 6);
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(v, (unsigned)d, 64);
        if (lane >= d) v = v > o ? v : o;
    }
    if (lane == 63) wave_tot[w] = v;
    __syncthreads();
    int prefix = -1;
    for (int k = 0; k < w; ++k) prefix = prefix > wave_tot[k] ? prefix : wave_tot[k];
    __syncthreads();
    return v > prefix ? v : prefix;
}

// Would a charged neighbour of the uncharged cell (x, y) fill its run?  The
// three arms of the serial kernel, for this cell's column only.
__device__ bool emap_activates(const uint32_t *bmap, const uint32_t *emap, int x, int y, int xc, int yc) {
    if (y + 1 < yc) {  // charged span below
        const uint32_t se = emap[(y + 1) * xc + x];
        const int sy = y + 1;
        const bool single = (se & EMAP_ENDS) == EMAP_ENDS;
        if ((se & 0x3Fu) == EMAP_CHARGED && !ladder(bmap, single, x, sy - 2, -1, xc, yc) &&
            ((se & EMAP_ENDS) || y + 1 == yc - 1 || wire_at(bmap, x - 1, y, xc, yc) ||
             wire_at(bmap, x + 1, y, xc, yc) || wire_at(bmap, x - 1, sy + 1, xc, yc) ||
             !wire_at(bmap, x, sy + 1, xc, yc) || wire_at(bmap, x + 1, sy + 1, xc, yc)))
            return true;
    }
    if (y > 0) {  // charged span above
        const uint32_t se = emap[(y - 1) * xc + x];
        const int sy = y - 1;
        const bool single = (se & EMAP_ENDS) == EMAP_ENDS;
        if ((se & 0x3Fu) == EMAP_CHARGED && !ladder(bmap, single, x, y + 1, 1, xc, yc) &&
            ((se & EMAP_ENDS) || wire_at(bmap, x - 1, y, xc, yc) || wire_at(bmap, x + 1, y, xc, yc) ||
             wire_at(bmap, x - 1, sy - 1, xc, yc) || !wire_at(bmap, x, sy - 1, xc, yc) ||
             wire_at(bmap, x + 1, sy - 1, xc, yc)))
            return true;
    }
    // ladder jumps from a singleton two rows away
    const bool up = y + 2 < yc && (emap[(y + 2) * xc + x] & EMAP_ENDS) == EMAP_ENDS;
    const bool down = y >= 2 && (emap[(y - 2) * xc + x] & EMAP_ENDS) == EMAP_ENDS;
    return ladder(bmap, up, x, y, -1, xc, yc) || ladder(bmap, down, x, y, 1, xc, yc);
}

__global__ __launch_bounds__(EMAP_THREADS) void mv_emap_closure(const uint32_t *__restrict__ bmap,
                                                               uint32_t *__restrict__ emap,
                                                               uint32_t *__restrict__ gate, int xc, int yc) {
    __shared__ int offs[EMAP_THREADS];
    __shared__ int act[2][EMAP_THREADS];
    __shared__ int wave_tot[EMAP_THREADS / 64];
    __shared__ uint32_t g_sh;
    const int x = (int)threadIdx.x;
    if (x == 0) {
        g_sh = *gate;
        *gate = 0u;
    }
    __syncthreads();
    if (g_sh != 3u) return;
    const bool col = x < xc;
    bool changed = true;
    for (int pass = 0; changed && pass < yc + 2; ++pass) {
        bool mine = false;
        for (int y = 0; y < yc; ++y) {
            const uint32_t c = (uint32_t)(y * xc + x);
            const bool off = col && wire_off(bmap, emap, c);
            int *a = act[y & 1];
            a[x] = 0;
            offs[x] = off ? 1 : 0;
            if (!__syncthreads_or(off ? 1 : 0)) continue;
            const bool first = off && (x == 0 || offs[x - 1] == 0);
            const bool last = off && (x + 1 == xc || offs[x + 1] == 0);
            const int rs = block_scan_max(first ? x : -1, wave_tot);
            if (off && emap_activates(bmap, emap, x, y, xc, yc)) atomicOr(&a[rs], 1);
            __syncthreads();
            if (off && a[rs]) {
                emap[c] = EMAP_CHARGED | (first ? EMAP_LEFT : 0u) | (last ? EMAP_RIGHT : 0u);
                mine = true;
            }
        }
        changed = __syncthreads_or(mine ? 1 : 0) != 0;
    }
}

__global__ void mv_emap_strip(uint32_t *__restrict__ emap, uint32_t cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cells) {
        const uint32_t v = emap[c];
        if (v & EMAP_ENDS) emap[c] = v & ~EMAP_ENDS;
    }
}

// Everything the movement pass needs rebuilt before the intent kernel, in
// one launch of max(P/4, 512) threads:
//  * cell_winner = 0 (clear_move_intents_kernel, mod.rs:268-273), 16-byte
//    stores;
//  * the planner's byte type map from pmap;
//  * the per-cell record MoveEnv::cellrec, after the air deposit and the
//    wire passes have written their grids;
//  * the allocation window's free bitmap cleared (the intent kernel fills
//    the words of the slots it covers) and the pass's list counters reset.
__global__ __launch_bounds__(BLK) void mv_prepare(uint4 *__restrict__ winner, const uint4 *__restrict__ pmap,
                                                  uint32_t *__restrict__ tmap, uint32_t n4,
                                                  const uint32_t *__restrict__ air_vx,
                                                  const uint32_t *__restrict__ air_vy,
                                                  const uint32_t *__restrict__ wall, const uint32_t *__restrict__ emap,
                                                  uint4 *__restrict__ rec, uint32_t cells,
                                                  uint32_t *__restrict__ spawn_free, uint32_t *__restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        counters[TPT_CTR_SAW] = 0u;
        counters[TPT_CTR_PAIR] = 0u;
        counters[TPT_CTR_CREATE] = 0u;
    }
    if (i < TPT_SPAWN_BITMAP_WORDS) spawn_free[i] = 0u;
    if (i < n4) {
        winner[i] = make_uint4(0u, 0u, 0u, 0u);
        const uint4 v = pmap[i];
        tmap[i] = (v.x & 0xFFu) | ((v.y & 0xFFu) << 8) | ((v.z & 0xFFu) << 16) | ((v.w & 0xFFu) << 24);
    }
    if (i < cells) rec[i] = make_uint4(air_vx[i], air_vy[i], wall[i], emap[i]);
}

// velocity_and_intent_kernel (velocity.rs:951-1157).
__device__ __forceinline__ void intent_one(Slots &s, const MoveEnv &e, uint32_t *__restrict__ winner, uint32_t i,
                                           uint32_t slots, uint32_t seed) {
    if (i >= slots) return;
    s.intent[i] = NO_INTENT;
    // Slot fields loaded together rather than one latency per test.
    const uint32_t ty = s.ty[i], flags = s.flags[i];
    const float px = s.x[i], py = s.y[i], vx0 = s.vx[i], vy0 = s.vy[i];
    const int32_t tmp0 = s.tmp[i];
    if (ty == 0u || ty >= 256u) return;
    if ((flags & 0x10u) || (flags & 0x02u)) return;
    const int xi = tpt_f2i(px + 0.5f), yi = tpt_f2i(py + 0.5f);
    if (xi < 0 || xi >= e.xr || yi < 0 || yi >= e.yr) {
        s.ty[i] = 0u;
        return;
    }
    const int cell = wcell(e, xi, yi);
    const uint32_t meta = ty * 12u;
    const uint32_t props = e.lut[meta], falldown = e.lut[meta + 1u], rank = e.lut[meta + 2u];
    const float advection = __uint_as_float(e.lut[meta + 4u]), loss = __uint_as_float(e.lut[meta + 7u]);
    const float collision = __uint_as_float(e.lut[meta + 8u]), gravity = __uint_as_float(e.lut[meta + 9u]);
    const float diffusion = __uint_as_float(e.lut[meta + 10u]);
    const uint4 crec = e.cellrec[cell];
    const uint32_t wall = crec.z, em = crec.w;
    const bool in_margin = !(xi < 4 || yi < 4 || xi >= e.xr - 4 || yi >= e.yr - 4);
    const bool stickman = ty == 59u || ty == 137u || ty == 159u;
    if (!in_margin || (!stickman && wall_kills_particle(wall, em, props))) {
        s.ty[i] = 0u;
        return;
    }
    if (wall == 18u && (em & 0xFFu) < 8u) return;
    float gx, gy;
    base_gravity(e, gravity, (float)xi, (float)yi, gx, gy);
    const float ax = __uint_as_float(crec.x), ay = __uint_as_float(crec.y);
    const uint32_t h1 = pcg_hash(seed ^ (i * 2654435769u)), h2 = pcg_hash(h1);
    const float loss_here = e.newton_grav ? 1.0f : loss;
    const float rdx = (float)(h1 & 0xFFFFu) * (2.0f / 65535.0f) - 1.0f;
    const float rdy = (float)(h2 & 0xFFFFu) * (2.0f / 65535.0f) - 1.0f;
    float vx = vx0 * loss_here + advection * ax + gx + diffusion * rdx;
    float vy = vy0 * loss_here + advection * ay + gy + diffusion * rdy;
    float mv = fabsf(vx);
    if (fabsf(vy) > mv) mv = fabsf(vy);
    if (mv > MAX_VELOCITY) {
        vx = vx * (MAX_VELOCITY / mv);
        vy = vy * (MAX_VELOCITY / mv);
    }
    const int sx = step_axis(vx, h1), sy = step_axis(vy, h2);
    uint32_t target;
    if (falldown == 0u)
        target = gas_destination(e, vx, vy, px, py, xi, yi, sx, sy, props, ty, i, collision, seed);
    else
        target = fall_down_destination(e, vx, vy, px, py, xi, yi, sx, sy, props, ty, falldown, flags, collision, h2,
                                       gravity, tmp0);
    if (target == DIE_INTENT) {
        s.ty[i] = 0u;
        return;
    }
    if (target != NO_INTENT) {
        s.intent[i] = target;
        if ((target & 0x80000000u) == 0u) {
            const uint32_t key = ((rank & 0x1FFu) << 23) | ((i + 1u) & 0x7FFFFFu);
            atomicMax(&winner[target & PIXEL_MASK], key);
            if (pair_tracked(ty)) e.pair_list[atomicAdd(&e.counters[TPT_CTR_PAIR], 1u)] = i;
        }
    }
    s.vx[i] = vx;
    s.vy[i] = vy;
}

// After its slots' intents (which kill out-of-bounds particles), a block
// over the device allocation window records which of them are free: empty
// and not portal-held (collect_gpu_free_slots_kernel,
// reaction_kernel/shared.rs:1197-1222), as begin_pending_creates samples
// them between the intent and pair launches (move_kernel/mod.rs:1203).
// A slot's type is written only by its own lane in this kernel.
// Eight waves per SIMD (64 VGPRs): the kernel waits on dependent random
// loads, and the extra waves hide more of that than the few spilled
// registers cost (measured).
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(8))) void mv_intent(Slots s, MoveEnv e, uint32_t *__restrict__ winner, uint32_t slots,
                                                 const uint32_t *__restrict__ seedp, uint32_t capacity,
                                                 uint32_t *__restrict__ spawn_free) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    intent_one(s, e, winner, i, slots, *seedp);
    const uint32_t first = blockIdx.x * blockDim.x;
    if (first + blockDim.x <= TPT_GPU_SPAWN_START || first >= TPT_GPU_SPAWN_END) return;
    __syncthreads();
    const bool in_window = i >= TPT_GPU_SPAWN_START && i < TPT_GPU_SPAWN_END;
    const bool free_slot = in_window && i < capacity && s.ty[i] == 0u && (s.flags[i] & 0x04u) == 0u;
    const uint64_t bits = __ballot(free_slot);
    const uint32_t lane0 = i & ~63u;
    if ((threadIdx.x & 63u) == 0u && lane0 >= TPT_GPU_SPAWN_START && lane0 < TPT_GPU_SPAWN_END) {
        const uint32_t w = (lane0 - TPT_GPU_SPAWN_START) >> 5;
        spawn_free[w] = (uint32_t)bits;
        spawn_free[w + 1u] = (uint32_t)(bits >> 32);
    }
}

// ---- e == 2 pair pass (reaction_kernel/energy_entry.rs:90-393) ------------
//
// Canonical form (the reference resolves pairs concurrently):
//  * every mover reads its own type as the intent pass left it: the
//    reference's transient claim of the occupant's type (PAIR_CLAIMED) is
//    not modelled, since each occupant has at most one mover (one winner per
//    destination pixel, one pixel per occupant);
//  * creation records are allocated in ascending requester slot order, the
//    k-th record taking the k-th free slot of the window in ascending order;
//    at most TPT_GPU_CREATE_COUNT records, the lowest requester slots.
// The arms touch disjoint data otherwise, so no other order matters.

// Element_FILT_getWavelengths (reaction_kernel/rays/emitter.rs:10-26).
__device__ __forceinline__ uint32_t filt_wavelengths(uint32_t ctype, float temp) {
    if ((ctype & 0x3FFFFFFFu) != 0u) return ctype & 0x3FFFFFFFu;
    int bin = tpt_f2i((temp - 273.0f) * 0.025f);
    bin = bin < 0 ? 0 : (bin > 25 ? 25 : bin);
    return 0x1Fu << bin;
}

// Element_FILT_interactWavelengths (reaction_kernel/rays/emitter.rs:28-80).
// Shift counts use their low five bits, as the reference's u32 shifts do.
__device__ uint32_t interact_wavelengths(int32_t filt_tmp, uint32_t filt_ctype, float filt_temp, uint32_t wl) {
    const uint32_t mask = 0x3FFFFFFFu;
    const uint32_t fw = filt_wavelengths(filt_ctype, filt_temp);
    switch (filt_tmp) {
    case 0: return fw;
    case 1: return wl & fw;
    case 2: return wl | fw;
    case 3: return wl & ~fw;
    case 4:
    case 5: {
        int shift = tpt_f2i((filt_temp - 273.0f) * 0.025f);
        if (shift <= 0) shift = 1;
        return (filt_tmp == 4 ? wl << ((uint32_t)shift & 31u) : wl >> ((uint32_t)shift & 31u)) & mask;
    }
    case 6: return wl;
    case 7: return wl ^ fw;
    case 8: return ~wl & mask;
    case 10: return (wl * (fw & (~fw + 1u))) & mask;
    case 11: {
        const uint32_t lsb = fw & (~fw + 1u);
        return lsb != 0u ? (wl / lsb) & mask : wl;
    }
    default: return fw;
    }
}

struct PairBufs {
    uint32_t *ctype;
    int32_t *life, *tmp2;
    float *temp;
    const float *air_pv;
    const uint32_t *defaults;  // spawn defaults, TPT_SPAWN_DEFAULT_WORDS per element
    uint32_t *rec;             // creation records
    uint32_t rec_cap;
};

__device__ __forceinline__ void queue_photon(const PairBufs &pb, uint32_t *counters, uint32_t slot, int x, int y,
                                             uint32_t ctype, float temp, float vx, float vy) {
    const uint32_t k = atomicAdd(&counters[TPT_CTR_CREATE], 1u);
    if (k >= pb.rec_cap) return;
    uint32_t *r = pb.rec + (size_t)k * TPT_CREATE_REC_WORDS;
    r[0] = slot;
    r[1] = (uint32_t)x | ((uint32_t)y << 16);
    r[2] = ctype;
    r[3] = __float_as_uint(temp);
    r[4] = __float_as_uint(vx);
    r[5] = __float_as_uint(vy);
}

// pair_resolve_kernel's body for mover i (energy_entry.rs:111-390).
__device__ void pair_one(Slots s, const PairBufs &pb, const uint32_t *__restrict__ pmap,
                         const uint32_t *__restrict__ winner, uint32_t *counters, uint32_t i, uint32_t slots,
                         int xres, int yres, int xc, uint32_t frame_seed) {
    const uint32_t ty = s.ty[i];
    if (ty == 0u || ty >= 256u || !pair_tracked(ty)) return;
    const uint32_t raw = s.intent[i];
    if (raw == NO_INTENT || (raw & 0x80000000u) != 0u) return;
    const uint32_t dest = raw & PIXEL_MASK;
    if (dest >= (uint32_t)(xres * yres)) return;
    const float xf = s.x[i], yf = s.y[i];
    const int xi = tpt_f2i(xf + 0.5f), yi = tpt_f2i(yf + 0.5f);
    const uint32_t own = (uint32_t)(yi * xres + xi);
    if (own == dest) return;
    const uint32_t m = pmap[dest];
    if (m < 512u) return;
    const uint32_t rty = m & 0x1FFu, r = (m >> 9) - 1u;
    if (r >= slots) return;
    // The winner test masks 27 bits (energy_entry.rs:171-175), so it also
    // requires the low four bits of the mover's weight rank to be zero.
    if ((winner[dest] & 0x07FFFFFFu) != ((i + 1u) & 0x07FFFFFFu)) return;
    // The occupant claim (a CAS from the map's type, energy_entry.rs:181-185)
    // fails on a slot the intent pass killed or a WOOD it sawed.  Matter
    // types change in this pass only through the H2 arm, on the thread's
    // own occupant, so the read is stable.
    if (s.ty[r] != rty) return;
    const uint32_t seed = fmix32(frame_seed ^ fmix32(i));
    uint32_t final_ty = rty;
    if (ty == PT_PHOT) {
        if (rty == PT_GLOW) {
            if (pb.life[r] == 0 && seed % 30u == 0u) {
                pb.life[r] = 120;
                int lr = -1;
                if (fmix32(seed ^ 0x68E31DA9u) % 2u == 0u) lr = 1;
                const float fx = xf - (float)lr * 0.3f * s.vy[i];
                const float fy = yf + (float)lr * 0.3f * s.vx[i];
                const int gx = tpt_f2i(fx + 0.5f), gy = tpt_f2i(fy + 0.5f);
                if (gx >= 0 && gx < xres && gy >= 0 && gy < yres) {
                    const uint32_t g = pmap[gy * xres + gx];
                    const uint32_t gslot = (g >> 9) - 1u;
                    if (g >= 512u && (g & 0x1FFu) == PT_GLOW && gslot < slots) {
                        int bin = tpt_f2i((pb.temp[gslot] - 273.0f) * 0.25f);
                        bin = bin < 0 ? 0 : (bin > 25 ? 25 : bin);
                        queue_photon(pb, counters, i, gx, gy, 0x1Fu << bin, pb.temp[gslot], s.vx[i], s.vy[i]);
                    }
                }
            }
        } else if (rty == PT_FILT) {
            pb.ctype[i] = interact_wavelengths(s.tmp[r], pb.ctype[r], pb.temp[r], pb.ctype[i]);
        } else if (rty == PT_C5) {
            const uint32_t rc = pb.ctype[r], ic = pb.ctype[i];
            if (pb.life[r] > 0 && (rc & ic & 0xFFFFFFC0u) != 0u) {
                const int32_t rt = s.tmp[r];
                const float cvx = (float)((int32_t)((uint32_t)rt << 16) >> 16) / 255.0f;
                const float cvy = (float)(rt >> 16) / 255.0f;
                const float ovx = s.vx[i], ovy = s.vy[i];
                if (ovx + cvx == 0.0f && ovy + cvy == 0.0f) {
                    s.vx[i] = cvy;
                    s.vy[i] = -cvx;
                } else {
                    pb.ctype[i] = (rc & ic) >> 6;
                    const float nvx = ovx + cvx, nvy = ovy + cvy;
                    const float speed = nvx * nvx + nvy * nvy;
                    const float scale = sqrtf((ovx * ovx + ovy * ovy) / speed);
                    s.vx[i] = nvx * scale;
                    s.vy[i] = nvy * scale;
                }
                pb.life[r] = 0;
                pb.ctype[r] = 0u;
            } else if (rc == 0u && (ic & 0xFFFFFFC0u) != 0u) {
                const float vx = s.vx[i], vy = s.vy[i];
                pb.life[r] = 1;
                pb.ctype[r] = ic;
                s.tmp[r] = (int32_t)((0xFFFFu & (uint32_t)tpt_f2i(vx * 255.0f)) |
                                     (0xFFFF0000u & (uint32_t)tpt_f2i(vy * 16711680.0f)));
                pb.tmp2[r] = (int32_t)((0xFFFFu & (uint32_t)tpt_f2i((xf - (float)xi) * 255.0f)) |
                                       (0xFFFF0000u & (uint32_t)tpt_f2i((yf - (float)yi) * 16711680.0f)));
                s.ty[i] = 0u;
            }
        } else if (rty == PT_INVIS) {
            float resistance = 4.0f;
            if (s.tmp[r] > 0) resistance = (float)s.tmp[r];
            const uint32_t px = dest % (uint32_t)xres, py = dest / (uint32_t)xres;
            const float pv = pb.air_pv[(py / 4u) * (uint32_t)xc + px / 4u];
            if (pv >= -resistance && pv <= resistance) {
                s.ty[i] = PT_NEUT;
                pb.ctype[i] = 0u;
            }
        } else if (rty == PT_BIZR || rty == PT_BIZRG || rty == PT_BIZRS) {
            s.ty[i] = PT_ELEC;
            pb.ctype[i] = 0u;
        } else if (rty == PT_H2) {
            if ((s.tmp[i] & 1) == 0) {
                s.ty[i] = PT_PROT;
                pb.ctype[i] = 0u;
                pb.tmp2[i] = 1;
                final_ty = PT_ELEC;
                s.x[r] = (float)xi;
                s.y[r] = (float)yi;
                s.intent[r] = NO_INTENT;
                // create_part(ID(r), x, y, PT_ELEC): DefaultProperties, then
                // ELEC's Create hook (life 680, speed 2 at a hashed angle).
                const uint32_t *d = pb.defaults + PT_ELEC * TPT_SPAWN_DEFAULT_WORDS;
                pb.temp[r] = __uint_as_float(d[0]);
                pb.life[r] = 680;
                pb.ctype[r] = d[2];
                s.tmp[r] = (int32_t)d[3];
                pb.tmp2[r] = (int32_t)d[4];
                const uint32_t roll = fmix32(frame_seed ^ ((r + 1u) << 7) ^ (PT_ELEC << 3));
                const float a = (float)(roll % 360u) / 360.0f * 6.28318530717958647692f;
                // Double-precision cos/sin rounded to float: the same value
                // on every IEEE platform for these 360 angles.
                s.vx[r] = 2.0f * (float)cos((double)a);
                s.vy[r] = 2.0f * (float)sin((double)a);
            }
        } else if (rty == PT_GPMP && pb.life[r] == 0) {
            s.ty[i] = PT_GRVT;
            s.tmp[i] = tpt_f2i(pb.temp[r] - 273.15f);
        }
    } else if (ty == PT_NEUT) {
        if ((rty == PT_GLAS || rty == PT_BGLA) && seed % 10u == 0u) {
            const uint32_t g = pmap[own];
            const uint32_t gty = g & 0x1FFu, gslot = (g >> 9) - 1u;
            if (g >= 512u && (gty == PT_GLAS || gty == PT_BGLA) && gslot < slots) {
                const float vx = s.vx[i], vy = s.vy[i];
                if (vx * vx + vy * vy >= 1.44f * 1.44f) {
                    // cherenkov_photon (reaction_kernel/neut/mod.rs:335-369).
                    const uint32_t roll = fmix32(seed ^ 0x9E3779B9u);
                    const int lr = (int)(roll % 2u) * 2 - 1;
                    const float rvx = vx - (float)lr * 2.5f * vy;
                    const float rvy = vy + (float)lr * 2.5f * vx;
                    const float speed = sqrtf(rvx * rvx + rvy * rvy);
                    const float scale = 1.269f / speed;
                    queue_photon(pb, counters, i, xi, yi, 0x00000F80u, pb.temp[gslot], rvx * scale, rvy * scale);
                }
            }
        }
    } else if (ty == PT_ELEC) {
        if (rty == PT_GLOW) {
            s.ty[i] = PT_PHOT;
            pb.ctype[i] = 0x3FFFFFFFu;
        }
    } else if (ty == PT_PROT) {
        if (rty == PT_INVIS) s.ty[i] = PT_NEUT;
    } else if (rty == PT_FILT) {  // BIZR / BIZRG
        pb.ctype[i] = interact_wavelengths(s.tmp[r], pb.ctype[r], pb.temp[r], pb.ctype[i]);
    }
    if (final_ty != rty) s.ty[r] = final_ty;
}

// The deferred WOOD -> SAWD of rejected probes, then the pair pass over the
// movers the intent kernel listed.  The two lists touch disjoint state: a
// sawed WOOD has no pair arm, so whether a pair test sees it as WOOD or
// SAWD changes nothing.
__global__ __launch_bounds__(BLK) void mv_saw_pair(Slots s, PairBufs pb, const uint32_t *__restrict__ pmap,
                                                   const uint32_t *__restrict__ winner,
                                                   const uint32_t *__restrict__ saw_list,
                                                   const uint32_t *__restrict__ pair_list,
                                                   uint32_t *__restrict__ counters, uint32_t slots, int xres, int yres,
                                                   int xc, const uint32_t *__restrict__ seedp) {
    const uint32_t n_saw = counters[TPT_CTR_SAW], n = counters[TPT_CTR_PAIR];
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    for (uint32_t k = t0; k < n_saw; k += stride) s.ty[saw_list[k]] = 180u;
    if (n == 0u) return;
    const uint32_t frame_seed = *seedp;
    for (uint32_t k = t0; k < n; k += stride)
        pair_one(s, pb, pmap, winner, counters, pair_list[k], slots, xres, yres, xc, frame_seed);
}

struct DrainBufs {
    int32_t *life, *tmp2, *tmp3, *tmp4;
    uint32_t *ctype, *dcolour;
    float *temp;
    const uint32_t *defaults;
    const uint32_t *rec;
    const uint32_t *spawn_free;
};

constexpr int DRAIN_BLK = 1024;
constexpr uint32_t DRAIN_CHUNK = 4096u;

// drain_pending_creates for the pair pass's photons (create_alloc_kernel +
// create_fields_kernel, reaction_kernel/shared.rs:516-912, with the record
// mask CR_NO_OCCUPANCY | CR_SET_CTYPE | CR_SET_TEMP | CR_SET_VX | CR_SET_VY
// both creators use): every record takes a slot from the window without an
// occupancy test; life, tmp..tmp4 come from PHOT's DefaultProperties,
// dcolour and flags are cleared.  The map words the reference writes here
// are rebuilt by the scatter that follows, and the active-element mask it
// updates has no reader in this library.  One workgroup: a record's
// allocation index is its rank by requester slot (slots are unique).
__device__ void drain_creates(Slots &s, const DrainBufs &d, const uint32_t *__restrict__ counters,
                              uint32_t rec_cap, int xres, int yres) {
    __shared__ uint32_t keys[DRAIN_CHUNK];
    __shared__ uint32_t prefix[TPT_SPAWN_BITMAP_WORDS + 1];
    const uint32_t total = counters[TPT_CTR_CREATE];
    const uint32_t n = total < rec_cap ? total : rec_cap;
    if (n == 0u) return;
    const uint32_t tid = threadIdx.x;
    // Exclusive prefix popcount of the free bitmap (sequential over 512
    // words by one thread: negligible next to the rank loop).
    if (tid == 0u) {
        uint32_t acc = 0u;
        for (uint32_t w = 0; w < TPT_SPAWN_BITMAP_WORDS; ++w) {
            prefix[w] = acc;
            acc += __popc(d.spawn_free[w]);
        }
        prefix[TPT_SPAWN_BITMAP_WORDS] = acc;
    }
    __syncthreads();
    const uint32_t avail = prefix[TPT_SPAWN_BITMAP_WORDS];
    for (uint32_t base = 0; base < n; base += DRAIN_BLK) {
        const uint32_t k = base + tid;
        const uint32_t mine = k < n ? d.rec[(size_t)k * TPT_CREATE_REC_WORDS] : 0xFFFFFFFFu;
        uint32_t rank = 0u;
        for (uint32_t c0 = 0; c0 < n; c0 += DRAIN_CHUNK) {
            const uint32_t len = n - c0 < DRAIN_CHUNK ? n - c0 : DRAIN_CHUNK;
            __syncthreads();
            for (uint32_t j = tid; j < len; j += DRAIN_BLK) keys[j] = d.rec[(size_t)(c0 + j) * TPT_CREATE_REC_WORDS];
            __syncthreads();
            for (uint32_t j = 0; j < len; ++j) rank += keys[j] < mine ? 1u : 0u;
        }
        if (k >= n || rank >= TPT_GPU_CREATE_COUNT || rank >= avail) continue;
        const uint32_t *r = d.rec + (size_t)k * TPT_CREATE_REC_WORDS;
        const int x = (int)(r[1] & 0xFFFFu), y = (int)(r[1] >> 16);
        if (x >= xres || y >= yres) continue;
        // The rank-th set bit of the bitmap, in ascending slot order.
        uint32_t lo = 0u, hi = TPT_SPAWN_BITMAP_WORDS - 1u;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1u) >> 1;
            if (prefix[mid] <= rank) lo = mid;
            else hi = mid - 1u;
        }
        uint32_t word = d.spawn_free[lo], skip = rank - prefix[lo];
        for (; skip > 0u; --skip) word &= word - 1u;
        const uint32_t slot = TPT_GPU_SPAWN_START + lo * 32u + (uint32_t)__ffs(word) - 1u;
        const uint32_t *def = d.defaults + PT_PHOT * TPT_SPAWN_DEFAULT_WORDS;
        s.ty[slot] = PT_PHOT;
        s.x[slot] = (float)x;
        s.y[slot] = (float)y;
        s.flags[slot] = 0u;
        d.life[slot] = (int32_t)def[1];
        d.ctype[slot] = r[2];
        d.temp[slot] = __uint_as_float(r[3]);
        s.tmp[slot] = (int32_t)def[3];
        d.tmp2[slot] = (int32_t)def[4];
        d.tmp3[slot] = (int32_t)def[5];
        d.tmp4[slot] = (int32_t)def[6];
        d.dcolour[slot] = 0u;
        s.vx[slot] = __uint_as_float(r[4]);
        s.vy[slot] = __uint_as_float(r[5]);
    }
}

struct CommitBufs {
    uint32_t *ctype;
    int32_t *life;
    float *temp;
    const uint32_t *photon_mask;
    const float *air_pv;
};

__device__ __forceinline__ bool void_admits(uint32_t ty, uint32_t vctype, int32_t vtmp) {
    if (vctype == 0u) return true;
    const bool matches = vctype == ty, inv = (vtmp & 1) != 0;
    return inv ? !matches : matches;
}

__device__ __forceinline__ uint32_t pmap_at(const uint32_t *pmap, int x, int y, int xr, int yr) {
    return (x < 0 || y < 0 || x >= xr || y >= yr) ? 0u : pmap[y * xr + x];
}

// commit_move_kernel (commit.rs:181-535).
__global__ __launch_bounds__(BLK) void mv_commit(Slots s, CommitBufs cb, const uint32_t *__restrict__ pmap,
                                                 const uint4 *__restrict__ cellrec, const uint32_t *__restrict__ lut,
                                                 const uint32_t *__restrict__ winner, uint32_t *__restrict__ photon_list,
                                                 uint32_t *__restrict__ counters, uint32_t slots, int xres, int yres,
                                                 int xc, uint32_t heat_enable) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    // Slot fields loaded together (only this thread writes flags[i]).
    const uint32_t ty = s.ty[i], raw = s.intent[i], flags = s.flags[i];
    if (ty == 0u || ty >= 256u) return;
    if (raw == NO_INTENT) {
        if (lut[ty * 12u + 1u] >= 2u && (flags & 0x11u) == 0u) s.flags[i] = flags | 0x01u;
        return;
    }
    const bool half_mirror = (raw & HALF_MIRROR) != 0u;
    const uint32_t target = raw & PIXEL_MASK;
    if (raw & 0x80000000u) {
        // A blocked photon's hit: deferred to mv_after_commit (canonical R10).
        if (ty == 29u) photon_list[atomicAdd(&counters[TPT_CTR_PHOT], 1u)] = i;
        return;
    }
    if (target >= (uint32_t)(xres * yres)) return;
    if (flags & 0x01u) s.flags[i] = flags & 0xFFFFFFFEu;
    // The claim, the destination's occupant and its wall in one round trip.
    const int nx = (int)(target % (uint32_t)xres), ny = (int)(target / (uint32_t)xres);
    const int wc = (ny / TPT_CELL) * xc + nx / TPT_CELL;
    const uint32_t claim = winner[target], dest = pmap[target], wall = cellrec[wc].z;
    if ((claim & 0x7FFFFFu) != i + 1u) return;
    if (half_mirror) {
        s.x[i] = (float)nx;
        s.y[i] = (float)ny;
        return;
    }
    const uint32_t dty = dest & 0x1FFu, docc = dest >> 9;
    const uint32_t mprops = lut[ty * 12u], dprops = lut[dty * 12u];
    uint32_t can = 0u, other = i;
    if (docc != 0u) {
        other = docc - 1u;
        can = lut[LUT_CAN + ty * 256u + dty];
        if (can == 3u) {  // resolve_can (commit.rs:76-128)
            if (dty == 58u && ty == 29u) can = cb.life[other] > 5 ? 2u : 0u;
            else if (dty == 43u && ty == 29u) can = cb.life[other] < 10 ? 2u : 0u;
            else if (dty == 124u) {
                const float res = s.tmp[other] > 0 ? (float)s.tmp[other] : 4.0f;
                const float pv = cb.air_pv[wc];
                can = (pv < -res || pv > res) ? 2u : 0u;
            } else if (dty == 75u)
                can = (cb.life[other] == 10 && void_admits(ty, cb.ctype[other], s.tmp[other])) ? 1u : 0u;
            else if (dty == 23u)
                can = void_admits(ty, cb.ctype[other], s.tmp[other]) ? 1u : 0u;
            else if (dty == 60u)
                can = ty == 151u ? (cb.life[other] >= 10 ? 2u : 0u) : 1u;
            else
                can = 1u;
        }
    }
    bool ehole_exit = false;
    if (wall == 12u && (cellrec[wc].w & 0xFFu) == 0u) {
        const int cx = tpt_f2i(s.x[i] + 0.5f) / TPT_CELL, cy = tpt_f2i(s.y[i] + 0.5f) / TPT_CELL;
        const int here = cy * xc + cx;
        ehole_exit = cellrec[here].z == 12u && (cellrec[here].w & 0xFFu) == 0u;
    }
    if (ehole_exit) return;
    if (docc != 0u && can == 0u) return;
    if (docc != 0u && can == 1u && other != i) {
        if (dty == 23u || dty == 75u) {
            s.ty[i] = 0u;
        } else if (dty == 91u || dty == 40u) {
            cb.temp[other] = tpt_clampf(cb.temp[other] + cb.temp[i] / 2.0f, 0.0f, 9999.0f);
            s.ty[i] = 0u;
        } else if (dty == 92u || dty == 41u) {
            if (ty == 122u) {
                cb.temp[other] = tpt_clampf(cb.temp[other] - (9999.0f - cb.temp[i]) / 2.0f, 0.0f, 9999.0f);
                s.ty[i] = 0u;
            }
        } else if (dty == 111u) {
            if (ty == 31u) {
                if (cb.life[other] < 6000) cb.life[other] += 1;
                cb.temp[other] = 0.0f;
                s.ty[i] = 0u;
            }
        } else if (dty == 165u || dty == 166u) {
            if (mprops & 0x10u) {
                s.tmp[other] += 20;
                s.ty[i] = 0u;
            }
        } else if (ty == 30u && (dprops & 0x100u)) {
            s.ty[i] = 0u;
        } else if (ty == 30u) {
            const int own_x = tpt_f2i(s.x[i] + 0.5f), own_y = tpt_f2i(s.y[i] + 0.5f);
            const int own_cell = (own_y / TPT_CELL) * xc + own_x / TPT_CELL;
            const uint32_t own = pmap[own_y * xres + own_x];
            const bool drag = own != 0u && (lut[(own & 0x1FFu) * 12u] & 0x80u) != 0u;
            if (drag && cellrec[own_cell].z != 15u) {
                const uint32_t under = (own >> 9) - 1u;
                s.x[under] = (float)nx;
                s.y[under] = (float)ny;
                s.x[other] = (float)own_x;
                s.y[other] = (float)own_y;
            }
            s.x[i] = (float)nx;
            s.y[i] = (float)ny;
        } else if (ty == 42u && cb.life[i] > 0) {
        } else {
            const uint32_t oraw = s.intent[other], otgt = oraw & PIXEL_MASK;
            const bool other_moves = oraw != NO_INTENT && (oraw & 0x80000000u) == 0u &&
                                     otgt < (uint32_t)(xres * yres) && (winner[otgt] & 0x7FFFFFu) == other + 1u;
            if (!other_moves) {
                const float ox = s.x[other], oy = s.y[other];
                s.x[other] = s.x[i];
                s.y[other] = s.y[i];
                s.x[i] = ox;
                s.y[i] = oy;
            }
        }
        return;
    }
    s.x[i] = (float)nx;
    s.y[i] = (float)ny;
}

// The blocked photons' hits on the pixels they could not enter
// (commit.rs:263-320): wavelength filter, PSCN/NSCN sparking, temperature
// exchange with the hit particle.  The reference runs these read-modify-
// writes from concurrent threads; here (canonical R10) they follow the rest
// of the commit, and hits on one particle apply in ascending photon slot
// order: one workgroup runs rounds in which the lowest pending photon of
// each hit particle (atomicMin into `owner`, all ones at rest) applies its
// hit and leaves the list.
__device__ void photon_hit(Slots &s, CommitBufs &cb, const uint32_t *__restrict__ pmap,
                           const uint32_t *__restrict__ lut, uint32_t i, uint32_t target, int xres, int yres,
                           uint32_t heat_enable) {
    const uint32_t hit = target & 0x7FFFFFFFu, r = pmap[hit];
    if (r >= 512u) {
        const uint32_t rty = r & 0x1FFu, rslot = (r >> 9) - 1u;
        uint32_t mask = cb.photon_mask[rty];
        if (rty == 190u) {
            const int bin = min(max((int)(cb.ctype[rslot] / 4u), 0), 25);
            mask = 0x1Fu << (uint32_t)bin;
        } else if (rty == 194u) {
            const uint32_t colour = (cb.ctype[rslot] >> 6) & 0x3Fu;
            if ((colour & 48u) == 0u) mask |= 1u << 25;
            if ((colour & 12u) == 0u) mask |= 1u << 15;
            if ((colour & 3u) == 0u) mask |= 1u << 5;
        }
        cb.ctype[i] &= mask;
        if (rty == 50u && cb.life[rslot] == 0) {
            const int hx = (int)(hit % (uint32_t)xres), hy = (int)(hit / (uint32_t)xres);
            if ((pmap_at(pmap, hx - 1, hy, xres, yres) & 0x1FFu) == 51u ||
                (pmap_at(pmap, hx + 1, hy, xres, yres) & 0x1FFu) == 51u ||
                (pmap_at(pmap, hx, hy - 1, xres, yres) & 0x1FFu) == 51u ||
                (pmap_at(pmap, hx, hy + 1, xres, yres) & 0x1FFu) == 51u) {
                cb.ctype[rslot] = 50u;
                s.ty[rslot] = 12u;
                cb.life[rslot] = 4;
            }
        }
        if (heat_enable) {
            if (rty == 62u || rty == 27u) cb.temp[rslot] = cb.temp[i];
            const uint32_t hc = lut[rty * 12u + 3u];
            const bool ins = hc == 0u || (rty == 68u && cb.life[rslot] != 10) ||
                             ((rty == 82u || rty == 162u) && (s.tmp[rslot] & 1) == 0);
            if (!ins && rty != 134u) {
                const float avg = tpt_clampf((cb.temp[rslot] + cb.temp[i]) * 0.5f, 0.0f, 9999.0f);
                cb.temp[i] = avg;
                cb.temp[rslot] = avg;
            }
        }
    }
}

__device__ void photon_hits(Slots &s, CommitBufs &cb, const uint32_t *__restrict__ pmap,
                            const uint32_t *__restrict__ lut, uint32_t *__restrict__ list,
                            uint32_t *__restrict__ counters, uint32_t *__restrict__ owner, int xres, int yres,
                            uint32_t heat_enable) {
    __shared__ int pending;
    const uint32_t n = counters[TPT_CTR_PHOT];
    for (;;) {
        if (threadIdx.x == 0) pending = 0;
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
            const uint32_t i = list[k];
            if (i == NO_INTENT) continue;
            const uint32_t r = pmap[(s.intent[i] & PIXEL_MASK) & 0x7FFFFFFFu];
            if (r < 512u) {
                list[k] = NO_INTENT;  // nothing there to hit
                continue;
            }
            atomicMin(&owner[(r >> 9) - 1u], i);
            pending = 1;
        }
        __syncthreads();
        if (!pending) break;
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
            const uint32_t i = list[k];
            if (i == NO_INTENT) continue;
            const uint32_t target = s.intent[i] & PIXEL_MASK;
            const uint32_t rslot = (pmap[target & 0x7FFFFFFFu] >> 9) - 1u;
            if (owner[rslot] != i) continue;
            photon_hit(s, cb, pmap, lut, i, target, xres, yres, heat_enable);
            owner[rslot] = 0xFFFFFFFFu;
            list[k] = NO_INTENT;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) counters[TPT_CTR_PHOT] = 0u;
}

// One workgroup after the commit pass: the deferred blocked-photon hits,
// then the drain of the pair pass's creations.  The drain writes only
// window slots that were free before the pair pass; the hits touch only
// particles that were live then.
__global__ __launch_bounds__(DRAIN_BLK) void mv_after_commit(Slots s, CommitBufs cb, DrainBufs d,
                                                             const uint32_t *__restrict__ pmap,
                                                             const uint32_t *__restrict__ lut,
                                                             uint32_t *__restrict__ list,
                                                             uint32_t *__restrict__ counters,
                                                             uint32_t *__restrict__ owner, uint32_t rec_cap, int xres,
                                                             int yres, uint32_t heat_enable) {
    photon_hits(s, cb, pmap, lut, list, counters, owner, xres, yres, heat_enable);
    __syncthreads();
    drain_creates(s, d, counters, rec_cap, xres, yres);
}

}  // namespace

int tpt_movement_step(tpt_ctx *ctx) {
    const uint32_t *seed = ctx->counters + TPT_CTR_SEED;
    ctx->maps_fresh = ctx->rebuild_fresh = false;
    const tpt_config &c = ctx->cfg;
    // The pair pass arms the creation queue, after which the slot bound
    // covers the device allocation window (gpu_spawn_used,
    // move_kernel/mod.rs:1212-1216).  The added slots are empty until the
    // drain fills them.
    const uint32_t spawn_end = c.capacity < TPT_GPU_SPAWN_END ? c.capacity : TPT_GPU_SPAWN_END;
    if (ctx->slot_bound < spawn_end) ctx->slot_bound = spawn_end;
    const uint32_t slots = ctx->slot_bound, P = ctx->pixels;
    const int xres = (int)c.xres, yres = (int)c.yres, xc = (int)ctx->xcells;
    Slots s{tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<float>(ctx, TPT_BUF_PART_X),
            tpt_buf<float>(ctx, TPT_BUF_PART_Y),   tpt_buf<float>(ctx, TPT_BUF_PART_VX),
            tpt_buf<float>(ctx, TPT_BUF_PART_VY),  tpt_buf<uint32_t>(ctx, TPT_BUF_PART_FLAGS),
            tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP), tpt_buf<uint32_t>(ctx, TPT_BUF_INTENT_TARGET)};
    const uint32_t *wall = tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP), *emap = tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP);
    const uint32_t *lut = tpt_buf<uint32_t>(ctx, TPT_BUF_MOVE_LUT);
    uint32_t *pmap = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP);
    uint32_t *winner = tpt_buf<uint32_t>(ctx, TPT_BUF_CELL_WINNER);
    float *air_vx = tpt_buf<float>(ctx, TPT_BUF_AIR_VX), *air_vy = tpt_buf<float>(ctx, TPT_BUF_AIR_VY);
    const dim3 sg((slots + BLK - 1) / BLK), cg((ctx->cells + BLK - 1) / BLK), b(BLK);
    const uint32_t n4 = P / 4;
    const dim3 pg((n4 + BLK - 1) / BLK);

    tpt_pass_begin(ctx, PASS_AIR_DEPOSIT);
    if (ctx->rec_cover) {
        hipLaunchKernelGGL(mv_air_deposit_list, dim3(256), b, 0, ctx->stream, s, wall, emap, ctx->cell_scratch_u,
                           ctx->offmap, ctx->counters, slots, xres, yres, xc);
        hipLaunchKernelGGL(mv_air_deposit_px, cg, b, 0, ctx->stream, s, (const ulonglong2 *)ctx->pix_rec, wall, emap,
                           lut, air_vx, air_vy, ctx->cell_scratch_u, slots, xres, xc, (int)ctx->ycells);
    } else {
        hipLaunchKernelGGL(mv_air_deposit, sg, b, 0, ctx->stream, s, wall, emap, ctx->cell_scratch_u, slots, xres,
                           yres, xc);
        hipLaunchKernelGGL(mv_air_deposit_cells, cg, b, 0, ctx->stream, s, lut, air_vx, air_vy, ctx->cell_scratch_u,
                           ctx->cells);
    }
    ctx->rec_cover = false;
    tpt_pass_end(ctx, PASS_AIR_DEPOSIT);
    // WL_DETECT charging + closure + marker strip, on canvases with wire
    // walls (the host sees every wall edit) or with end markers uploaded.
    if (ctx->has_wire || ctx->emap_markers) {
        uint32_t *emap_w = tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP);
        tpt_pass_begin(ctx, PASS_MOVE_DETECT);
        if (ctx->has_wire) {
            hipLaunchKernelGGL(mv_detect_span, sg, b, 0, ctx->stream, s, wall, emap_w, ctx->cell_scratch_u2, slots,
                               xres, yres, xc);
            hipLaunchKernelGGL(mv_detect_fill, cg, b, 0, ctx->stream, emap_w, ctx->cell_scratch_u2, xc, ctx->cells);
            hipLaunchKernelGGL(mv_emap_gate, cg, b, 0, ctx->stream, wall, emap_w, ctx->counters + TPT_CTR_EMAP_GATE,
                               ctx->cells);
            hipLaunchKernelGGL(mv_emap_closure, dim3(1), dim3(EMAP_THREADS), 0, ctx->stream, wall, emap_w,
                               ctx->counters + TPT_CTR_EMAP_GATE, xc, (int)ctx->ycells);
        }
        hipLaunchKernelGGL(mv_emap_strip, cg, b, 0, ctx->stream, emap_w, ctx->cells);
        ctx->emap_markers = false;
        tpt_pass_end(ctx, PASS_MOVE_DETECT);
    }
    uint32_t *spawn_free = ctx->create_rec + (size_t)c.capacity * TPT_CREATE_REC_WORDS;
    tpt_pass_begin(ctx, PASS_MOVE_CLEAR);
    const uint32_t n_prep = n4 > TPT_SPAWN_BITMAP_WORDS ? n4 : TPT_SPAWN_BITMAP_WORDS;
    hipLaunchKernelGGL(mv_prepare, dim3((n_prep + BLK - 1) / BLK), b, 0, ctx->stream, (uint4 *)winner,
                       (const uint4 *)pmap, (uint32_t *)ctx->pix_type, n4, (const uint32_t *)air_vx,
                       (const uint32_t *)air_vy, wall, emap, ctx->cell_rec, ctx->cells, spawn_free, ctx->counters);
    tpt_pass_end(ctx, PASS_MOVE_CLEAR);
    MoveEnv e{pmap, ctx->pix_type, ctx->cell_rec, lut, xres, yres, xc, c.edge_mode, c.gravity_mode,
              c.custom_gravity_x, c.custom_gravity_y, c.newton_grav, ctx->slot_scratch_u2, ctx->counters + TPT_CTR_SAW,
              tpt_buf<uint32_t>(ctx, TPT_BUF_PART_CTYPE), ctx->slot_scratch_u, ctx->counters};
    tpt_pass_begin(ctx, PASS_MOVE_INTENT);
    hipLaunchKernelGGL(mv_intent, sg, b, 0, ctx->stream, s, e, winner, slots, seed, c.capacity, spawn_free);
    tpt_pass_end(ctx, PASS_MOVE_INTENT);
    const uint32_t *defaults = tpt_buf<uint32_t>(ctx, TPT_BUF_SPAWN_DEFAULTS);
    PairBufs pb{tpt_buf<uint32_t>(ctx, TPT_BUF_PART_CTYPE), tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE),
                tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP2), tpt_buf<float>(ctx, TPT_BUF_PART_TEMP),
                tpt_buf<float>(ctx, TPT_BUF_AIR_PV), defaults, ctx->create_rec, c.capacity};
    tpt_pass_begin(ctx, PASS_MOVE_SAW);
    hipLaunchKernelGGL(mv_saw_pair, dim3(64), b, 0, ctx->stream, s, pb, pmap, winner, ctx->slot_scratch_u2,
                       ctx->slot_scratch_u, ctx->counters, slots, xres, yres, xc, seed);
    tpt_pass_end(ctx, PASS_MOVE_SAW);
    CommitBufs cb{tpt_buf<uint32_t>(ctx, TPT_BUF_PART_CTYPE), tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE),
                  tpt_buf<float>(ctx, TPT_BUF_PART_TEMP), tpt_buf<uint32_t>(ctx, TPT_BUF_PHOTON_MASK),
                  tpt_buf<float>(ctx, TPT_BUF_AIR_PV)};
    tpt_pass_begin(ctx, PASS_MOVE_COMMIT);
    hipLaunchKernelGGL(mv_commit, sg, b, 0, ctx->stream, s, cb, pmap, ctx->cell_rec, lut, winner, ctx->slot_scratch_u2,
                       ctx->counters, slots, xres, yres, xc, c.heat_enable ? 1u : 0u);
    tpt_pass_end(ctx, PASS_MOVE_COMMIT);
    DrainBufs db{tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE),    tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP2),
                 tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP3),    tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP4),
                 tpt_buf<uint32_t>(ctx, TPT_BUF_PART_CTYPE),  tpt_buf<uint32_t>(ctx, TPT_BUF_PART_DCOLOUR),
                 tpt_buf<float>(ctx, TPT_BUF_PART_TEMP),      defaults,
                 ctx->create_rec,                              spawn_free};
    tpt_pass_begin(ctx, PASS_MOVE_DRAIN);
    hipLaunchKernelGGL(mv_after_commit, dim3(1), dim3(DRAIN_BLK), 0, ctx->stream, s, cb, db, pmap, lut,
                       ctx->slot_scratch_u2, ctx->counters, ctx->slot_owner, c.capacity, xres, yres,
                       c.heat_enable ? 1u : 0u);
    tpt_pass_end(ctx, PASS_MOVE_DRAIN);
    int rc = tpt_check_launch("movement pass");
    if (rc) return rc;
    // clear_spatial_maps_kernel + scatter_spatial_maps_kernel (commit.rs:541-603),
    // built as the pixel records RecalcFreeParticles uses.
    return tpt_move_rebuild_maps(ctx);
}
