/*
 * Oracle: Eulerian air step -- TEST INFRASTRUCTURE ONLY (see tpt_oracle.h).
 *
 * Restates crates/tpt-sim/src/air_kernel/{mod,pressure,advect,heat,settle}.rs
 * kernel by kernel, in launch order (air_kernel/mod.rs:150-291).  Each kernel
 * reads the buffers the reference binds and writes its outputs only after
 * every cell has been computed, which is what a data-parallel launch does when
 * no thread reads what another writes.
 *
 * Canonical race rule R4: the reference heat kernel writes vx/vy in place
 * while neighbouring threads read them for smoothing (heat.rs:93-95 vs
 * 249-250).  Here every read sees the pre-kernel vx/vy (Jacobi).
 */
#include "tpt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_util.h"

#define BORDER 2u
#define ADVECT_DIST 0.7f

static uint32_t in_border(uint32_t x, uint32_t y, uint32_t xc, uint32_t yc) {
    return (x < BORDER || y < BORDER || x + BORDER >= xc || y + BORDER >= yc) ? 1u : 0u;
}

static uint32_t interior(uint32_t x, uint32_t y, uint32_t xc, uint32_t yc) {
    return (x >= 1 && y >= 1 && x + 1 < xc && y + 1 < yc) ? 1u : 0u;
}

/* pressure.rs:29-36 */
static uint32_t blocked_at(const uint32_t *b, int x, int y, uint32_t xc, uint32_t yc) {
    if (x >= 0 && y >= 0 && x < (int)xc && y < (int)yc) return b[(uint32_t)y * xc + (uint32_t)x];
    return 0;
}

/* pressure.rs:42-62 */
static uint32_t wall_cleared_vx(const uint32_t *b, uint32_t x, uint32_t y, uint32_t xc, uint32_t yc) {
    uint32_t out = 0;
    if (y >= 1 && y + 1 < yc) {
        if (x >= 1 && x + 1 < xc && blocked_at(b, (int)x, (int)y, xc, yc) != 0) out = 1;
        if (x >= 2 && blocked_at(b, (int)x - 1, (int)y, xc, yc) != 0) out = 1;
        if (x + 2 < xc && blocked_at(b, (int)x + 1, (int)y, xc, yc) != 0) out = 1;
    }
    return out;
}

/* pressure.rs:65-85 */
static uint32_t wall_cleared_vy(const uint32_t *b, uint32_t x, uint32_t y, uint32_t xc, uint32_t yc) {
    uint32_t out = 0;
    if (x >= 1 && x + 1 < xc) {
        if (y >= 1 && y + 1 < yc && blocked_at(b, (int)x, (int)y, xc, yc) != 0) out = 1;
        if (y >= 2 && blocked_at(b, (int)x, (int)y - 1, xc, yc) != 0) out = 1;
        if (y + 2 < yc && blocked_at(b, (int)x, (int)y + 1, xc, yc) != 0) out = 1;
    }
    return out;
}

/* pressure.rs:91-114 */
static uint32_t adjacent_blocked_x(const uint32_t *b, uint32_t x, uint32_t y, uint32_t xc, uint32_t yc) {
    return (blocked_at(b, (int)x - 1, (int)y, xc, yc) || blocked_at(b, (int)x, (int)y, xc, yc) ||
            blocked_at(b, (int)x + 1, (int)y, xc, yc)) ? 1u : 0u;
}
static uint32_t adjacent_blocked_y(const uint32_t *b, uint32_t x, uint32_t y, uint32_t xc, uint32_t yc) {
    return (blocked_at(b, (int)x, (int)y - 1, xc, yc) || blocked_at(b, (int)x, (int)y, xc, yc) ||
            blocked_at(b, (int)x, (int)y + 1, xc, yc)) ? 1u : 0u;
}

/* pressure.rs:118-157 */
static float effective_vx(const float *vx, const uint32_t *b, uint32_t x, uint32_t y, uint32_t xc,
                          uint32_t yc, float evx) {
    float v = vx[y * xc + x];
    if (in_border(x, y, xc, yc)) v = evx + (v - evx) * 0.9f;
    if (wall_cleared_vx(b, x, y, xc, yc)) v = 0.0f;
    return v;
}
static float effective_vy(const float *vy, const uint32_t *b, uint32_t x, uint32_t y, uint32_t xc,
                          uint32_t yc, float evy) {
    float v = vy[y * xc + x];
    if (in_border(x, y, xc, yc)) v = evy + (v - evy) * 0.9f;
    if (wall_cleared_vy(b, x, y, xc, yc)) v = 0.0f;
    return v;
}

void oracle_air_pressure(oracle_state *s) {
    const uint32_t xc = s->xres / 4, yc = s->yres / 4;
    const float ep = s->edge_pressure;
#pragma omp parallel for schedule(static)
    for (uint32_t y = 0; y < yc; ++y)
        for (uint32_t x = 0; x < xc; ++x) {
            uint32_t idx = y * xc + x;
            float cur = s->air_pv[idx];
            if (in_border(x, y, xc, yc)) cur = ep + (cur - ep) * 0.8f;
            float next = cur;
            if (interior(x, y, xc, yc)) {
                float dp = effective_vx(s->air_vx, s->block_air, x - 1, y, xc, yc, s->edge_velocity_x) -
                           effective_vx(s->air_vx, s->block_air, x + 1, y, xc, yc, s->edge_velocity_x) +
                           effective_vy(s->air_vy, s->block_air, x, y - 1, xc, yc, s->edge_velocity_y) -
                           effective_vy(s->air_vy, s->block_air, x, y + 1, xc, yc, s->edge_velocity_y);
                next = ep + (cur - ep) * 0.9999f + dp * 0.15f;
            }
            s->air_opv[idx] = next;
        }
}

void oracle_air_velocity(oracle_state *s) {
    const uint32_t xc = s->xres / 4, yc = s->yres / 4;
    const float evx = s->edge_velocity_x, evy = s->edge_velocity_y;
    const float *opv = s->air_opv;
#pragma omp parallel for schedule(static)
    for (uint32_t y = 0; y < yc; ++y)
        for (uint32_t x = 0; x < xc; ++x) {
            uint32_t idx = y * xc + x;
            float cvx = s->air_vx[idx], cvy = s->air_vy[idx];
            if (in_border(x, y, xc, yc)) {
                cvx = evx + (cvx - evx) * 0.9f;
                cvy = evy + (cvy - evy) * 0.9f;
            }
            if (wall_cleared_vx(s->block_air, x, y, xc, yc)) cvx = 0.0f;
            if (wall_cleared_vy(s->block_air, x, y, xc, yc)) cvy = 0.0f;
            float nvx = cvx, nvy = cvy;
            if (interior(x, y, xc, yc)) {
                float dx = opv[y * xc + x - 1] - opv[y * xc + x + 1];
                float dy = opv[(y - 1) * xc + x] - opv[(y + 1) * xc + x];
                nvx = evx + (cvx - evx) * 0.999f + dx * 0.2f;
                nvy = evy + (cvy - evy) * 0.999f + dy * 0.2f;
                if (adjacent_blocked_x(s->block_air, x, y, xc, yc)) nvx = 0.0f;
                if (adjacent_blocked_y(s->block_air, x, y, xc, yc)) nvy = 0.0f;
            }
            s->air_ovx[idx] = nvx;
            s->air_ovy[idx] = nvy;
        }
}

/* advect.rs:10-42 */
static float vorticity_bmap(const float *vx, const float *vy, const uint32_t *b, uint32_t xc, uint32_t yc,
                            uint32_t x, uint32_t y) {
    float out = 0.0f;
    if (x > 1 && x + 2 < xc && y > 1 && y + 2 < yc) {
        uint32_t row = y * xc, rm = (y - 1) * xc, rp = (y + 1) * xc;
        int here = b[row + x] != 0;
        float dvydx = vy[row + x + 1] - vy[row + x - 1];
        if (here || b[row + x + 1] != 0 || b[row + x - 1] != 0) dvydx = 0.0f;
        float dvxdy = vx[rp + x] - vx[rm + x];
        if (here || b[rp + x] != 0 || b[rm + x] != 0) dvxdy = 0.0f;
        out = (dvydx - dvxdy) * 0.5f;
    }
    return out;
}

/* advect.rs:46-52 */
static uint32_t advect_blocked(const uint32_t *b, uint32_t idx, uint32_t heat) {
    return heat ? (b[idx] & 0x8u) : b[idx];
}

void oracle_far_advect_target(uint32_t x, uint32_t y, float dx, float dy, const uint32_t *block,
                              uint32_t xc, uint32_t yc, uint32_t heat, float *otx, float *oty) {
    float tx = (float)x - dx * ADVECT_DIST;
    float ty = (float)y - dy * ADVECT_DIST;
    float adx = dx * ADVECT_DIST;
    if (adx < 0.0f) adx = 0.0f - adx;
    float ady = dy * ADVECT_DIST;
    if (ady < 0.0f) ady = 0.0f - ady;
    float limit_x = (float)xc - 2.0f, limit_y = (float)yc - 2.0f;
    if ((adx > 1.0f || ady > 1.0f) && tx >= 2.0f && tx < limit_x && ty >= 2.0f && ty < limit_y) {
        float abs_dx = dx < 0.0f ? 0.0f - dx : dx;
        float abs_dy = dy < 0.0f ? 0.0f - dy : dy;
        float step_x, step_y;
        int step_limit;
        if (abs_dx > abs_dy) {
            step_x = dx < 0.0f ? 1.0f : -1.0f;
            step_y = 0.0f - dy / abs_dx;
            step_limit = ora_f2i(adx);
        } else {
            step_x = 0.0f - dx / abs_dy;
            step_y = dy < 0.0f ? 1.0f : -1.0f;
            step_limit = ora_f2i(ady);
        }
        tx = (float)x;
        ty = (float)y;
        int step = 0, stopped = 0;
        while (step < step_limit) {
            tx += step_x;
            ty += step_y;
            int cx = ora_f2i(tx + 0.5f), cy = ora_f2i(ty + 0.5f);
            int hit = 1;
            if (cx >= 0 && cy >= 0 && cx < (int)xc && cy < (int)yc) {
                hit = advect_blocked(block, (uint32_t)cy * xc + (uint32_t)cx, heat) != 0;
            }
            if (hit) {
                tx -= step_x;
                ty -= step_y;
                stopped = 1;
                step = step_limit;
            } else {
                step += 1;
            }
        }
        if (!stopped) {
            tx = (float)x - dx * ADVECT_DIST;
            ty = (float)y - dy * ADVECT_DIST;
        }
    }
    *otx = tx;
    *oty = ty;
}

static float gauss_w(uint32_t dx, uint32_t dy) {
    float w = 0.01831564f;
    if (dx == 1) w = 0.13533528f;
    if (dy == 1) w = 0.13533528f;
    if (dx == 1 && dy == 1) w = 1.0f;
    return w;
}

void oracle_air_smooth_advect(oracle_state *s) {
    const uint32_t xc = s->xres / 4, yc = s->yres / 4, n = xc * yc;
    const float *ovx = s->air_ovx, *ovy = s->air_ovy, *opv = s->air_opv;
    const uint32_t *ba = s->block_air;
#pragma omp parallel for schedule(static)
    for (uint32_t y = 0; y < yc; ++y)
        for (uint32_t x = 0; x < xc; ++x) {
            uint32_t idx = y * xc + x;
            float cvx = ovx[idx], cvy = ovy[idx], cpv = opv[idx];
            float sx_ = 0.0f, sy_ = 0.0f, sp_ = 0.0f, ws = 0.0f;
            for (uint32_t dy = 0; dy < 3; ++dy)
                for (uint32_t dx = 0; dx < 3; ++dx) {
                    uint32_t nx = x + dx, ny = y + dy;
                    nx = nx >= 1 ? nx - 1 : 0;
                    ny = ny >= 1 ? ny - 1 : 0;
                    float w = gauss_w(dx, dy);
                    int tap = nx >= 1 && ny >= 1 && nx + 1 < xc && ny + 1 < yc;
                    uint32_t ni = ny * xc + nx;
                    float a = cvx, b = cvy, c = cpv;
                    if (tap && ba[ni] == 0) {
                        a = ovx[ni];
                        b = ovy[ni];
                        c = opv[ni];
                    }
                    sx_ += a * w;
                    sy_ += b * w;
                    sp_ += c * w;
                    ws += w;
                }
            float nvx = sx_ / ws, nvy = sy_ / ws, npv = sp_ / ws;
            float tx0, ty0;
            oracle_far_advect_target(x, y, nvx, nvy, ba, xc, yc, 0, &tx0, &ty0);
            int i0 = ora_f2i(tx0), j0 = ora_f2i(ty0);
            float fx = tx0 - (float)i0, fy = ty0 - (float)j0;
            if (ba[idx] == 0 && i0 >= 2 && i0 < (int)xc - 3 && j0 >= 2 && j0 < (int)yc - 3 && fx >= 0.0f &&
                fy >= 0.0f) {
                uint32_t iu = (uint32_t)i0, r0 = (uint32_t)j0 * xc, r1 = ((uint32_t)j0 + 1) * xc;
                float w00 = (1.0f - fx) * (1.0f - fy);
                float w10 = fx * (1.0f - fy);
                float w01 = (1.0f - fx) * fy;
                float w11 = fx * fy;
                nvx = nvx * 0.7f + 0.3f * (ovx[r0 + iu] * w00 + ovx[r0 + iu + 1] * w10 + ovx[r1 + iu] * w01 +
                                           ovx[r1 + iu + 1] * w11);
                nvy = nvy * 0.7f + 0.3f * (ovy[r0 + iu] * w00 + ovy[r0 + iu + 1] * w10 + ovy[r1 + iu] * w01 +
                                           ovy[r1 + iu + 1] * w11);
            }
            if (s->vorticity_coeff > 0.0f && x > 1 && x + 2 < xc && y > 1 && y + 2 < yc) {
                float w = vorticity_bmap(ovx, ovy, ba, xc, yc, x, y);
                float wxp = vorticity_bmap(ovx, ovy, ba, xc, yc, x + 1, y);
                float wxm = vorticity_bmap(ovx, ovy, ba, xc, yc, x - 1, y);
                float wyp = vorticity_bmap(ovx, ovy, ba, xc, yc, x, y + 1);
                float wym = vorticity_bmap(ovx, ovy, ba, xc, yc, x, y - 1);
                if (wxp < 0.0f) wxp = 0.0f - wxp;
                if (wxm < 0.0f) wxm = 0.0f - wxm;
                if (wyp < 0.0f) wyp = 0.0f - wyp;
                if (wym < 0.0f) wym = 0.0f - wym;
                float dwx = (wxp - wxm) * 0.5f;
                float dwy = (wyp - wym) * 0.5f;
                float norm = sqrtf(dwx * dwx + dwy * dwy);
                float k = s->vorticity_coeff / 5.0f;
                nvx += k * dwy / (norm + 0.001f) * w;
                nvy += k * (0.0f - dwx) / (norm + 0.001f) * w;
            }
            if (s->wall_bmap[idx] == 5u) {
                nvx += s->fan_vx[idx];
                nvy += s->fan_vy[idx];
            }
            nvx = ora_clampf(nvx, -256.0f, 256.0f);
            nvy = ora_clampf(nvy, -256.0f, 256.0f);
            npv = ora_clampf(npv, -256.0f, 256.0f);
            if (s->air_mode == 1) npv = 0.0f;
            if (s->air_mode == 2) { nvx = 0.0f; nvy = 0.0f; }
            if (s->air_mode == 3) { nvx = 0.0f; nvy = 0.0f; npv = 0.0f; }
            s->air_vx[idx] = nvx;
            s->air_vy[idx] = nvy;
            s->air_pv[idx] = npv;
        }
    (void)n;
}

void oracle_air_heat_convect(oracle_state *s) {
    const uint32_t xc = s->xres / 4, yc = s->yres / 4, n = xc * yc;
    const float *hv = s->air_hv;
    const uint32_t *bh = s->block_airh;
    /* Jacobi (rule R4): snapshot vx/vy, write into fresh arrays. */
    float *vx0 = (float *)malloc(sizeof(float) * n), *vy0 = (float *)malloc(sizeof(float) * n);
    memcpy(vx0, s->air_vx, sizeof(float) * n);
    memcpy(vy0, s->air_vy, sizeof(float) * n);
    const float amb = s->ambient_air_temp;
#pragma omp parallel for schedule(static)
    for (uint32_t y = 0; y < yc; ++y)
        for (uint32_t x = 0; x < xc; ++x) {
            uint32_t idx = y * xc + x;
            float center = hv[idx], cvx = vx0[idx], cvy = vy0[idx];
            float sh = 0.0f, sx_ = 0.0f, sy_ = 0.0f, ws = 0.0f;
            for (uint32_t dy = 0; dy < 3; ++dy)
                for (uint32_t dx = 0; dx < 3; ++dx) {
                    uint32_t nx = x + dx, ny = y + dy;
                    nx = nx >= 1 ? nx - 1 : 0;
                    ny = ny >= 1 ? ny - 1 : 0;
                    float w = gauss_w(dx, dy);
                    int tap = nx >= 1 && ny >= 1 && nx + 1 < xc && ny + 1 < yc;
                    uint32_t ni = ny * xc + nx;
                    float a = center, b = cvx, c = cvy;
                    if (tap && (bh[ni] & 0x8u) == 0) {
                        a = hv[ni];
                        b = vx0[ni];
                        c = vy0[ni];
                    }
                    sh += a * w;
                    sx_ += b * w;
                    sy_ += c * w;
                    ws += w;
                }
            float dh = sh / ws, svx = sx_ / ws, svy = sy_ / ws;
            float tx, ty;
            oracle_far_advect_target(x, y, svx, svy, bh, xc, yc, 1, &tx, &ty);
            int i0 = ora_f2i(tx), j0 = ora_f2i(ty);
            float fx = tx - (float)i0, fy = ty - (float)j0;
            if ((bh[idx] & 0x8u) == 0 && i0 >= 0 && i0 < (int)xc - 1 && j0 >= 0 && j0 < (int)yc - 1 &&
                fx >= 0.0f && fy >= 0.0f) {
                uint32_t iu = (uint32_t)i0, r0 = (uint32_t)j0 * xc, r1 = ((uint32_t)j0 + 1) * xc;
                float s00 = (bh[r0 + iu] & 0x8u) ? dh : hv[r0 + iu];
                float s10 = (bh[r0 + iu + 1] & 0x8u) ? dh : hv[r0 + iu + 1];
                float s01 = (bh[r1 + iu] & 0x8u) ? dh : hv[r1 + iu];
                float s11 = (bh[r1 + iu + 1] & 0x8u) ? dh : hv[r1 + iu + 1];
                float w00 = (1.0f - fx) * (1.0f - fy);
                float w10 = fx * (1.0f - fy);
                float w01 = (1.0f - fx) * fy;
                float w11 = fx * fy;
                dh = dh * 0.7f + 0.3f * (s00 * w00 + s10 * w10 + s01 * w01 + s11 * w11);
            }
            if (bh[idx] & 0x8u) dh = center;
            dh = ora_clampf(dh, 0.0f, 9999.0f);
            if (in_border(x, y, xc, yc)) dh = amb;
            if (s->aheat_enable == 0) dh = amb;
            s->air_ohv[idx] = dh;

            if (x >= 2 && x + 2 < xc && y >= 2 && y + 2 < yc) {
                float gx = 0.0f, gy = 0.0f;
                if (s->gravity_mode == 0) {
                    gy = 1.0f;
                } else if (s->gravity_mode == 2) {
                    float rx = (float)xc * 0.5f - ((float)x + 0.5f);
                    float ry = (float)yc * 0.5f - ((float)y + 0.5f);
                    float rd = sqrtf(rx * rx + ry * ry);
                    if (rd > 0.0f) {
                        gx = rx / rd;
                        gy = ry / rd;
                    }
                } else if (s->gravity_mode == 3) {
                    gx = s->custom_gravity_x;
                    gy = s->custom_gravity_y;
                }
                float cx_ = 0.0f - gx, cy_ = 0.0f - gy;
                float nvx = cvx, nvy = cvy;
                if (s->convection_mode == 1) {
                    uint32_t rm = (y - 1) * xc;
                    float weight = ((center - hv[idx - 1]) * cx_ + (center - hv[rm + x]) * cy_) / 5000.0f;
                    if (weight > 0.0f && (bh[rm + x] & 0x8u) == 0) {
                        nvx += weight * cx_;
                        nvy += weight * cy_;
                    }
                } else if (s->convection_mode == 2) {
                    float gm = sqrtf(cx_ * cx_ + cy_ * cy_);
                    if (gm > 10.0f) {
                        cx_ /= 0.1f * gm;
                        cy_ /= 0.1f * gm;
                    }
                    float weight = (center - amb) / 10000.0f;
                    if (weight > 0.01f) weight = 0.01f;
                    nvx += weight * cx_;
                    nvy += weight * cy_;
                }
                nvx = ora_clampf(nvx, -256.0f, 256.0f);
                nvy = ora_clampf(nvy, -256.0f, 256.0f);
                s->air_vx[idx] = nvx;
                s->air_vy[idx] = nvy;
            }
        }
    free(vx0);
    free(vy0);
}

void oracle_air_copy_back(oracle_state *s) {
    const uint32_t n = (s->xres / 4) * (s->yres / 4);
    memcpy(s->air_hv, s->air_ohv, sizeof(float) * n);
}

/* settle.rs:51-93; WL_WALLELEC=1, WL_EWALL=2, WL_WALL=8, WL_GRAV=14, WL_BLOCKAIR=16 */
void oracle_air_block_masks(oracle_state *s) {
    const uint32_t n = (s->xres / 4) * (s->yres / 4);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t wall = s->wall_bmap[i] & 0xFFu;
        uint32_t raw = s->wall_emap[i];
        if (raw != 0) s->wall_emap[i] = raw - 1;
        uint32_t emap = s->wall_emap[i] & 0xFFu;
        uint32_t blocked = (wall == 8 || wall == 1 || wall == 16 || (wall == 2 && emap == 0)) ? 1u : 0u;
        s->block_air[i] = blocked | s->elem_block_air[i];
        uint32_t heat = (blocked || wall == 14) ? 0x8u : 0u;
        s->block_airh[i] = heat | s->elem_block_airh[i];
    }
}

/* settle.rs:119-160; PT_GAS=10, PT_NBLE=97, hot_air is word 11 of the row */
void oracle_hot_air_pressure(oracle_state *s) {
    const uint32_t xc = s->xres / 4, yc = s->yres / 4;
#pragma omp parallel for schedule(static)
    for (uint32_t cy = 0; cy < yc; ++cy)
        for (uint32_t cx = 0; cx < xc; ++cx) {
            uint32_t cell = cy * xc + cx;
            for (uint32_t dy = 0; dy < 4; ++dy)
                for (uint32_t dx = 0; dx < 4; ++dx) {
                    uint32_t px = cx * 4 + dx, py = cy * 4 + dy;
                    if (px >= s->xres || py >= s->yres) continue;
                    uint32_t e = s->pmap[py * s->xres + px];
                    if (e == 0) continue;
                    uint32_t ty = e & 0x1FFu;
                    float hot = ora_bits_f(s->elem_props[ty * 32 + 11]);
                    if (hot != 0.0f) {
                        if (ty == 10 || ty == 97) {
                            if (s->air_pv[cell] < 3.5f) s->air_pv[cell] += 4.0f * hot * (3.5f - s->air_pv[cell]);
                        } else {
                            s->air_pv[cell] += 4.0f * hot;
                        }
                    }
                }
        }
}

/* air_kernel/mod.rs:120-299 */
void oracle_step_air(oracle_state *s) {
    if (s->air_mode != 4) {
        oracle_air_pressure(s);
        oracle_air_velocity(s);
        oracle_air_smooth_advect(s);
    }
    if (s->aheat_enable) {
        oracle_air_heat_convect(s);
        oracle_air_copy_back(s);
    }
    oracle_air_block_masks(s);
    oracle_hot_air_pressure(s);
}
