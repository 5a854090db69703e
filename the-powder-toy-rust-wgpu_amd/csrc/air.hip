// Eulerian air step for gfx950: air_kernel::step_air_gpu
// (crates/tpt-sim/src/air_kernel/mod.rs:120-299) as four launches instead of
// the reference's seven.
//
//   air_pv_kernel       pressure.rs:174-238 + pressure.rs:248-299, fused: the
//                       pressure of the tile plus a one-cell halo is computed
//                       into LDS, then the velocity reads its gradient there.
//   air_advect_kernel   advect.rs:176-360 (smoothing, far advection,
//                       vorticity, fans, caps, air modes).
//   air_heat_kernel     heat.rs:31-253 with the Jacobi rule for vx/vy: the
//                       new velocity goes to the (dead) ovx/ovy scratch.
//   air_settle_kernel   settle.rs:22-28 + settle.rs:51-93 + settle.rs:119-160:
//                       hv/vx/vy copy-back, emap decay, wall masks, HotAir.
//
// Tiles are 64 cells wide and 4 cells tall so that every wavefront reads a
// contiguous 256-byte row segment of each grid.

#include "tpt_internal.h"

namespace {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int BORDER = 2;
constexpr float ADVECT_DIST = 0.7f;

struct AirDims {
    int xc, yc;
};

__device__ __forceinline__ bool in_border(int x, int y, AirDims d) {
    return x < BORDER || y < BORDER || x + BORDER >= d.xc || y + BORDER >= d.yc;
}

__device__ __forceinline__ bool interior(int x, int y, AirDims d) {
    return x >= 1 && y >= 1 && x + 1 < d.xc && y + 1 < d.yc;
}

// A grid read through a window: element (x, y) of the canvas grid sits at
// p[(y - oy) * s + (x - ox)].  The pressure/velocity kernel reads its
// inputs through LDS copies of its tile plus a 3-cell halo.
template <typename T> struct GridView {
    const T *p;
    int ox, oy, s;
    __device__ __forceinline__ T at(int x, int y) const { return p[(y - oy) * s + (x - ox)]; }
};

// A canvas grid in global memory, read as at(x, y).
template <typename T> struct GlobalView {
    const T *p;
    int xc;
    __device__ __forceinline__ T at(int x, int y) const { return p[y * xc + x]; }
};

// Stages the window of `n` grids (origin x0 - h, y0 - h, (TX + 2h) x
// (TY + 2h)) into LDS, all loads in flight before the first store;
// cells off the canvas are staged as zero and never read.
template <int H, typename T0, typename T1, typename T2, typename T3>
__device__ __forceinline__ void stage4(const T0 *__restrict__ g0, const T1 *__restrict__ g1,
                                       const T2 *__restrict__ g2, const T3 *__restrict__ g3, T0 *s0, T1 *s1, T2 *s2,
                                       T3 *s3, int x0, int y0, AirDims d) {
    constexpr int W = TX + 2 * H, N = W * (TY + 2 * H), K = (N + TX * TY - 1) / (TX * TY);
    const int tid = threadIdx.y * TX + threadIdx.x;
    T0 a[K];
    T1 b[K];
    T2 c[K];
    T3 e[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = tid + k * TX * TY, ly = i / W, lx = i - ly * W;
        const int gx = x0 + lx - H, gy = y0 + ly - H;
        const bool in = i < N && gx >= 0 && gy >= 0 && gx < d.xc && gy < d.yc;
        const int g = in ? gy * d.xc + gx : 0;
        a[k] = in ? g0[g] : T0(0);
        b[k] = in ? g1[g] : T1(0);
        c[k] = in ? g2[g] : T2(0);
        e[k] = in ? g3[g] : T3(0);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = tid + k * TX * TY;
        if (i < N) {
            s0[i] = a[k];
            s1[i] = b[k];
            s2[i] = c[k];
            s3[i] = e[k];
        }
    }
}

__device__ __forceinline__ uint32_t blocked_at(const GridView<uint32_t> &b, int x, int y, AirDims d) {
    return (x >= 0 && y >= 0 && x < d.xc && y < d.yc) ? b.at(x, y) : 0u;
}

__device__ __forceinline__ bool wall_cleared_vx(const GridView<uint32_t> &b, int x, int y, AirDims d) {
    bool out = false;
    if (y >= 1 && y + 1 < d.yc) {
        if (x >= 1 && x + 1 < d.xc && blocked_at(b, x, y, d)) out = true;
        if (x >= 2 && blocked_at(b, x - 1, y, d)) out = true;
        if (x + 2 < d.xc && blocked_at(b, x + 1, y, d)) out = true;
    }
    return out;
}

__device__ __forceinline__ bool wall_cleared_vy(const GridView<uint32_t> &b, int x, int y, AirDims d) {
    bool out = false;
    if (x >= 1 && x + 1 < d.xc) {
        if (y >= 1 && y + 1 < d.yc && blocked_at(b, x, y, d)) out = true;
        if (y >= 2 && blocked_at(b, x, y - 1, d)) out = true;
        if (y + 2 < d.yc && blocked_at(b, x, y + 1, d)) out = true;
    }
    return out;
}

__device__ __forceinline__ float effective_v(const GridView<float> &v, const GridView<uint32_t> &b, int x, int y,
                                             AirDims d, float edge, bool along_x) {
    float val = v.at(x, y);
    if (in_border(x, y, d)) val = edge + (val - edge) * 0.9f;
    if (along_x ? wall_cleared_vx(b, x, y, d) : wall_cleared_vy(b, x, y, d)) val = 0.0f;
    return val;
}

// Pressure of pressure.rs:186-236 at (x, y).
__device__ __forceinline__ float pressure_at(const GridView<float> &pv, const GridView<float> &vx,
                                             const GridView<float> &vy, const GridView<uint32_t> &ba, int x, int y,
                                             AirDims d, float ep, float evx, float evy) {
    float cur = pv.at(x, y);
    if (in_border(x, y, d)) cur = ep + (cur - ep) * 0.8f;
    float next = cur;
    if (interior(x, y, d)) {
        float dp = effective_v(vx, ba, x - 1, y, d, evx, true) - effective_v(vx, ba, x + 1, y, d, evx, true) +
                   effective_v(vy, ba, x, y - 1, d, evy, false) - effective_v(vy, ba, x, y + 1, d, evy, false);
        next = ep + (cur - ep) * 0.9999f + dp * 0.15f;
    }
    return next;
}

// pressure.rs:174-238 then pressure.rs:248-299.  The tile's inputs plus a
// 3-cell halo (the velocity's pressure gradient reads the pressure one cell
// out, which reads velocity one more cell out, whose wall test reads the
// mask one more) are staged in LDS with every load in flight at once; the
// tile's pressure plus a one-cell halo is then computed into LDS, and the
// velocity reads its gradient there.
constexpr int PH = 3, SW = TX + 2 * PH, SH = TY + 2 * PH;

__global__ __launch_bounds__(TX *TY) void air_pv_kernel(const float *__restrict__ pv, const float *__restrict__ vx,
                                                        const float *__restrict__ vy,
                                                        const uint32_t *__restrict__ ba, float *__restrict__ opv,
                                                        float *__restrict__ ovx, float *__restrict__ ovy, AirDims d,
                                                        float ep, float evx, float evy) {
    __shared__ float s_in_pv[SH * SW], s_in_vx[SH * SW], s_in_vy[SH * SW];
    __shared__ uint32_t s_in_ba[SH * SW];
    __shared__ float s_pv[TY + 2][TX + 2];
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    const int tid = threadIdx.y * TX + threadIdx.x;
    constexpr int STAGE = (SH * SW + TX * TY - 1) / (TX * TY);
    {
        float a[STAGE], b[STAGE], c[STAGE];
        uint32_t m[STAGE];
#pragma unroll
        for (int k = 0; k < STAGE; ++k) {
            const int i = tid + k * TX * TY, ly = i / SW, lx = i - ly * SW;
            const int gx = x0 + lx - PH, gy = y0 + ly - PH;
            const bool in = i < SH * SW && gx >= 0 && gy >= 0 && gx < d.xc && gy < d.yc;
            const int g = in ? gy * d.xc + gx : 0;
            a[k] = in ? pv[g] : 0.0f;
            b[k] = in ? vx[g] : 0.0f;
            c[k] = in ? vy[g] : 0.0f;
            m[k] = in ? ba[g] : 0u;
        }
#pragma unroll
        for (int k = 0; k < STAGE; ++k) {
            const int i = tid + k * TX * TY;
            if (i < SH * SW) {
                s_in_pv[i] = a[k];
                s_in_vx[i] = b[k];
                s_in_vy[i] = c[k];
                s_in_ba[i] = m[k];
            }
        }
    }
    __syncthreads();
    const GridView<float> P{s_in_pv, x0 - PH, y0 - PH, SW}, VX{s_in_vx, x0 - PH, y0 - PH, SW},
        VY{s_in_vy, x0 - PH, y0 - PH, SW};
    const GridView<uint32_t> B{s_in_ba, x0 - PH, y0 - PH, SW};
    // Pressure for the tile and its one-cell halo.
    for (int i = tid; i < (TY + 2) * (TX + 2); i += TX * TY) {
        int ly = i / (TX + 2), lx = i % (TX + 2);
        int gx = x0 + lx - 1, gy = y0 + ly - 1;
        float p = 0.0f;
        if (gx >= 0 && gy >= 0 && gx < d.xc && gy < d.yc) p = pressure_at(P, VX, VY, B, gx, gy, d, ep, evx, evy);
        s_pv[ly][lx] = p;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= d.xc || y >= d.yc) return;
    const int idx = y * d.xc + x;
    const int lx = threadIdx.x + 1, ly = threadIdx.y + 1;
    opv[idx] = s_pv[ly][lx];

    float cvx = VX.at(x, y), cvy = VY.at(x, y);
    if (in_border(x, y, d)) {
        cvx = evx + (cvx - evx) * 0.9f;
        cvy = evy + (cvy - evy) * 0.9f;
    }
    if (wall_cleared_vx(B, x, y, d)) cvx = 0.0f;
    if (wall_cleared_vy(B, x, y, d)) cvy = 0.0f;
    float nvx = cvx, nvy = cvy;
    if (interior(x, y, d)) {
        float dx = s_pv[ly][lx - 1] - s_pv[ly][lx + 1];
        float dy = s_pv[ly - 1][lx] - s_pv[ly + 1][lx];
        nvx = evx + (cvx - evx) * 0.999f + dx * 0.2f;
        nvy = evy + (cvy - evy) * 0.999f + dy * 0.2f;
        if (blocked_at(B, x - 1, y, d) || blocked_at(B, x, y, d) || blocked_at(B, x + 1, y, d)) nvx = 0.0f;
        if (blocked_at(B, x, y - 1, d) || blocked_at(B, x, y, d) || blocked_at(B, x, y + 1, d)) nvy = 0.0f;
    }
    ovx[idx] = nvx;
    ovy[idx] = nvy;
}

// advect.rs:66-152.
template <typename B>
__device__ __forceinline__ void far_advect_target(int x, int y, float dx, float dy, const B &block, AirDims d,
                                                  bool heat, float &otx, float &oty) {
    float tx = (float)x - dx * ADVECT_DIST;
    float ty = (float)y - dy * ADVECT_DIST;
    float adx = dx * ADVECT_DIST;
    if (adx < 0.0f) adx = 0.0f - adx;
    float ady = dy * ADVECT_DIST;
    if (ady < 0.0f) ady = 0.0f - ady;
    const float limit_x = (float)d.xc - 2.0f, limit_y = (float)d.yc - 2.0f;
    if ((adx > 1.0f || ady > 1.0f) && tx >= 2.0f && tx < limit_x && ty >= 2.0f && ty < limit_y) {
        float abs_dx = dx < 0.0f ? 0.0f - dx : dx;
        float abs_dy = dy < 0.0f ? 0.0f - dy : dy;
        float step_x, step_y;
        int step_limit;
        if (abs_dx > abs_dy) {
            step_x = dx < 0.0f ? 1.0f : -1.0f;
            step_y = 0.0f - dy / abs_dx;
            step_limit = tpt_f2i(adx);
        } else {
            step_x = 0.0f - dx / abs_dy;
            step_y = dy < 0.0f ? 1.0f : -1.0f;
            step_limit = tpt_f2i(ady);
        }
        tx = (float)x;
        ty = (float)y;
        bool stopped = false;
        for (int step = 0; step < step_limit; ++step) {
            tx += step_x;
            ty += step_y;
            int cx = tpt_f2i(tx + 0.5f), cy = tpt_f2i(ty + 0.5f);
            bool hit = true;
            if (cx >= 0 && cy >= 0 && cx < d.xc && cy < d.yc) {
                uint32_t b = block.at(cx, cy);
                hit = (heat ? (b & 0x8u) : b) != 0u;
            }
            if (hit) {
                tx -= step_x;
                ty -= step_y;
                stopped = true;
                break;
            }
        }
        if (!stopped) {
            tx = (float)x - dx * ADVECT_DIST;
            ty = (float)y - dy * ADVECT_DIST;
        }
    }
    otx = tx;
    oty = ty;
}

__device__ __forceinline__ float gauss_w(int dx, int dy) {
    float w = 0.01831564f;
    if (dx == 1) w = 0.13533528f;
    if (dy == 1) w = 0.13533528f;
    if (dx == 1 && dy == 1) w = 1.0f;
    return w;
}

// advect.rs:10-42.
__device__ __forceinline__ float vorticity_bmap(const GridView<float> &vx, const GridView<float> &vy,
                                                const GridView<uint32_t> &b, AirDims d, int x, int y) {
    float out = 0.0f;
    if (x > 1 && x + 2 < d.xc && y > 1 && y + 2 < d.yc) {
        bool here = b.at(x, y) != 0u;
        float dvydx = vy.at(x + 1, y) - vy.at(x - 1, y);
        if (here || b.at(x + 1, y) != 0u || b.at(x - 1, y) != 0u) dvydx = 0.0f;
        float dvxdy = vx.at(x, y + 1) - vx.at(x, y - 1);
        if (here || b.at(x, y + 1) != 0u || b.at(x, y - 1) != 0u) dvxdy = 0.0f;
        out = (dvydx - dvxdy) * 0.5f;
    }
    return out;
}

__global__ __launch_bounds__(TX *TY) void air_advect_kernel(
    const float *__restrict__ ovx, const float *__restrict__ ovy, const float *__restrict__ opv,
    const uint32_t *__restrict__ ba, const uint32_t *__restrict__ wall, const float *__restrict__ fan_vx,
    const float *__restrict__ fan_vy, float *__restrict__ vx, float *__restrict__ vy, float *__restrict__ pv,
    AirDims d, float vorticity_coeff, uint32_t air_mode) {
    constexpr int H = 2, W = TX + 2 * H, N = W * (TY + 2 * H);
    __shared__ float s_vx[N], s_vy[N], s_pv[N];
    __shared__ uint32_t s_ba[N];
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    stage4<H>(ovx, ovy, opv, ba, s_vx, s_vy, s_pv, s_ba, x0, y0, d);
    __syncthreads();
    // The smoothing, the vorticity and the cell itself read the LDS window;
    // the far-advection walk and its bilinear taps, which can leave it, read
    // global memory.
    const GridView<float> OVX{s_vx, x0 - H, y0 - H, W}, OVY{s_vy, x0 - H, y0 - H, W}, OPV{s_pv, x0 - H, y0 - H, W};
    const GridView<uint32_t> BA{s_ba, x0 - H, y0 - H, W};
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= d.xc || y >= d.yc) return;
    const int idx = y * d.xc + x;
    const float cvx = OVX.at(x, y), cvy = OVY.at(x, y), cpv = OPV.at(x, y);
    float sx = 0.0f, sy = 0.0f, sp = 0.0f, ws = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            int nx = x + dx - 1, ny = y + dy - 1;
            if (nx < 0) nx = 0;
            if (ny < 0) ny = 0;
            const float w = gauss_w(dx, dy);
            float a = cvx, b = cvy, c = cpv;
            if (nx >= 1 && ny >= 1 && nx + 1 < d.xc && ny + 1 < d.yc) {
                if (BA.at(nx, ny) == 0u) {
                    a = OVX.at(nx, ny);
                    b = OVY.at(nx, ny);
                    c = OPV.at(nx, ny);
                }
            }
            sx += a * w;
            sy += b * w;
            sp += c * w;
            ws += w;
        }
    float nvx = sx / ws, nvy = sy / ws, npv = sp / ws;

    float tx0, ty0;
    far_advect_target(x, y, nvx, nvy, GlobalView<uint32_t>{ba, d.xc}, d, false, tx0, ty0);
    const int i0 = tpt_f2i(tx0), j0 = tpt_f2i(ty0);
    const float fx = tx0 - (float)i0, fy = ty0 - (float)j0;
    if (BA.at(x, y) == 0u && i0 >= 2 && i0 < d.xc - 3 && j0 >= 2 && j0 < d.yc - 3 && fx >= 0.0f && fy >= 0.0f) {
        const int r0 = j0 * d.xc, r1 = (j0 + 1) * d.xc;
        const float w00 = (1.0f - fx) * (1.0f - fy);
        const float w10 = fx * (1.0f - fy);
        const float w01 = (1.0f - fx) * fy;
        const float w11 = fx * fy;
        nvx = nvx * 0.7f +
              0.3f * (ovx[r0 + i0] * w00 + ovx[r0 + i0 + 1] * w10 + ovx[r1 + i0] * w01 + ovx[r1 + i0 + 1] * w11);
        nvy = nvy * 0.7f +
              0.3f * (ovy[r0 + i0] * w00 + ovy[r0 + i0 + 1] * w10 + ovy[r1 + i0] * w01 + ovy[r1 + i0 + 1] * w11);
    }

    if (vorticity_coeff > 0.0f && x > 1 && x + 2 < d.xc && y > 1 && y + 2 < d.yc) {
        const float w = vorticity_bmap(OVX, OVY, BA, d, x, y);
        float wxp = vorticity_bmap(OVX, OVY, BA, d, x + 1, y);
        float wxm = vorticity_bmap(OVX, OVY, BA, d, x - 1, y);
        float wyp = vorticity_bmap(OVX, OVY, BA, d, x, y + 1);
        float wym = vorticity_bmap(OVX, OVY, BA, d, x, y - 1);
        if (wxp < 0.0f) wxp = 0.0f - wxp;
        if (wxm < 0.0f) wxm = 0.0f - wxm;
        if (wyp < 0.0f) wyp = 0.0f - wyp;
        if (wym < 0.0f) wym = 0.0f - wym;
        const float dwx = (wxp - wxm) * 0.5f;
        const float dwy = (wyp - wym) * 0.5f;
        const float norm = sqrtf(dwx * dwx + dwy * dwy);
        const float k = vorticity_coeff / 5.0f;
        nvx += k * dwy / (norm + 0.001f) * w;
        nvy += k * (0.0f - dwx) / (norm + 0.001f) * w;
    }
    if (wall[idx] == 5u) {
        nvx += fan_vx[idx];
        nvy += fan_vy[idx];
    }
    nvx = tpt_clampf(nvx, -256.0f, 256.0f);
    nvy = tpt_clampf(nvy, -256.0f, 256.0f);
    npv = tpt_clampf(npv, -256.0f, 256.0f);
    if (air_mode == 1u) npv = 0.0f;
    if (air_mode == 2u) {
        nvx = 0.0f;
        nvy = 0.0f;
    }
    if (air_mode == 3u) {
        nvx = 0.0f;
        nvy = 0.0f;
        npv = 0.0f;
    }
    vx[idx] = nvx;
    vy[idx] = nvy;
    pv[idx] = npv;
}

struct HeatParams {
    float ambient;
    uint32_t aheat_enable, convection_mode, gravity_mode;
    float custom_gx, custom_gy;
};

// heat.rs:31-253.  vx/vy are read only; the new velocity (or the unchanged
// one outside the convection window) goes to nvx/nvy.
__global__ __launch_bounds__(TX *TY) void air_heat_kernel(const float *__restrict__ hv,
                                                          const float *__restrict__ vx,
                                                          const float *__restrict__ vy,
                                                          const uint32_t *__restrict__ bh, float *__restrict__ ohv,
                                                          float *__restrict__ nvx_out, float *__restrict__ nvy_out,
                                                          AirDims d, HeatParams p) {
    constexpr int H = 2, W = TX + 2 * H, N = W * (TY + 2 * H);
    __shared__ float s_hv[N], s_vx[N], s_vy[N];
    __shared__ uint32_t s_bh[N];
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    stage4<H>(hv, vx, vy, bh, s_hv, s_vx, s_vy, s_bh, x0, y0, d);
    __syncthreads();
    // The smoothing, the convection taps and the cell itself read the LDS
    // window; the far-advection walk and its bilinear taps read global memory.
    const GridView<float> HV{s_hv, x0 - H, y0 - H, W}, VX{s_vx, x0 - H, y0 - H, W}, VY{s_vy, x0 - H, y0 - H, W};
    const GridView<uint32_t> BH{s_bh, x0 - H, y0 - H, W};
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= d.xc || y >= d.yc) return;
    const int idx = y * d.xc + x;
    const float center = HV.at(x, y), cvx = VX.at(x, y), cvy = VY.at(x, y);
    float sh = 0.0f, sx = 0.0f, sy = 0.0f, ws = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            int nx = x + dx - 1, ny = y + dy - 1;
            if (nx < 0) nx = 0;
            if (ny < 0) ny = 0;
            const float w = gauss_w(dx, dy);
            float a = center, b = cvx, c = cvy;
            if (nx >= 1 && ny >= 1 && nx + 1 < d.xc && ny + 1 < d.yc) {
                if ((BH.at(nx, ny) & 0x8u) == 0u) {
                    a = HV.at(nx, ny);
                    b = VX.at(nx, ny);
                    c = VY.at(nx, ny);
                }
            }
            sh += a * w;
            sx += b * w;
            sy += c * w;
            ws += w;
        }
    float dh = sh / ws;
    const float svx = sx / ws, svy = sy / ws;
    float tx, ty;
    far_advect_target(x, y, svx, svy, GlobalView<uint32_t>{bh, d.xc}, d, true, tx, ty);
    const int i0 = tpt_f2i(tx), j0 = tpt_f2i(ty);
    const float fx = tx - (float)i0, fy = ty - (float)j0;
    const uint32_t own = BH.at(x, y);
    if ((own & 0x8u) == 0u && i0 >= 0 && i0 < d.xc - 1 && j0 >= 0 && j0 < d.yc - 1 && fx >= 0.0f && fy >= 0.0f) {
        const int r0 = j0 * d.xc, r1 = (j0 + 1) * d.xc;
        const float s00 = (bh[r0 + i0] & 0x8u) ? dh : hv[r0 + i0];
        const float s10 = (bh[r0 + i0 + 1] & 0x8u) ? dh : hv[r0 + i0 + 1];
        const float s01 = (bh[r1 + i0] & 0x8u) ? dh : hv[r1 + i0];
        const float s11 = (bh[r1 + i0 + 1] & 0x8u) ? dh : hv[r1 + i0 + 1];
        const float w00 = (1.0f - fx) * (1.0f - fy);
        const float w10 = fx * (1.0f - fy);
        const float w01 = (1.0f - fx) * fy;
        const float w11 = fx * fy;
        dh = dh * 0.7f + 0.3f * (s00 * w00 + s10 * w10 + s01 * w01 + s11 * w11);
    }
    if (own & 0x8u) dh = center;
    dh = tpt_clampf(dh, 0.0f, 9999.0f);
    if (in_border(x, y, d)) dh = p.ambient;
    if (p.aheat_enable == 0u) dh = p.ambient;
    ohv[idx] = dh;

    float nvx = cvx, nvy = cvy;
    if (x >= 2 && x + 2 < d.xc && y >= 2 && y + 2 < d.yc) {
        float gx = 0.0f, gy = 0.0f;
        if (p.gravity_mode == 0u) {
            gy = 1.0f;
        } else if (p.gravity_mode == 2u) {
            const float rx = (float)d.xc * 0.5f - ((float)x + 0.5f);
            const float ry = (float)d.yc * 0.5f - ((float)y + 0.5f);
            const float rd = sqrtf(rx * rx + ry * ry);
            if (rd > 0.0f) {
                gx = rx / rd;
                gy = ry / rd;
            }
        } else if (p.gravity_mode == 3u) {
            gx = p.custom_gx;
            gy = p.custom_gy;
        }
        float cx = 0.0f - gx, cy = 0.0f - gy;
        if (p.convection_mode == 1u) {
            const float weight = ((center - HV.at(x - 1, y)) * cx + (center - HV.at(x, y - 1)) * cy) / 5000.0f;
            if (weight > 0.0f && (BH.at(x, y - 1) & 0x8u) == 0u) {
                nvx += weight * cx;
                nvy += weight * cy;
            }
        } else if (p.convection_mode == 2u) {
            const float gm = sqrtf(cx * cx + cy * cy);
            if (gm > 10.0f) {
                cx /= 0.1f * gm;
                cy /= 0.1f * gm;
            }
            float weight = (center - p.ambient) / 10000.0f;
            if (weight > 0.01f) weight = 0.01f;
            nvx += weight * cx;
            nvy += weight * cy;
        }
        nvx = tpt_clampf(nvx, -256.0f, 256.0f);
        nvy = tpt_clampf(nvy, -256.0f, 256.0f);
    }
    nvx_out[idx] = nvx;
    nvy_out[idx] = nvy;
}

// settle.rs:22-28, 51-93, 119-160 for one cell.
__global__ __launch_bounds__(TX *TY) void air_settle_kernel(
    const float *__restrict__ ohv, const float *__restrict__ nvx, const float *__restrict__ nvy,
    float *__restrict__ hv, float *__restrict__ vx, float *__restrict__ vy, const uint32_t *__restrict__ wall_bmap,
    uint32_t *__restrict__ wall_emap, uint32_t *__restrict__ ba, uint32_t *__restrict__ bh,
    const uint32_t *__restrict__ eba, const uint32_t *__restrict__ ebh, const uint32_t *__restrict__ pmap,
    const uint32_t *__restrict__ elem_props, float *__restrict__ pv, AirDims d, uint32_t xres, uint32_t copy_heat,
    tpt_tick_begin tb) {
    const int x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    // No kernel of the air pass reads the counter block, so the tick can
    // open here instead of in a launch of its own.
    if (tb.counters && x == 0 && y == 0) tpt_tick_begin_dev(tb);
    if (x >= d.xc || y >= d.yc) return;
    const int idx = y * d.xc + x;
    if (copy_heat) {
        hv[idx] = ohv[idx];
        vx[idx] = nvx[idx];
        vy[idx] = nvy[idx];
    }
    const uint32_t wall = wall_bmap[idx] & 0xFFu;
    uint32_t em = wall_emap[idx];
    if (em != 0u) {
        em -= 1u;
        wall_emap[idx] = em;
    }
    const uint32_t emap = em & 0xFFu;
    const uint32_t blocked = (wall == 8u || wall == 1u || wall == 16u || (wall == 2u && emap == 0u)) ? 1u : 0u;
    ba[idx] = blocked | eba[idx];
    bh[idx] = ((blocked || wall == 14u) ? 0x8u : 0u) | ebh[idx];

    // HotAir: one cell owns its 4x4 pixels; each row is one 16-byte load.
    float p = pv[idx];
    bool changed = false;
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
        const uint4 row = *reinterpret_cast<const uint4 *>(pmap + (size_t)(y * 4 + dy) * xres + x * 4);
#pragma unroll
        for (int dx = 0; dx < 4; ++dx) {
            const uint32_t e = dx == 0 ? row.x : dx == 1 ? row.y : dx == 2 ? row.z : row.w;
            if (e == 0u) continue;
            const uint32_t ty = e & 0x1FFu;
            const float hot = __uint_as_float(elem_props[ty * 32u + 11u]);
            if (hot != 0.0f) {
                if (ty == 10u || ty == 97u) {
                    if (p < 3.5f) p += 4.0f * hot * (3.5f - p);
                } else {
                    p += 4.0f * hot;
                }
                changed = true;
            }
        }
    }
    if (changed) pv[idx] = p;
}

}  // namespace

int tpt_air_step(tpt_ctx *ctx, const tpt_tick_begin *tb) {
    const AirDims d{(int)ctx->xcells, (int)ctx->ycells};
    const dim3 block(TX, TY);
    const dim3 grid((ctx->xcells + TX - 1) / TX, (ctx->ycells + TY - 1) / TY);
    const tpt_config &c = ctx->cfg;
    float *pv = tpt_buf<float>(ctx, TPT_BUF_AIR_PV), *vx = tpt_buf<float>(ctx, TPT_BUF_AIR_VX),
          *vy = tpt_buf<float>(ctx, TPT_BUF_AIR_VY), *hv = tpt_buf<float>(ctx, TPT_BUF_AIR_HV);
    float *opv = tpt_buf<float>(ctx, TPT_BUF_AIR_OPV), *ovx = tpt_buf<float>(ctx, TPT_BUF_AIR_OVX),
          *ovy = tpt_buf<float>(ctx, TPT_BUF_AIR_OVY), *ohv = tpt_buf<float>(ctx, TPT_BUF_AIR_OHV);
    uint32_t *ba = tpt_buf<uint32_t>(ctx, TPT_BUF_BLOCK_AIR), *bh = tpt_buf<uint32_t>(ctx, TPT_BUF_BLOCK_AIRH);
    int rc;
    if (c.air_mode != 4u) {
        tpt_pass_begin(ctx, PASS_AIR_PV);
        hipLaunchKernelGGL(air_pv_kernel, grid, block, 0, ctx->stream, pv, vx, vy, ba, opv, ovx, ovy, d,
                           c.edge_pressure, c.edge_velocity_x, c.edge_velocity_y);
        tpt_pass_end(ctx, PASS_AIR_PV);
        if ((rc = tpt_check_launch("air_pv_kernel"))) return rc;
        tpt_pass_begin(ctx, PASS_AIR_ADVECT);
        hipLaunchKernelGGL(air_advect_kernel, grid, block, 0, ctx->stream, ovx, ovy, opv, ba,
                           tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP), tpt_buf<float>(ctx, TPT_BUF_FAN_VX),
                           tpt_buf<float>(ctx, TPT_BUF_FAN_VY), vx, vy, pv, d, c.vorticity_coeff, c.air_mode);
        tpt_pass_end(ctx, PASS_AIR_ADVECT);
        if ((rc = tpt_check_launch("air_advect_kernel"))) return rc;
    }
    if (c.aheat_enable) {
        HeatParams hp{c.ambient_air_temp, c.aheat_enable, c.convection_mode, c.gravity_mode, c.custom_gravity_x,
                      c.custom_gravity_y};
        tpt_pass_begin(ctx, PASS_AIR_HEAT);
        hipLaunchKernelGGL(air_heat_kernel, grid, block, 0, ctx->stream, hv, vx, vy, bh, ohv, ovx, ovy, d, hp);
        tpt_pass_end(ctx, PASS_AIR_HEAT);
        if ((rc = tpt_check_launch("air_heat_kernel"))) return rc;
    }
    tpt_pass_begin(ctx, PASS_AIR_SETTLE);
    hipLaunchKernelGGL(air_settle_kernel, grid, block, 0, ctx->stream, ohv, ovx, ovy, hv, vx, vy,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP), tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP), ba, bh,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_BLOCK_AIR), tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_BLOCK_AIRH),
                       tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP), tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS), pv, d,
                       c.xres, c.aheat_enable ? 1u : 0u, tb ? *tb : tpt_tick_begin{nullptr, 0u, 0u, 0u});
    tpt_pass_end(ctx, PASS_AIR_SETTLE);
    return tpt_check_launch("air_settle_kernel");
}
