// Device kernels of the learned air surrogate: a 3x3 convolution as an
// implicit GEMM on gfx950's bf16 MFMA (v_mfma_f32_16x16x32_bf16, fp32
// accumulation), its weight gradient, and the elementwise U-Net glue.
//
// Layouts
//  * activations: NHWC bf16, one grid (batch 1), channels contiguous;
//  * a layer's parameters: one bf16 matrix W[co][k] of `kpad` columns with
//    k = tap * cin + ci (tap = 3 * dy + dx over the 3x3 window) and the bias
//    in column 9 * cin, which meets a constant-1 input; every other padding
//    column is zero.  The fp32 master copy, the Adam moments and the
//    gradient share that layout, so one update covers weights and bias;
//  * the data gradient runs the same convolution with W transposed and
//    flipped (Wt[ci][tap' * cout + co] = W[co][(8 - tap') * cin + ci]).
//
// bf16 values are stored as uint16_t and converted with round-to-nearest-even
// (gfx950's v_cvt_pk_bf16_f32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sur {

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16_t f2bf(float f) { return __builtin_bit_cast(bf16_t, (__bf16)f); }
// Two floats to packed bf16 (low half first), one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t{lo, hi}), bf16x2_t));
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))), evaluated
// as z * sigmoid(2 sqrt(2/pi) (z + 0.044715 z^3)) (0.5 (1 + tanh u) =
// sigmoid(2 u)) with the hardware exponential and reciprocal.
__device__ __forceinline__ float gelu_s(float z) {
    const float u = 1.5957691216f * (z + 0.044715f * z * z * z);
    return __builtin_amdgcn_rcpf(1.0f + __expf(-u));
}
__device__ __forceinline__ float gelu(float z) { return z * gelu_s(z); }
// d/dz: s + z s (1 - s) * 2 sqrt(2/pi) (1 + 3 * 0.044715 z^2).
__device__ __forceinline__ float gelu_grad(float z) {
    const float s = gelu_s(z);
    return s + z * s * (1.0f - s) * (1.5957691216f * (1.0f + 3.0f * 0.044715f * z * z));
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// 8 bf16 <-> 8 floats, 16-byte vectors.
__device__ __forceinline__ void unpack8(uint4 v, float (&f)[8]) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f[2 * i] = __uint_as_float(w[i] << 16);
        f[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
    }
}
__device__ __forceinline__ uint4 pack8(const float (&f)[8]) {
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = f2bf2(f[2 * i], f[2 * i + 1]);
    return make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ __forceinline__ uint4 ld8(const bf16_t *p) { return *reinterpret_cast<const uint4 *>(p); }

// ---------------------------------------------------------------------------
// Convolution: 256 threads (4 waves).  The MFMA's A operand is the weight
// matrix (rows = output channels), its B operand the input patches (columns
// = pixels), so that a lane's accumulator holds 4 consecutive output
// channels of one pixel and leaves as one 8-byte store.  Each wave owns the
// 64 pixels of one row segment (4 sub-tiles of 16 pixels) and 16 * NT output
// channels (NT sub-tiles of 16).  KS waves share one pixel tile and split its
// K chunks (split K for the small grids of the coarse levels); their partial
// sums meet in LDS and each of them finishes a quarter or half of the tile.
// A workgroup covers 4 / KS rows of a 64-pixel segment.  The input tile
// (rows y0-1 .. y0+4/KS, columns x0-1 .. x0+64) sits in LDS with a pixel
// stride of cin + 8 bf16, which spreads the 16 pixels of a fragment read
// over the banks.  The K loop is unrolled by two with separate fragment
// registers, so the loads of one chunk are in flight while the other
// chunk's MFMAs issue.

constexpr int CONV_THREADS = 256;
constexpr int CONV_SEG = 64;
constexpr int CONV_LCOLS = CONV_SEG + 2;

// How a conv stages its input tile:
//  STAGE_PLAIN  in is H x W x cin;
//  STAGE_POOL   in is 2H x 2W x cin and the tile holds its 2x2 average
//               (pool2's arithmetic and rounding);
//  STAGE_UPADD  in is the skip (H x W x cin), in2 the low level
//               (H/2 x W/2 x cin); the tile holds up2_add's sum;
//  STAGE_PREP   (LO kernels, cin 8) the tile holds the network's input
//               built from the air planes (prep_channels), in two bf16 halves.
enum { STAGE_PLAIN = 0, STAGE_POOL = 1, STAGE_UPADD = 2, STAGE_PREP = 3 };

// The 8 input channels of one cell: pressure, velocity, heat above ambient
// (each over its scale), wall mask, heat-source temperature above ambient
// over the heat scale, two zeros.
__device__ __forceinline__ void prep_channels(float pv, float vx, float vy, float hv, float wall, float src,
                                              float sp, float sv, float sh, float amb, float (&v)[8]) {
    v[0] = pv / sp;
    v[1] = vx / sv;
    v[2] = vy / sv;
    v[3] = (hv - amb) / sh;
    v[4] = wall;
    v[5] = src > 0.0f ? (src - amb) / sh : 0.0f;
    v[6] = 0.0f;
    v[7] = 0.0f;
}

struct ConvArgs {
    const bf16_t *in;     // H x W x cin (see the staging modes)
    const bf16_t *in_lo;  // LO kernels: the low bf16 half of the input (in + in_lo = the fp32 value)
    const bf16_t *in2;    // STAGE_UPADD: the low-level input
    int stage;            // STAGE_*
    // STAGE_PREP: the air planes, wall word per cell, source temperatures, scales
    const float *pv, *vx, *vy, *hv, *src;
    const uint32_t *wall;
    float sp, sv, sh, amb;
    const bf16_t *w;      // cout rows (padded to 64) x kpad
    bf16_t *out;          // H x W x cout, bf16 (may be null when out_f32 is set)
    bf16_t *zout;         // forward: pre-activation, or null
    const bf16_t *zref;   // data gradient: pre-activation whose GELU derivative scales the result, or null
    float *out_f32;       // fp32 copy of the output, or null
    int H, W, cin, lcin, cout;
    int kpad;             // weight row stride (columns)
    int kchunks;          // k columns the forward sums, / 32 (9 * cin + 1 rounded up)
    int gelu;             // forward: apply GELU
    int bias_col;         // the k = 9 * cin column meets a constant 1
};

__host__ __device__ inline size_t conv_lds_bytes(int cin, int ks, int nt, bool lo = false) {
    const size_t tile = (size_t)(lo ? 2 : 1) * (4 / ks + 2) * CONV_LCOLS * (cin + 8) * sizeof(bf16_t);
    const size_t part = ks > 1 ? (size_t)4 * 4 * nt * 64 * sizeof(f32x4) : 0;
    return tile > part ? tile : part;
}

template <int NT>
__device__ __forceinline__ void conv_fetch(const bf16_t *tile, const bf16_t *const *wrow, const ConvArgs &a, int kc,
                                           int g, int m, int row, int ps, int bias, uint4 (&xv)[4],
                                           uint4 (&wv)[NT]) {
    const int k0 = 32 * kc + 8 * g, tap = k0 >> a.lcin, ci0 = k0 & (a.cin - 1);
#pragma unroll
    for (int t = 0; t < NT; ++t) wv[t] = *reinterpret_cast<const uint4 *>(wrow[t] + k0);
    if (tap < 9) {
        const int dy = tap / 3, dx = tap - 3 * dy;
        const bf16_t *base = tile + ((row + dy) * CONV_LCOLS + m + dx) * ps + ci0;
#pragma unroll
        for (int j = 0; j < 4; ++j) xv[j] = *reinterpret_cast<const uint4 *>(base + 16 * j * ps);
    } else {
        const uint4 z = (tap == 9 && ci0 == 0 && bias) ? make_uint4(0x3F80u, 0u, 0u, 0u)  // bf16 1.0
                                                             : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int j = 0; j < 4; ++j) xv[j] = z;
    }
}

// conv_fetch for cin >= 32: a K chunk (32 columns) then lies inside one
// tap, so the tap, its window offset and the chunk's first channel are the
// same for every lane (scalar arithmetic), and a lane's LDS and weight
// addresses are its fixed bases plus that offset.  Same operands as
// conv_fetch.
template <int NT>
__device__ __forceinline__ void conv_fetch_wide(const bf16_t *lane_tile, const bf16_t *const *lane_w, const ConvArgs &a,
                                                int kc, int ps, int g, int bias, uint4 (&xv)[4], uint4 (&wv)[NT]) {
    const int k0 = 32 * kc, tap = k0 >> a.lcin, cib = k0 & (a.cin - 1);
#pragma unroll
    for (int t = 0; t < NT; ++t) wv[t] = *reinterpret_cast<const uint4 *>(lane_w[t] + k0);
    if (tap < 9) {
        const int dy = tap / 3, dx = tap - 3 * dy;
        const bf16_t *base = lane_tile + (dy * CONV_LCOLS + dx) * ps + cib;
#pragma unroll
        for (int j = 0; j < 4; ++j) xv[j] = *reinterpret_cast<const uint4 *>(base + 16 * j * ps);
    } else {
        // the bias chunk: column 9 * cin (lane group 0's first element) meets a 1
        const uint4 z = (g == 0 && bias) ? make_uint4(0x3F80u, 0u, 0u, 0u) : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int j = 0; j < 4; ++j) xv[j] = z;
    }
}

template <int NT>
__device__ __forceinline__ void conv_mfma(f32x4 (&acc)[4][NT], const uint4 (&xv)[4], const uint4 (&wv)[NT]) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const bf16x8 w = as_bf16x8(wv[t]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            acc[j][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, as_bf16x8(xv[j]), acc[j][t], 0, 0, 0);
    }
}

// Output channels co .. co + 3 of pixel (x, y): the pre-activation (zout),
// GELU or the data gradient's GELU derivative, the fp32 and bf16 outputs.
__device__ __forceinline__ void conv_store(const ConvArgs &a, f32x4 v, int x, int y, int co) {
    if (co >= a.cout || x >= a.W || y >= a.H) return;
    const size_t o = ((size_t)y * a.W + x) * a.cout + co;
    if (a.zout) {
        *reinterpret_cast<uint2 *>(a.zout + o) = make_uint2(f2bf2(v[0], v[1]), f2bf2(v[2], v[3]));
    }
    if (a.gelu) {
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = gelu(v[i]);
    }
    if (a.zref) {
        const uint2 zr = *reinterpret_cast<const uint2 *>(a.zref + o);
        v[0] *= gelu_grad(bf2f((bf16_t)(zr.x & 0xFFFFu)));
        v[1] *= gelu_grad(bf2f((bf16_t)(zr.x >> 16)));
        v[2] *= gelu_grad(bf2f((bf16_t)(zr.y & 0xFFFFu)));
        v[3] *= gelu_grad(bf2f((bf16_t)(zr.y >> 16)));
    }
    if (a.out_f32) *reinterpret_cast<float4 *>(a.out_f32 + o) = make_float4(v[0], v[1], v[2], v[3]);
    if (a.out) *reinterpret_cast<uint2 *>(a.out + o) = make_uint2(f2bf2(v[0], v[1]), f2bf2(v[2], v[3]));
}

// Stages the input tile (rows y0-1 .. y0+rows, columns x0-1 .. x0+64, zero
// outside the grid; 8 channels of one pixel per element) into LDS.  Each
// thread issues the loads of STAGE_BATCH elements before it combines and
// stores any of them, so a tile costs a few memory latencies rather than
// one per element.
constexpr int STAGE_BATCH = 4;

template <bool LO>
__device__ __forceinline__ void conv_stage(const ConvArgs &a, bf16_t *tile, int tile_elems, int ps, int x0, int y0,
                                           int rows) {
    // cin is a power of two (8 .. 128): an element's 8-channel group and
    // column come from shifts and masks
    const int lv8 = a.lcin - 3, v8 = 1 << lv8, total = (rows + 2) * CONV_LCOLS * v8;
    for (int base = threadIdx.x; base < total; base += CONV_THREADS * STAGE_BATCH) {
        int dst[STAGE_BATCH];
        bool ok[STAGE_BATCH];
        size_t src[STAGE_BATCH];
        int gy[STAGE_BATCH], gx[STAGE_BATCH], ch[STAGE_BATCH];
#pragma unroll
        for (int k = 0; k < STAGE_BATCH; ++k) {
            const int e = base + k * CONV_THREADS;
            const int v = e & (v8 - 1), rc = e >> lv8, r = rc / CONV_LCOLS, c = rc - r * CONV_LCOLS;
            gy[k] = y0 - 1 + r;
            gx[k] = x0 - 1 + c;
            ch[k] = 8 * v;
            dst[k] = e < total ? (r * CONV_LCOLS + c) * ps + 8 * v : -1;
            ok[k] = e < total && gy[k] >= 0 && gy[k] < a.H && gx[k] >= 0 && gx[k] < a.W;
            src[k] = ((size_t)gy[k] * a.W + gx[k]) * a.cin + 8 * v;
        }
        uint4 val[STAGE_BATCH], vlo[STAGE_BATCH];
        if (a.stage == STAGE_POOL) {
            uint4 q[STAGE_BATCH][4];
            const size_t W2 = 2 * (size_t)a.W;
#pragma unroll
            for (int k = 0; k < STAGE_BATCH; ++k) {
                const size_t b = ((2 * (size_t)gy[k]) * W2 + 2 * gx[k]) * a.cin + ch[k];
                const uint4 z = make_uint4(0u, 0u, 0u, 0u);
                q[k][0] = ok[k] ? ld8(a.in + b) : z;
                q[k][1] = ok[k] ? ld8(a.in + b + a.cin) : z;
                q[k][2] = ok[k] ? ld8(a.in + b + W2 * a.cin) : z;
                q[k][3] = ok[k] ? ld8(a.in + b + W2 * a.cin + a.cin) : z;
            }
#pragma unroll
            for (int k = 0; k < STAGE_BATCH; ++k) {
                float f0[8], f1[8], f2[8], f3[8], o[8];
                unpack8(q[k][0], f0);
                unpack8(q[k][1], f1);
                unpack8(q[k][2], f2);
                unpack8(q[k][3], f3);
#pragma unroll
                for (int i = 0; i < 8; ++i) o[i] = 0.25f * (((f0[i] + f1[i]) + f2[i]) + f3[i]);
                val[k] = ok[k] ? pack8(o) : make_uint4(0u, 0u, 0u, 0u);
            }
        } else if (a.stage == STAGE_UPADD) {
            uint4 lq[STAGE_BATCH], kq[STAGE_BATCH];
#pragma unroll
            for (int k = 0; k < STAGE_BATCH; ++k) {
                const uint4 z = make_uint4(0u, 0u, 0u, 0u);
                lq[k] = ok[k] ? ld8(a.in2 + ((size_t)(gy[k] / 2) * (a.W / 2) + gx[k] / 2) * a.cin + ch[k]) : z;
                kq[k] = ok[k] ? ld8(a.in + src[k]) : z;
            }
#pragma unroll
            for (int k = 0; k < STAGE_BATCH; ++k) {
                float l[8], kk[8];
                unpack8(lq[k], l);
                unpack8(kq[k], kk);
#pragma unroll
                for (int i = 0; i < 8; ++i) l[i] += kk[i];
                val[k] = ok[k] ? pack8(l) : make_uint4(0u, 0u, 0u, 0u);
            }
        } else if (LO && a.stage == STAGE_PREP) {
            float pv[STAGE_BATCH], vx[STAGE_BATCH], vy[STAGE_BATCH], hv[STAGE_BATCH], sr[STAGE_BATCH];
            uint32_t wl[STAGE_BATCH];
#pragma unroll
            for (int k = 0; k < STAGE_BATCH; ++k) {
                const size_t cell = ok[k] ? (size_t)gy[k] * a.W + gx[k] : 0;
                pv[k] = a.pv[cell];
                vx[k] = a.vx[cell];
                vy[k] = a.vy[cell];
                hv[k] = a.hv[cell];
                sr[k] = a.src[cell];
                wl[k] = a.wall[cell];
            }
#pragma unroll
            for (int k = 0; k < STAGE_BATCH; ++k) {
                float c8[8], hi[8], lo[8];
                prep_channels(pv[k], vx[k], vy[k], hv[k], (float)(wl[k] != 0u), sr[k], a.sp, a.sv, a.sh, a.amb, c8);
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    hi[i] = bf2f(f2bf(c8[i]));
                    lo[i] = c8[i] - hi[i];
                }
                val[k] = ok[k] ? pack8(hi) : make_uint4(0u, 0u, 0u, 0u);
                vlo[k] = ok[k] ? pack8(lo) : make_uint4(0u, 0u, 0u, 0u);
            }
        } else {
#pragma unroll
            for (int k = 0; k < STAGE_BATCH; ++k) {
                const uint4 z = make_uint4(0u, 0u, 0u, 0u);
                val[k] = ok[k] ? ld8(a.in + src[k]) : z;
                if (LO) vlo[k] = ok[k] ? ld8(a.in_lo + src[k]) : z;
            }
        }
#pragma unroll
        for (int k = 0; k < STAGE_BATCH; ++k) {
            if (dst[k] < 0) continue;
            *reinterpret_cast<uint4 *>(tile + dst[k]) = val[k];
            if (LO) *reinterpret_cast<uint4 *>(tile + tile_elems + dst[k]) = vlo[k];
        }
    }
}

// LO: the input comes as two bf16 halves (in, in_lo) that meet the same
// weights, so the first layer sees its fp32 input to about 16 bits.
template <int NT, int KS, bool LO>
__global__ __launch_bounds__(CONV_THREADS) void conv3x3(ConvArgs a) {
    extern __shared__ __align__(16) bf16_t tile[];
    constexpr int ROWS = 4 / KS;
    const int ps = a.cin + 8, tile_elems = (ROWS + 2) * CONV_LCOLS * ps;
    const int x0 = blockIdx.x * CONV_SEG, y0 = blockIdx.y * ROWS, nbase = blockIdx.z * 16 * NT;
    const int tid = threadIdx.x;
    conv_stage<LO>(a, tile, tile_elems, ps, x0, y0, ROWS);
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, m = lane & 15, g = lane >> 4;
    const int row = wave / KS, kg = wave % KS;
    const int kbeg = kg * a.kchunks / KS, kend = (kg + 1) * a.kchunks / KS;
    f32x4 acc[4][NT];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[j][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const bf16_t *wrow[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) wrow[t] = a.w + (size_t)(nbase + 16 * t + m) * a.kpad;
    uint4 x0v[4], w0v[NT], x1v[4], w1v[NT];
    int kc = kbeg;
    if (LO) {
        // short K (the first layer): fetch, then both halves through the MFMAs
        for (; kc < kend; ++kc) {
            conv_fetch<NT>(tile, wrow, a, kc, g, m, row, ps, a.bias_col, x0v, w0v);
            conv_fetch<NT>(tile + tile_elems, wrow, a, kc, g, m, row, ps, 0, x1v, w1v);
            conv_mfma<NT>(acc, x0v, w0v);
            conv_mfma<NT>(acc, x1v, w0v);
        }
    } else if (a.cin >= 32) {
        const bf16_t *lane_tile = tile + (row * CONV_LCOLS + m) * ps + 8 * g;
        const bf16_t *lane_w[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) lane_w[t] = wrow[t] + 8 * g;
        if (kc < kend) conv_fetch_wide<NT>(lane_tile, lane_w, a, kc, ps, g, a.bias_col, x0v, w0v);
        for (; kc + 1 < kend; kc += 2) {
            conv_fetch_wide<NT>(lane_tile, lane_w, a, kc + 1, ps, g, a.bias_col, x1v, w1v);
            conv_mfma<NT>(acc, x0v, w0v);
            if (kc + 2 < kend) conv_fetch_wide<NT>(lane_tile, lane_w, a, kc + 2, ps, g, a.bias_col, x0v, w0v);
            conv_mfma<NT>(acc, x1v, w1v);
        }
        if (kc < kend) conv_mfma<NT>(acc, x0v, w0v);
    } else {
        if (kc < kend) conv_fetch<NT>(tile, wrow, a, kc, g, m, row, ps, a.bias_col, x0v, w0v);
        for (; kc + 1 < kend; kc += 2) {
            conv_fetch<NT>(tile, wrow, a, kc + 1, g, m, row, ps, a.bias_col, x1v, w1v);
            conv_mfma<NT>(acc, x0v, w0v);
            if (kc + 2 < kend) conv_fetch<NT>(tile, wrow, a, kc + 2, g, m, row, ps, a.bias_col, x0v, w0v);
            conv_mfma<NT>(acc, x1v, w1v);
        }
        if (kc < kend) conv_mfma<NT>(acc, x0v, w0v);
    }
    const int y = y0 + row;
    // lane: pixel x0 + 16 j + m, output channels nbase + 16 t + 4 g + (0..3)
    if (KS == 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int t = 0; t < NT; ++t) conv_store(a, acc[j][t], x0 + 16 * j + m, y, nbase + 16 * t + 4 * g);
        return;
    }
    // split K: every wave parks its partial sums, then wave kg of a row adds
    // the KS partials of pixel sub-tiles j = kg, kg + KS, ...
    f32x4 *part = reinterpret_cast<f32x4 *>(tile);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) part[((wave * 4 + j) * NT + t) * 64 + lane] = acc[j][t];
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < 4 / KS; ++jj) {
        const int j = kg + KS * jj;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            f32x4 s = part[(((row * KS) * 4 + j) * NT + t) * 64 + lane];
#pragma unroll
            for (int q = 1; q < KS; ++q) {
                const f32x4 u = part[(((row * KS + q) * 4 + j) * NT + t) * 64 + lane];
                s[0] += u[0];
                s[1] += u[1];
                s[2] += u[2];
                s[3] += u[3];
            }
            conv_store(a, s, x0 + 16 * j + m, y, nbase + 16 * t + 4 * g);
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: G[co][k] = sum over pixels p of dZ[p][co] * X(p, k), with
// X(p, k) the input at p shifted by tap(k) (zero outside the grid), 1 for
// the bias column, 0 for padding.  An MFMA GEMM with M = output channels,
// N = k columns, and the pixels as the reduction: each workgroup sums the
// pixels of WG_ROWS rows of one 64-pixel row segment for a 64-column k range
// (16 per wave) and all output channels, and writes its partial sums; a
// second kernel adds the partials.  dZ is staged transposed ([co][pixel]) so
// that an A fragment is one 16-byte read; the B fragment gathers 8 pixels of
// one input channel.

constexpr int WG_THREADS = 256;
constexpr int WG_ROWS = 8;
constexpr int WG_SEG = 64;

struct WgradArgs {
    const bf16_t *x;    // H x W x cin
    const bf16_t *dz;   // H x W x cout
    float *part;        // groups x cout16 x kpad partial sums
    int H, W, cin, lcin, cout, cout16, kpad;
};

__host__ __device__ inline size_t wgrad_lds_bytes(int cin, int cout16) {
    return ((size_t)3 * (WG_SEG + 2) * (cin + 8) + (size_t)cout16 * (WG_SEG + 8)) * sizeof(bf16_t);
}

template <int MT>  // output-channel sub-tiles of 16 (cout16 = 16 * MT)
__global__ __launch_bounds__(WG_THREADS) void conv3x3_wgrad(WgradArgs a) {
    extern __shared__ __align__(16) bf16_t lds[];
    const int ps = a.cin + 8, dzs = WG_SEG + 8;
    bf16_t *xt = lds, *dzt = lds + 3 * (WG_SEG + 2) * ps;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, m = lane & 15, g = lane >> 4;
    const int kbase = blockIdx.x * 64 + 16 * wave;  // this wave's 16 k columns
    const int x0 = blockIdx.z * WG_SEG, ry0 = blockIdx.y * WG_ROWS;
    // the wave's k column: tap and channel of lane m's column
    const int kcol = kbase + m, tap = kcol >> a.lcin, ci = kcol & (a.cin - 1);
    const int dy = tap < 9 ? tap / 3 : 0, dx = tap < 9 ? tap - 3 * (tap / 3) : 0;
    f32x4 acc[MT];
#pragma unroll
    for (int s = 0; s < MT; ++s) acc[s] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int v8 = a.cin >> 3;
    for (int yy = 0; yy < WG_ROWS; ++yy) {
        const int y = ry0 + yy;
        if (y >= a.H) break;
        __syncthreads();
        for (int e = tid; e < 3 * (WG_SEG + 2) * v8; e += WG_THREADS) {
            const int v = e % v8, rc = e / v8, c = rc % (WG_SEG + 2), r = rc / (WG_SEG + 2);
            const int gy = y - 1 + r, gx = x0 - 1 + c;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
                val = *reinterpret_cast<const uint4 *>(a.x + ((size_t)gy * a.W + gx) * a.cin + 8 * v);
            *reinterpret_cast<uint4 *>(xt + (r * (WG_SEG + 2) + c) * ps + 8 * v) = val;
        }
        // dZ of row y, transposed to [co][pixel] (zeros past cout and the grid)
        for (int e = tid; e < a.cout16 * WG_SEG; e += WG_THREADS) {
            const int px = e % WG_SEG, co = e / WG_SEG, gx = x0 + px;
            bf16_t val = 0;
            if (co < a.cout && gx < a.W) val = a.dz[((size_t)y * a.W + gx) * a.cout + co];
            dzt[co * dzs + px] = val;
        }
        __syncthreads();
#pragma unroll
        for (int kc = 0; kc < WG_SEG / 32; ++kc) {
            const int p0 = 32 * kc + 8 * g;  // the lane's 8 pixels
            // B fragment: X(p0 + i, kcol) for i < 8
            bf16x8 b;
            if (tap < 9) {
                const bf16_t *src = xt + (dy * (WG_SEG + 2) + p0 + dx) * ps + ci;
#pragma unroll
                for (int i = 0; i < 8; ++i) b[i] = __builtin_bit_cast(__bf16, src[i * ps]);
            } else {
                const bool bias = tap == 9 && ci == 0;
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    b[i] = __builtin_bit_cast(__bf16, (bf16_t)(bias && x0 + p0 + i < a.W ? 0x3F80u : 0u));
            }
#pragma unroll
            for (int s = 0; s < MT; ++s) {
                const bf16x8 av = as_bf16x8(*reinterpret_cast<const uint4 *>(dzt + (16 * s + m) * dzs + p0));
                acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b, acc[s], 0, 0, 0);
            }
        }
    }
    // lane holds rows (output channels) 16 s + 4 g + i of column kbase + m
    const size_t group = (size_t)blockIdx.y * gridDim.z + blockIdx.z;
    float *dst = a.part + group * (size_t)a.cout16 * a.kpad;
#pragma unroll
    for (int s = 0; s < MT; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i) dst[(size_t)(16 * s + 4 * g + i) * a.kpad + kbase + m] = acc[s][i];
}

// grad[co][k] = sum of the partials (co < cout; rows past it are padding).
__global__ void wgrad_reduce(const float *__restrict__ part, float *__restrict__ grad, int groups, int cout,
                             int cout16, int kpad, int accumulate) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= cout * kpad) return;
    const size_t stride = (size_t)cout16 * kpad;
    float s = 0.0f;
    for (int gI = 0; gI < groups; ++gI) s += part[gI * stride + idx];
    grad[idx] = accumulate ? grad[idx] + s : s;
}

// ---------------------------------------------------------------------------
// U-Net glue (NHWC bf16).  One thread per 8 channels of a pixel (C is a
// multiple of 8), 16-byte loads and stores.


// 2x2 average pool: out (H/2 x W/2 x C).
__global__ void pool2(const bf16_t *__restrict__ in, bf16_t *__restrict__ out, int H, int W, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, h = H / 2, w = W / 2, c8 = C >> 3;
    if (idx >= h * w * c8) return;
    const int c = 8 * (idx % c8), p = idx / c8, x = p % w, y = p / w;
    const size_t b = ((size_t)(2 * y) * W + 2 * x) * C + c;
    float f0[8], f1[8], f2[8], f3[8], o[8];
    unpack8(ld8(in + b), f0);
    unpack8(ld8(in + b + C), f1);
    unpack8(ld8(in + b + (size_t)W * C), f2);
    unpack8(ld8(in + b + (size_t)W * C + C), f3);
#pragma unroll
    for (int i = 0; i < 8; ++i) o[i] = 0.25f * (((f0[i] + f1[i]) + f2[i]) + f3[i]);
    *reinterpret_cast<uint4 *>(out + (size_t)p * C + c) = pack8(o);
}

// out = nearest-neighbour 2x upsample of lo (H/2 x W/2 x C) + skip (H x W x C).
__global__ void up2_add(const bf16_t *__restrict__ lo, const bf16_t *__restrict__ skip, bf16_t *__restrict__ out,
                        int H, int W, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, c8 = C >> 3;
    if (idx >= H * W * c8) return;
    const int c = 8 * (idx % c8), p = idx / c8, x = p % W, y = p / W;
    float l[8], k[8];
    unpack8(ld8(lo + ((size_t)(y / 2) * (W / 2) + x / 2) * C + c), l);
    unpack8(ld8(skip + (size_t)p * C + c), k);
#pragma unroll
    for (int i = 0; i < 8; ++i) l[i] += k[i];
    *reinterpret_cast<uint4 *>(out + (size_t)p * C + c) = pack8(l);
}

// Backward of up2_add for the low branch: out_lo = (sum over each 2x2 of
// d_hi) * gelu'(z_lo).
__global__ void up2_bwd_gelu(const bf16_t *__restrict__ dhi, const bf16_t *__restrict__ zlo,
                             bf16_t *__restrict__ out, int H, int W, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, h = H / 2, w = W / 2, c8 = C >> 3;
    if (idx >= h * w * c8) return;
    const int c = 8 * (idx % c8), p = idx / c8, x = p % w, y = p / w;
    const size_t b = ((size_t)(2 * y) * W + 2 * x) * C + c;
    float f0[8], f1[8], f2[8], f3[8], z[8], o[8];
    unpack8(ld8(dhi + b), f0);
    unpack8(ld8(dhi + b + C), f1);
    unpack8(ld8(dhi + b + (size_t)W * C), f2);
    unpack8(ld8(dhi + b + (size_t)W * C + C), f3);
    unpack8(ld8(zlo + (size_t)p * C + c), z);
#pragma unroll
    for (int i = 0; i < 8; ++i) o[i] = (((f0[i] + f1[i]) + f2[i]) + f3[i]) * gelu_grad(z[i]);
    *reinterpret_cast<uint4 *>(out + (size_t)p * C + c) = pack8(o);
}

// Backward of pool2 plus a skip gradient: out = (0.25 d_lo[up] + d_skip) *
// gelu'(z), at full resolution.
__global__ void pool2_bwd_add_gelu(const bf16_t *__restrict__ dlo, const bf16_t *__restrict__ dskip,
                                   const bf16_t *__restrict__ z, bf16_t *__restrict__ out, int H, int W, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, c8 = C >> 3;
    if (idx >= H * W * c8) return;
    const int c = 8 * (idx % c8), p = idx / c8, x = p % W, y = p / W;
    float l[8], k[8], zz[8];
    unpack8(ld8(dlo + ((size_t)(y / 2) * (W / 2) + x / 2) * C + c), l);
    unpack8(ld8(dskip + (size_t)p * C + c), k);
    unpack8(ld8(z + (size_t)p * C + c), zz);
#pragma unroll
    for (int i = 0; i < 8; ++i) l[i] = (0.25f * l[i] + k[i]) * gelu_grad(zz[i]);
    *reinterpret_cast<uint4 *>(out + (size_t)p * C + c) = pack8(l);
}

// Wt[ci][tap' * cout + co] = W[co][(8 - tap') * cin + ci] (no bias column);
// every other entry of Wt's kpad_t columns is zero.
__global__ void transpose_flip(const bf16_t *__restrict__ w, bf16_t *__restrict__ wt, int cin, int cout, int kpad,
                               int rows_t, int kpad_t) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows_t * kpad_t) return;
    const int kt = idx % kpad_t, ci = idx / kpad_t;
    bf16_t v = 0;
    if (ci < cin && kt < 9 * cout) {
        const int tapt = kt / cout, co = kt - tapt * cout;
        v = w[(size_t)co * kpad + (8 - tapt) * cin + ci];
    }
    wt[idx] = v;
}

// Block sums of NC values per thread: block b writes red[b * NC + c]
// (blockDim.x = 256); sum_blocks adds them up in a fixed order.
template <int NC>
__device__ __forceinline__ void block_sum(float (&v)[NC], float *__restrict__ red) {
    __shared__ float wsum[4][NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
        for (int off = 32; off > 0; off >>= 1) v[c] += __shfl_down(v[c], off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NC; ++c) wsum[wave][c] = v[c];
    __syncthreads();
    if (threadIdx.x < NC)
        red[(size_t)blockIdx.x * NC + threadIdx.x] =
            ((wsum[0][threadIdx.x] + wsum[1][threadIdx.x]) + wsum[2][threadIdx.x]) + wsum[3][threadIdx.x];
}

// acc[c] += sum over b < nblocks of red[b * nc + c], for c < nc (one block
// of 256 threads).
__global__ void sum_blocks(const float *__restrict__ red, int nblocks, int nc, float *__restrict__ acc, int add) {
    __shared__ float part[256];
    for (int c = 0; c < nc; ++c) {
        float v = 0.0f;
        for (int b = threadIdx.x; b < nblocks; b += 256) v += red[(size_t)b * nc + c];
        part[threadIdx.x] = v;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) acc[c] = add ? acc[c] + part[0] : part[0];
        __syncthreads();
    }
}

// Block sums of squares of the gradient vector (256 threads per block).
__global__ void grad_sumsq(const float *__restrict__ grad, int n, float *__restrict__ red) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float v[1] = {i < n ? grad[i] * grad[i] : 0.0f};
    block_sum<1>(v, red);
}

// Adam on the fp32 master parameters, then the bf16 copy.  The gradient is
// the sum over the batch times `gscale`, scaled down so that its global L2
// norm (sqrt(*sumsq) * gscale) is at most `clip` when clip > 0.
__global__ void adam_step(float *__restrict__ wf, bf16_t *__restrict__ wb, const float *__restrict__ grad,
                          float *__restrict__ m1, float *__restrict__ m2, int n, float lr, float b1, float b2,
                          float eps, float c1, float c2, float gscale, const float *__restrict__ sumsq, float clip) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float f = gscale;
    if (clip > 0.0f) {
        const float norm = sqrtf(*sumsq) * gscale;
        if (norm > clip) f *= clip / norm;
    }
    const float gr = grad[i] * f;
    const float a = b1 * m1[i] + (1.0f - b1) * gr;
    const float b = b2 * m2[i] + (1.0f - b2) * gr * gr;
    m1[i] = a;
    m2[i] = b;
    const float w = wf[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    wf[i] = w;
    wb[i] = f2bf(w);
}

__global__ void f32_to_bf16(const float *__restrict__ src, bf16_t *__restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = f2bf(src[i]);
}

}  // namespace sur
