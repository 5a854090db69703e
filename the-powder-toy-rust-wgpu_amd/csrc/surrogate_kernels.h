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
// bf16 values are stored as uint16_t and converted with round-to-nearest-even.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sur {

typedef uint16_t bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (bf16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).
__device__ __forceinline__ float gelu_t(float z) {
    const float u = 0.7978845608f * (z + 0.044715f * z * z * z);
    return tanhf(u);
}
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + gelu_t(z)); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float t = gelu_t(z);
    const float du = 0.7978845608f * (1.0f + 3.0f * 0.044715f * z * z);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// ---------------------------------------------------------------------------
// Convolution: 256 threads (4 waves).  Each wave owns 16 * MS consecutive
// pixels of one row (MS 16-pixel M sub-tiles) and 16 * NT output channels
// (NT 16-channel N sub-tiles), so a workgroup covers MS rows of a 64-pixel
// segment.  The input tile (rows y0-1 .. y0+MS, columns x0-1 .. x0+64)
// sits in LDS with a pixel stride of cin + 8 bf16, which spreads the 16
// pixels of a fragment read over the banks.  The K loop fetches the next
// chunk's fragments (LDS patches, weight rows from L2) before it issues the
// current chunk's MFMAs.

constexpr int CONV_THREADS = 256;
constexpr int CONV_SEG = 64;
constexpr int CONV_LCOLS = CONV_SEG + 2;

struct ConvArgs {
    const bf16_t *in;     // H x W x cin
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

__host__ __device__ inline size_t conv_lds_bytes(int cin, int ms) {
    return (size_t)(ms + 2) * CONV_LCOLS * (cin + 8) * sizeof(bf16_t);
}

template <int NT, int MS>
__device__ __forceinline__ void conv_fetch(const bf16_t *tile, const bf16_t *const *wrow, const ConvArgs &a, int kc,
                                           int g, int m, int row, int px0, int ps, uint4 (&av)[MS],
                                           uint4 (&bv)[NT]) {
    const int k0 = 32 * kc + 8 * g, tap = k0 >> a.lcin, ci0 = k0 & (a.cin - 1);
    if (tap < 9) {
        const int dy = tap / 3, dx = tap - 3 * dy;
        const bf16_t *base = tile + ((row + dy) * CONV_LCOLS + px0 + m + dx) * ps + ci0;
#pragma unroll
        for (int j = 0; j < MS; ++j) av[j] = *reinterpret_cast<const uint4 *>(base + 16 * j * ps);
    } else {
        const uint4 z = (tap == 9 && ci0 == 0 && a.bias_col) ? make_uint4(0x3F80u, 0u, 0u, 0u)  // bf16 1.0
                                                             : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int j = 0; j < MS; ++j) av[j] = z;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) bv[t] = *reinterpret_cast<const uint4 *>(wrow[t] + k0);
}

template <int NT, int MS>
__global__ __launch_bounds__(CONV_THREADS) void conv3x3(ConvArgs a) {
    extern __shared__ __align__(16) bf16_t tile[];
    constexpr int WAVES_PER_ROW = 4 / MS;
    const int ps = a.cin + 8;
    const int x0 = blockIdx.x * CONV_SEG, y0 = blockIdx.y * MS, nbase = blockIdx.z * 16 * NT;
    const int tid = threadIdx.x;
    // stage the input tile with a zero halo ('same' padding)
    const int v8 = a.cin >> 3, total = (MS + 2) * CONV_LCOLS * v8;
    for (int e = tid; e < total; e += CONV_THREADS) {
        const int v = e % v8, rc = e / v8, c = rc % CONV_LCOLS, r = rc / CONV_LCOLS;
        const int gy = y0 - 1 + r, gx = x0 - 1 + c;
        uint4 val = make_uint4(0u, 0u, 0u, 0u);
        if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W)
            val = *reinterpret_cast<const uint4 *>(a.in + ((size_t)gy * a.W + gx) * a.cin + 8 * v);
        *reinterpret_cast<uint4 *>(tile + (r * CONV_LCOLS + c) * ps + 8 * v) = val;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, m = lane & 15, g = lane >> 4;
    const int row = wave / WAVES_PER_ROW, px0 = (wave % WAVES_PER_ROW) * 16 * MS;
    f32x4 acc[MS][NT];
#pragma unroll
    for (int j = 0; j < MS; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[j][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const bf16_t *wrow[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) wrow[t] = a.w + (size_t)(nbase + 16 * t + m) * a.kpad;
    uint4 av[MS], bv[NT];
    conv_fetch<NT, MS>(tile, wrow, a, 0, g, m, row, px0, ps, av, bv);
    for (int kc = 0; kc < a.kchunks; ++kc) {
        uint4 an[MS], bn[NT];
        if (kc + 1 < a.kchunks) conv_fetch<NT, MS>(tile, wrow, a, kc + 1, g, m, row, px0, ps, an, bn);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 b = as_bf16x8(bv[t]);
#pragma unroll
            for (int j = 0; j < MS; ++j)
                acc[j][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(av[j]), b, acc[j][t], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < MS; ++j) av[j] = an[j];
#pragma unroll
        for (int t = 0; t < NT; ++t) bv[t] = bn[t];
    }
    const int y = y0 + row;
    if (y >= a.H) return;
#pragma unroll
    for (int j = 0; j < MS; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = nbase + 16 * t + m;
            if (co >= a.cout) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int x = x0 + px0 + 16 * j + 4 * g + i;
                if (x >= a.W) continue;
                const size_t o = ((size_t)y * a.W + x) * a.cout + co;
                float v = acc[j][t][i];
                if (a.zout) a.zout[o] = f2bf(v);
                if (a.gelu) v = gelu(v);
                if (a.zref) v *= gelu_grad(bf2f(a.zref[o]));
                if (a.out_f32) a.out_f32[o] = v;
                if (a.out) a.out[o] = f2bf(v);
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
// U-Net glue (NHWC bf16).

// 2x2 average pool: out (H/2 x W/2 x C).
__global__ void pool2(const bf16_t *__restrict__ in, bf16_t *__restrict__ out, int H, int W, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, h = H / 2, w = W / 2;
    if (idx >= h * w * C) return;
    const int c = idx % C, p = idx / C, x = p % w, y = p / w;
    const size_t b = ((size_t)(2 * y) * W + 2 * x) * C + c;
    const float s = bf2f(in[b]) + bf2f(in[b + C]) + bf2f(in[b + (size_t)W * C]) + bf2f(in[b + (size_t)W * C + C]);
    out[idx] = f2bf(0.25f * s);
}

// out = nearest-neighbour 2x upsample of lo (H/2 x W/2 x C) + skip (H x W x C).
__global__ void up2_add(const bf16_t *__restrict__ lo, const bf16_t *__restrict__ skip, bf16_t *__restrict__ out,
                        int H, int W, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= H * W * C) return;
    const int c = idx % C, p = idx / C, x = p % W, y = p / W;
    out[idx] = f2bf(bf2f(lo[((size_t)(y / 2) * (W / 2) + x / 2) * C + c]) + bf2f(skip[idx]));
}

// Backward of up2_add for the low branch: out_lo = (sum over each 2x2 of
// d_hi) * gelu'(z_lo).
__global__ void up2_bwd_gelu(const bf16_t *__restrict__ dhi, const bf16_t *__restrict__ zlo,
                             bf16_t *__restrict__ out, int H, int W, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x, h = H / 2, w = W / 2;
    if (idx >= h * w * C) return;
    const int c = idx % C, p = idx / C, x = p % w, y = p / w;
    const size_t b = ((size_t)(2 * y) * W + 2 * x) * C + c;
    const float s = bf2f(dhi[b]) + bf2f(dhi[b + C]) + bf2f(dhi[b + (size_t)W * C]) + bf2f(dhi[b + (size_t)W * C + C]);
    out[idx] = f2bf(s * gelu_grad(bf2f(zlo[idx])));
}

// Backward of pool2 plus a skip gradient: out = (0.25 d_lo[up] + d_skip) *
// gelu'(z), at full resolution.
__global__ void pool2_bwd_add_gelu(const bf16_t *__restrict__ dlo, const bf16_t *__restrict__ dskip,
                                   const bf16_t *__restrict__ z, bf16_t *__restrict__ out, int H, int W, int C) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= H * W * C) return;
    const int c = idx % C, p = idx / C, x = p % W, y = p / W;
    const float d = 0.25f * bf2f(dlo[((size_t)(y / 2) * (W / 2) + x / 2) * C + c]) + bf2f(dskip[idx]);
    out[idx] = f2bf(d * gelu_grad(bf2f(z[idx])));
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

// Sum of squares of the gradient vector into *acc (one atomic per wave).
__global__ void grad_sumsq(const float *__restrict__ grad, int n, float *__restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float v = i < n ? grad[i] * grad[i] : 0.0f;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(acc, v);
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
