// Newtonian gravity for gfx950: gravity_kernel::step_gravity_gpu
// (crates/tpt-sim/src/gravity_kernel.rs:826-1016) with GravityFft::solve
// (crates/tpt-sim/src/gravity_gpu/mod.rs:114-278).
//
// The reference transforms the whole zero-padded W x H grid with one launch
// per radix-2 stage (22 full-grid passes per 2-D transform at 1024^2, three
// transforms per tick).  Here every 1-D transform is a Stockham radix-4
// transform (one radix-2 stage for odd lengths) inside LDS, with the
// twiddle table staged in LDS, and the solve is three launches:
//
//   grav_rows_fwd   rows 0..ycells-1 only (the other rows of the padded mass
//                   grid are zero, so is their transform): mass -> float,
//                   mask, zero-pad and forward row FFT, fused.
//   grav_cols       a tile of adjacent columns per workgroup: forward column
//                   FFT over the ycells non-zero rows, product with the
//                   cached kernel spectrum, inverse column FFT; only the
//                   ycells rows the extraction reads are stored.
//   grav_rows_inv   the inverse row FFT of one row + extraction into
//                   grav_fx / grav_fy under the gravity mask.
//
// The two force components share one transform.  The x and y kernels are
// real, so the field kx + i ky transforms to Kx + i Ky, and with the real
// mass spectrum F the products F Kx and F Ky are Hermitian: their inverse
// transforms are real (up to rounding), so the inverse of F (Kx + i Ky) is
// fx + i fy.  One kernel spectrum, one product and one inverse transform
// replace two of each.
//
// The kernel spectra (build_kernel_field_kernel, fft.rs:209-236, forward
// transformed) depend on the canvas only and are built once per context.
// Twiddles are exp(-2 pi i k / N) evaluated in double and rounded once.

#include "tpt_internal.h"

#include <math.h>
#include <mutex>
#include <stdlib.h>

namespace {

constexpr int FFT_THREADS = 256;
// Complex values one FFT workgroup holds in LDS: columns per workgroup in the
// column pass (and sequences in the row passes) x transform length.
constexpr int FFT_ELEMS = 4096;
constexpr int R4_PER_THREAD = FFT_ELEMS / 4 / FFT_THREADS;  // radix-4 butterflies per thread
constexpr int R2_PER_THREAD = FFT_ELEMS / 2 / FFT_THREADS;  // radix-2 butterflies per thread
constexpr float M_GRAV = 6.673e-1f;
constexpr int GRAV_BLOCK = 256;

// exp(-2 pi i e / N) for e in [0, N) from the half table tab[t] =
// exp(-2 pi i t / N), t < N/2 (the upper half is its negation); conjugated
// for the inverse transform (sign = -1).
__device__ __forceinline__ float2 twiddle(const float2 *tab, int e, int half, float sign) {
    float2 w = e < half ? tab[e] : make_float2(-tab[e - half].x, -tab[e - half].y);
    w.y *= sign;
    return w;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 w) {
    return make_float2(w.x * a.x - w.y * a.y, w.x * a.y + w.y * a.x);
}

// Copies the half twiddle table for length n = 2^logn out of the global
// table for length 2^tw_bits.
__device__ void load_twiddles(float2 *tab, const float2 *__restrict__ tw, int logn, int tw_bits) {
    const int half = (1 << logn) >> 1, shift = tw_bits - logn;
    for (int t = threadIdx.x; t < half; t += blockDim.x) tab[t] = tw[t << shift];
}

// In-place Stockham FFT of C = 2^logc interleaved sequences of length
// n = 2^logn held in LDS (element i of sequence c at buf[i * C + c]), natural
// order in and out: radix-4 stages, then one radix-2 stage when logn is odd.
// Each stage reads its butterflies' inputs into registers, waits, and writes
// the outputs where the next stage expects them.  C * n <= FFT_ELEMS.
__device__ void lds_fft(float2 *buf, const float2 *tab, int logn, int logc, float sign) {
    const int n = 1 << logn, C = 1 << logc, half = n >> 1;
    int s = 0;
    for (; s + 2 <= logn; s += 2) {
        const int L = 1 << s, quarter = (n >> 2) * C, nb = quarter;
        float2 y[R4_PER_THREAD][4];
#pragma unroll
        for (int u = 0; u < R4_PER_THREAD; ++u) {
            const int b = threadIdx.x + u * FFT_THREADS;
            if (b < nb) {
                const int k = (b >> logc) & (L - 1);
                const float2 a0 = buf[b], a1r = buf[b + quarter], a2r = buf[b + 2 * quarter],
                             a3r = buf[b + 3 * quarter];
                const int e = k << (logn - s - 2);
                const float2 a1 = cmul(a1r, twiddle(tab, e, half, sign));
                const float2 a2 = cmul(a2r, twiddle(tab, 2 * e, half, sign));
                const float2 a3 = cmul(a3r, twiddle(tab, 3 * e, half, sign));
                const float2 t0 = make_float2(a0.x + a2.x, a0.y + a2.y), t1 = make_float2(a0.x - a2.x, a0.y - a2.y);
                const float2 t2 = make_float2(a1.x + a3.x, a1.y + a3.y), t3 = make_float2(a1.x - a3.x, a1.y - a3.y);
                y[u][0] = make_float2(t0.x + t2.x, t0.y + t2.y);
                y[u][2] = make_float2(t0.x - t2.x, t0.y - t2.y);
                y[u][1] = make_float2(t1.x + sign * t3.y, t1.y - sign * t3.x);
                y[u][3] = make_float2(t1.x - sign * t3.y, t1.y + sign * t3.x);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < R4_PER_THREAD; ++u) {
            const int b = threadIdx.x + u * FFT_THREADS;
            if (b < nb) {
                const int c = b & (C - 1), j = b >> logc, k = j & (L - 1);
                const int base = (((j - k) << 2) + k) * C + c;
                buf[base] = y[u][0];
                buf[base + L * C] = y[u][1];
                buf[base + 2 * L * C] = y[u][2];
                buf[base + 3 * L * C] = y[u][3];
            }
        }
        __syncthreads();
    }
    if (s < logn) {
        const int L = 1 << s, nb = half * C;
        float2 y[R2_PER_THREAD][2];
#pragma unroll
        for (int u = 0; u < R2_PER_THREAD; ++u) {
            const int b = threadIdx.x + u * FFT_THREADS;
            if (b < nb) {
                const int k = (b >> logc) & (L - 1);
                const float2 a0 = buf[b];
                const float2 a1 = cmul(buf[b + nb], twiddle(tab, k << (logn - s - 1), half, sign));
                y[u][0] = make_float2(a0.x + a1.x, a0.y + a1.y);
                y[u][1] = make_float2(a0.x - a1.x, a0.y - a1.y);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < R2_PER_THREAD; ++u) {
            const int b = threadIdx.x + u * FFT_THREADS;
            if (b < nb) {
                const int c = b & (C - 1), j = b >> logc, k = j & (L - 1);
                const int base = (((j - k) << 1) + k) * C + c;
                buf[base] = y[u][0];
                buf[base + L * C] = y[u][1];
            }
        }
        __syncthreads();
    }
}

struct FftDims {
    int xc, yc, W, H, logW, logH, tw_bits, logcols;
};

// Column tile of this workgroup.  Workgroups are dealt to the 8 XCDs
// round-robin; numbering the tiles so that each XCD gets a contiguous run
// of columns keeps a row's 64-byte segment within one XCD's L2.
__device__ __forceinline__ int col_tile(const FftDims &d) {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    return (nb % 8) ? b : (b % 8) * (nb / 8) + b / 8;
}

// Dynamic LDS of the FFT kernels: the data plus the half twiddle table.
static inline size_t rows_lds(const FftDims &d, int seqs) { return ((size_t)d.W * seqs + d.W / 2) * sizeof(float2); }
static inline size_t cols_lds(const FftDims &d) {
    return (((size_t)d.H << d.logcols) + d.H / 2) * sizeof(float2);
}

// build_padded_mass_kernel (fft.rs:153-178) + forward row transform, one
// row per workgroup.
__global__ __launch_bounds__(FFT_THREADS) void grav_rows_fwd(const int32_t *__restrict__ mass,
                                                             const uint32_t *__restrict__ mask,
                                                             float2 *__restrict__ work,
                                                             const float2 *__restrict__ tw, FftDims d) {
    extern __shared__ float2 dyn[];
    float2 *sm = dyn, *tab = dyn + d.W;
    const int row = blockIdx.x;
    load_twiddles(tab, tw, d.logW, d.tw_bits);
    for (int i = threadIdx.x; i < d.W; i += FFT_THREADS) {
        float v = 0.0f;
        if (i < d.xc) {
            const int c = row * d.xc + i;
            if (mask[c] != 0u) v = (float)mass[c] * (1.0f / 65536.0f);
        }
        sm[i] = make_float2(v, 0.0f);
    }
    __syncthreads();
    lds_fft(sm, tab, d.logW, 0, 1.0f);
    for (int i = threadIdx.x; i < d.W; i += FFT_THREADS) work[(size_t)row * d.W + i] = sm[i];
}

// Forward column transform, product with the kernel spectrum
// (complex_mul_kernel, fft.rs:129-140) and inverse column transform, for
// 2^logcols adjacent columns per workgroup.  Only rows < rows_in of the
// input are non-zero and only rows < rows_out of the output are stored.
// Four waves per SIMD (128 VGPRs): the 1024 column workgroups then fit in
// one round on 256 CUs instead of one and a third (measured 44 -> 33 us at
// 2048x2048, with two products then).
__global__ __launch_bounds__(FFT_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void grav_cols(const float2 *__restrict__ work,
                                                         const float2 *__restrict__ kc, float2 *__restrict__ spec,
                                                         const float2 *__restrict__ tw, FftDims d, int rows_in,
                                                         int rows_out) {
    extern __shared__ float2 dyn[];
    const int logc = d.logcols, cols = 1 << logc, c0 = col_tile(d) * cols, total = d.H << logc;
    float2 *sm = dyn, *tab = dyn + total;
    load_twiddles(tab, tw, d.logH, d.tw_bits);
    for (int e = threadIdx.x; e < total; e += FFT_THREADS) {
        const int r = e >> logc, c = e & (cols - 1);
        sm[e] = r < rows_in ? work[(size_t)r * d.W + c0 + c] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    lds_fft(sm, tab, d.logH, logc, 1.0f);
    for (int e = threadIdx.x; e < total; e += FFT_THREADS) {
        const float2 k = kc[(size_t)(e >> logc) * d.W + c0 + (e & (cols - 1))];
        const float2 a = sm[e];
        sm[e] = make_float2(a.x * k.x - a.y * k.y, a.x * k.y + a.y * k.x);
    }
    __syncthreads();
    lds_fft(sm, tab, d.logH, logc, -1.0f);
    for (int e = threadIdx.x; e < (rows_out << logc); e += FFT_THREADS)
        spec[(size_t)(e >> logc) * d.W + c0 + (e & (cols - 1))] = sm[e];
}

// Inverse row transform of the product + extract_force_kernel
// (fft.rs:182-200), one row per workgroup: the real part is the x force,
// the imaginary part the y force.  Also packs {fx, fy, wall_bmap, wall_emap}
// per cell for grav_apply, and zeroes the mass grid for the tick's deposit
// when asked to.
__global__ __launch_bounds__(FFT_THREADS) void grav_rows_inv(const float2 *__restrict__ spec,
                                                             const uint32_t *__restrict__ mask,
                                                             float *__restrict__ fx, float *__restrict__ fy,
                                                             const uint32_t *__restrict__ wall,
                                                             const uint32_t *__restrict__ emap,
                                                             uint4 *__restrict__ rec, int32_t *__restrict__ mass_clear,
                                                             int32_t *__restrict__ prev_clear,
                                                             const float2 *__restrict__ tw, FftDims d) {
    extern __shared__ float2 dyn[];
    float2 *sm = dyn, *tab = dyn + d.W;
    const int row = blockIdx.x;
    load_twiddles(tab, tw, d.logW, d.tw_bits);
    for (int i = threadIdx.x; i < d.W; i += FFT_THREADS) sm[i] = spec[(size_t)row * d.W + i];
    __syncthreads();
    lds_fft(sm, tab, d.logW, 0, -1.0f);
    for (int i = threadIdx.x; i < d.xc; i += FFT_THREADS) {
        const int c = row * d.xc + i;
        const bool on = mask[c] != 0u;
        const float gx = on ? sm[i].x : 0.0f, gy = on ? sm[i].y : 0.0f;
        fx[c] = gx;
        fy[c] = gy;
        if (rec) rec[c] = make_uint4(__float_as_uint(gx), __float_as_uint(gy), wall[c], emap[c]);
        // The mass grid this tick's deposit accumulates into (the solve
        // reads the previous tick's snapshot, consumed by now, which the
        // deposit refills alongside); null for a bare solve.
        if (mass_clear) {
            mass_clear[c] = 0;
            prev_clear[c] = 0;
        }
    }
}

// exp(-2 pi i k / n) for k < len, evaluated in double and rounded once.
__global__ void grav_twiddles(float2 *__restrict__ tw, uint32_t n, uint32_t len) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= len) return;
    const double a = -2.0 * 3.14159265358979323846 * (double)k / (double)n;
    tw[k] = make_float2((float)cos(a), (float)sin(a));
}

// Kernel-field construction (once per context): kx + i ky, the x kernel in
// the real part and the y kernel in the imaginary part.
__global__ void grav_kernel_field(float2 *__restrict__ k, FftDims d, float scale) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= d.W * d.H) return;
    const int px = id % d.W, py = id / d.W;
    float dx = (float)px, dy = (float)py;
    if (px * 2 >= d.W) dx = (float)px - (float)d.W;
    if (py * 2 >= d.H) dy = (float)py - (float)d.H;
    float vx = 0.0f, vy = 0.0f;
    if (dx != 0.0f || dy != 0.0f) {
        const float dist = sqrtf(dx * dx + dy * dy);
        const float inv = scale / (dist * dist * dist);
        vx = dx * inv;
        vy = dy * inv;
    }
    k[id] = make_float2(vx, vy);
}

__global__ __launch_bounds__(FFT_THREADS) void grav_rows_inplace(float2 *__restrict__ data,
                                                                 const float2 *__restrict__ tw, FftDims d) {
    extern __shared__ float2 dyn[];
    float2 *sm = dyn, *tab = dyn + d.W;
    const size_t base = (size_t)blockIdx.x * d.W;
    load_twiddles(tab, tw, d.logW, d.tw_bits);
    for (int i = threadIdx.x; i < d.W; i += FFT_THREADS) sm[i] = data[base + i];
    __syncthreads();
    lds_fft(sm, tab, d.logW, 0, 1.0f);
    for (int i = threadIdx.x; i < d.W; i += FFT_THREADS) data[base + i] = sm[i];
}

__global__ __launch_bounds__(FFT_THREADS) void grav_cols_inplace(float2 *__restrict__ data,
                                                                 const float2 *__restrict__ tw, FftDims d) {
    extern __shared__ float2 dyn[];
    const int logc = d.logcols, cols = 1 << logc, c0 = col_tile(d) * cols, total = d.H << logc;
    float2 *sm = dyn, *tab = dyn + total;
    load_twiddles(tab, tw, d.logH, d.tw_bits);
    for (int e = threadIdx.x; e < total; e += FFT_THREADS)
        sm[e] = data[(size_t)(e >> logc) * d.W + c0 + (e & (cols - 1))];
    __syncthreads();
    lds_fft(sm, tab, d.logH, logc, 1.0f);
    for (int e = threadIdx.x; e < total; e += FFT_THREADS)
        data[(size_t)(e >> logc) * d.W + c0 + (e & (cols - 1))] = sm[e];
}

// apply_gravity_kernel (gravity_kernel.rs:434-507), NewtonianGravity = 1.
// rec: the per-cell {fx, fy, wall_bmap, wall_emap} grav_rows_inv packed.
// The force step of step_gravity_gpu for one slot (gravity_kernel.rs).
__device__ __forceinline__ void apply_one(uint32_t i, uint32_t ty, float x, float y, float *__restrict__ vx_a,
                                          float *__restrict__ vy_a, const uint32_t *__restrict__ flags_a,
                                          const uint4 *__restrict__ rec, const uint32_t *__restrict__ lut, int xres,
                                          int yres, int xc) {
    if (ty >= (uint32_t)TPT_PT_NUM) return;
    const uint32_t flags = flags_a[i];
    if (flags & 0x10u) return;
    const int xi = tpt_f2i(x + 0.5f), yi = tpt_f2i(y + 0.5f);
    if (xi < 4 || yi < 4 || xi >= xres - 4 || yi >= yres - 4) return;
    const int cell = (yi / TPT_CELL) * xc + xi / TPT_CELL;
    const uint4 r = rec[cell];
    if (r.z == 18u && (r.w & 0xFFu) < 8u) return;
    const uint32_t props = lut[ty * 12u];
    const float loss = __uint_as_float(lut[ty * 12u + 7u]);
    float vx = vx_a[i], vy = vy_a[i];
    if (!(ty == 78u && (flags & 0x08u))) {  // PT_SPNG with FLAG_MOVABLE keeps its velocity
        vx *= loss;
        vy *= loss;
    }
    if ((props & 0x04u) == 0u) {  // TYPE_SOLID receives no force
        vx += 1.0f * __uint_as_float(r.x);
        vy += 1.0f * __uint_as_float(r.y);
    }
    vx_a[i] = vx;
    vy_a[i] = vy;
}

// build_gravity_mass_kernel (gravity_kernel.rs:275-362) for one slot.  mode
// 1 deposits the element weight of every live particle.  `prev`, when set,
// receives the same deposit (the snapshot the next tick's solve reads).
__device__ __forceinline__ void mass_one(uint32_t i, uint32_t ty, float x, float y, const int32_t *__restrict__ life_a,
                                         const int32_t *__restrict__ tmp_a, const float *__restrict__ temp_a,
                                         const uint32_t *__restrict__ mask, const uint32_t *__restrict__ props,
                                         int32_t *__restrict__ mass, int32_t *__restrict__ prev, int xres, int yres,
                                         int xc, int yc, uint32_t mode) {
    float m = 0.0f;
    bool contributes = false;
    if (mode == 1u) {
        if (ty < (uint32_t)TPT_PT_NUM) {
            m = (float)(int32_t)props[ty * 32u + 2u];
            contributes = true;
        }
    } else if (ty == 39u) {  // GRVT
        m = 0.2f * (float)min(max(tmp_a[i], -100), 100);
        contributes = true;
    } else if (ty == 40u || ty == 41u) {  // NBHL / NWHL
        float hole = 0.1f;
        const int32_t tmp = tmp_a[i];
        if (tmp != 0) hole = tpt_clampf(0.001f * (float)tmp, 0.1f, 51.2f);
        if (ty == 41u) hole = -hole;
        m = hole;
        contributes = true;
    } else if (ty == 42u) {  // GBMB
        const int32_t life = life_a[i];
        if (life > 20) {
            m = 20.0f;
            contributes = true;
        } else if (life >= 1) {
            m = -80.0f;
            contributes = true;
        }
    } else if (ty == 43u) {  // GPMP
        if (life_a[i] == 10) {
            m = 0.2f * (temp_a[i] - 273.15f);
            contributes = true;
        }
    }
    if (!contributes) return;
    const int xi = tpt_f2i(x + 0.5f), yi = tpt_f2i(y + 0.5f);
    if (xi < 0 || yi < 0 || xi >= xres || yi >= yres) return;
    const int cx = xi / TPT_CELL, cy = yi / TPT_CELL;
    if (cx >= xc || cy >= yc) return;
    const int cell = cy * xc + cx;
    if (mask[cell] == 0u) return;
    float scaled = m * 65536.0f;
    if (scaled >= 0.0f) scaled += 0.5f;
    else scaled -= 0.5f;
    const int32_t q = tpt_f2i(scaled);
    if (q != 0) {
        atomicAdd(&mass[cell], q);
        if (prev) atomicAdd(&prev[cell], q);
    }
}

// The force step and the next mass deposit in one pass over the slots: they
// read the same type and position and write disjoint arrays.
struct MassArgs {
    const int32_t *life, *tmp;
    const float *temp;
    const uint32_t *mask, *props;
    int32_t *mass, *prev;
    int yc;
    uint32_t mode;
};

__global__ __launch_bounds__(GRAV_BLOCK) void grav_apply_mass(const uint32_t *__restrict__ ty_a,
                                                              const float *__restrict__ px_a,
                                                              const float *__restrict__ py_a, float *__restrict__ vx_a,
                                                              float *__restrict__ vy_a,
                                                              const uint32_t *__restrict__ flags_a,
                                                              const uint4 *__restrict__ rec,
                                                              const uint32_t *__restrict__ lut, MassArgs m,
                                                              uint32_t slots, int xres, int yres, int xc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint32_t ty = ty_a[i];
    if (ty == 0u) return;
    const float x = px_a[i], y = py_a[i];
    apply_one(i, ty, x, y, vx_a, vy_a, flags_a, rec, lut, xres, yres, xc);
    mass_one(i, ty, x, y, m.life, m.tmp, m.temp, m.mask, m.props, m.mass, m.prev, xres, yres, xc, m.yc, m.mode);
}

// seed_gravity_tool_mass_kernel (gravity_kernel.rs:241-262): a painted
// well replaces the snapshot's cell with its rounded fixed-point mass and is
// consumed; cells without a well keep the particle mass.
__global__ void grav_tool_seed(int32_t *__restrict__ prev, float *__restrict__ tool, uint32_t cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const float t = tool[c];
    if (t != 0.0f) {
        float scaled = t * 65536.0f;
        if (scaled >= 0.0f) scaled += 0.5f;
        else scaled -= 0.5f;
        prev[c] = tpt_f2i(scaled);
        tool[c] = 0.0f;
    }
}

// apply_grav_mass_override_kernel (gravity_kernel.rs:379-391): a flagged
// cell's assigned mass replaces what the deposit built.  `prev` is the
// snapshot when the deposit wrote both grids (it is then the copy the
// reference makes next).
__global__ void grav_override_apply(int32_t *__restrict__ mass, int32_t *__restrict__ prev,
                                    const int32_t *__restrict__ ov, const uint32_t *__restrict__ set,
                                    uint32_t cells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells || set[c] == 0u) return;
    const int32_t v = ov[c];
    mass[c] = v;
    if (prev) prev[c] = v;
}

__global__ void fill_u32(uint32_t *__restrict__ p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void copy_u32(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

uint32_t npow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

int ilog2(uint32_t v) {
    int b = 0;
    while ((1u << b) < v) ++b;
    return b;
}

FftDims dims_of(const tpt_ctx *ctx) {
    FftDims d;
    d.xc = (int)ctx->xcells;
    d.yc = (int)ctx->ycells;
    d.W = (int)ctx->fft.width;
    d.H = (int)ctx->fft.height;
    d.logW = ilog2(ctx->fft.width);
    d.logH = ilog2(ctx->fft.height);
    d.tw_bits = ilog2(ctx->fft.width > ctx->fft.height ? ctx->fft.width : ctx->fft.height);
    // One column per workgroup: W workgroups (1024 at the largest canvas)
    // keep four or more in flight per CU, which the latency of the LDS
    // stages needs more than wider row segments.
    d.logcols = 0;
    return d;
}

// The twiddle table and the kernel spectrum of the padded grid `d`: the
// Green's function samples (gravity_gpu/mod.rs:170-200), x kernel + i y
// kernel, transformed in place.  The inverse transform is unnormalised; the
// kernel carries 1/(W*H) (gravity_gpu/mod.rs:176-180).
void build_spectra(float2 *tw, float2 *kc, const FftDims &d, uint32_t ntw, uint32_t tw_len, hipStream_t stream) {
    hipLaunchKernelGGL(grav_twiddles, dim3((tw_len + 255) / 256), dim3(256), 0, stream, tw, ntw, tw_len);
    const size_t grid = (size_t)d.W * d.H;
    const float scale = -M_GRAV / (float)(d.W * d.H);
    hipLaunchKernelGGL(grav_kernel_field, dim3((grid + 255) / 256), dim3(256), 0, stream, kc, d, scale);
    hipLaunchKernelGGL(grav_rows_inplace, dim3(d.H), dim3(FFT_THREADS), rows_lds(d, 1), stream, kc, tw, d);
    hipLaunchKernelGGL(grav_cols_inplace, dim3(d.W >> d.logcols), dim3(FFT_THREADS), cols_lds(d), stream, kc, tw,
                       d);
}

// The three solve launches: padded forward rows, columns with the kernel
// product and inverse columns, inverse rows with the force extraction.
// `prof` (nullable) receives the pass timings.
void launch_solve(const int32_t *mass, const uint32_t *mask, float *fx, float *fy, const uint32_t *wall,
                  const uint32_t *emap, uint4 *rec, int32_t *mass_clear, int32_t *prev_clear, float2 *work,
                  float2 *spec, const float2 *kc, const float2 *tw, const FftDims &d, hipStream_t stream,
                  tpt_ctx *prof) {
    if (prof) tpt_pass_begin(prof, PASS_GRAV_FFT_ROWS);
    hipLaunchKernelGGL(grav_rows_fwd, dim3(d.yc), dim3(FFT_THREADS), rows_lds(d, 1), stream, mass, mask, work, tw,
                       d);
    if (prof) tpt_pass_end(prof, PASS_GRAV_FFT_ROWS);
    if (prof) tpt_pass_begin(prof, PASS_GRAV_FFT_COLS);
    hipLaunchKernelGGL(grav_cols, dim3(d.W >> d.logcols), dim3(FFT_THREADS), cols_lds(d), stream, work, kc, spec,
                       tw, d, d.yc, d.yc);
    if (prof) tpt_pass_end(prof, PASS_GRAV_FFT_COLS);
    if (prof) tpt_pass_begin(prof, PASS_GRAV_FFT_INV);
    hipLaunchKernelGGL(grav_rows_inv, dim3(d.yc), dim3(FFT_THREADS), rows_lds(d, 1), stream, spec, mask, fx, fy,
                       wall, emap, rec, mass_clear, prev_clear, tw, d);
    if (prof) tpt_pass_end(prof, PASS_GRAV_FFT_INV);
}

int ensure_fft(tpt_ctx *ctx) {
    tpt_fft_state &f = ctx->fft;
    if (f.kernels_ready) return TPT_OK;
    f.xcells = ctx->xcells;
    f.ycells = ctx->ycells;
    f.width = npow2(ctx->xcells * 2);
    f.height = npow2(ctx->ycells * 2);
    const size_t grid = (size_t)f.width * f.height;
    const size_t rows_grid = (size_t)f.width * ctx->ycells;
    const uint32_t ntw = f.width > f.height ? f.width : f.height;
    const uint32_t tw_len = ntw / 2 ? ntw / 2 : 1;
    if (hipMalloc(&f.work, rows_grid * sizeof(float2)) != hipSuccess ||
        hipMalloc(&f.spec, rows_grid * sizeof(float2)) != hipSuccess ||
        hipMalloc(&f.kernel, grid * sizeof(float2)) != hipSuccess ||
        hipMalloc(&f.twiddle, tw_len * sizeof(float2)) != hipSuccess)
        return tpt_set_error(TPT_ENOMEM, "gravity FFT buffers (%ux%u) could not be allocated", f.width, f.height);
    build_spectra(f.twiddle, f.kernel, dims_of(ctx), ntw, tw_len, ctx->stream);
    int rc = tpt_check_launch("gravity kernel spectra");
    if (rc) return rc;
    f.kernels_ready = true;
    return TPT_OK;
}

int solve(tpt_ctx *ctx, bool clear_mass) {
    int rc = ensure_fft(ctx);
    if (rc) return rc;
    const tpt_fft_state &f = ctx->fft;
    const FftDims d = dims_of(ctx);
    launch_solve(tpt_buf<int32_t>(ctx, TPT_BUF_GRAV_MASS_PREV), tpt_buf<uint32_t>(ctx, TPT_BUF_GRAV_MASK),
                 tpt_buf<float>(ctx, TPT_BUF_GRAV_FX), tpt_buf<float>(ctx, TPT_BUF_GRAV_FY),
                 tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP), tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP), ctx->cell_rec,
                 clear_mass ? tpt_buf<int32_t>(ctx, TPT_BUF_GRAV_MASS) : nullptr,
                 tpt_buf<int32_t>(ctx, TPT_BUF_GRAV_MASS_PREV), f.work, f.spec, f.kernel,
                 f.twiddle, d, ctx->stream, ctx);
    return tpt_check_launch("gravity solve");
}

}  // namespace

int tpt_gravity_solve_only(tpt_ctx *ctx) { return solve(ctx, false); }

int tpt_gravity_step(tpt_ctx *ctx, uint32_t) {
    const uint32_t cells = ctx->cells;
    const dim3 cgrid((cells + 255) / 256), blk(256);
    if (!ctx->cfg.newton_grav) {
        // SimBuffers::zero_gravity_forces (gpu_sync.rs:358-372).
        hipLaunchKernelGGL(fill_u32, cgrid, blk, 0, ctx->stream, tpt_buf<uint32_t>(ctx, TPT_BUF_GRAV_FX), 0u, cells);
        hipLaunchKernelGGL(fill_u32, cgrid, blk, 0, ctx->stream, tpt_buf<uint32_t>(ctx, TPT_BUF_GRAV_FY), 0u, cells);
        return tpt_check_launch("gravity force clear");
    }
    if (ctx->tool_armed) {
        // Step 0 of step_gravity_gpu (gravity_kernel.rs:895-913): the wells
        // join the snapshot this solve reads.
        tpt_pass_begin(ctx, PASS_GRAV_MASS);
        hipLaunchKernelGGL(grav_tool_seed, cgrid, blk, 0, ctx->stream, tpt_buf<int32_t>(ctx, TPT_BUF_GRAV_MASS_PREV),
                           tpt_buf<float>(ctx, TPT_BUF_GRAV_TOOL_MASS), cells);
        tpt_pass_end(ctx, PASS_GRAV_MASS);
        ctx->tool_armed = false;
    }
    int rc = solve(ctx, true);
    if (rc) return rc;
    const uint32_t slots = ctx->slot_bound;
    const dim3 sgrid((slots + GRAV_BLOCK - 1) / GRAV_BLOCK);
    // The force step and the deposit of this tick's mass in one launch.  In
    // the reference's mass mode the deposits are rare and go to both grids,
    // which the solve has just cleared; with every particle depositing
    // (mode 1) the snapshot is one copy instead of twice the atomics.
    const bool dual = ctx->cfg.grav_mass_mode == 0u;
    int32_t *mass = tpt_buf<int32_t>(ctx, TPT_BUF_GRAV_MASS);
    const MassArgs m{tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE), tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP),
                     tpt_buf<float>(ctx, TPT_BUF_PART_TEMP),   tpt_buf<uint32_t>(ctx, TPT_BUF_GRAV_MASK),
                     tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS), mass,
                     dual ? tpt_buf<int32_t>(ctx, TPT_BUF_GRAV_MASS_PREV) : nullptr, (int)ctx->ycells,
                     ctx->cfg.grav_mass_mode};
    tpt_pass_begin(ctx, PASS_GRAV_APPLY);
    hipLaunchKernelGGL(grav_apply_mass, sgrid, dim3(GRAV_BLOCK), 0, ctx->stream,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<float>(ctx, TPT_BUF_PART_X),
                       tpt_buf<float>(ctx, TPT_BUF_PART_Y), tpt_buf<float>(ctx, TPT_BUF_PART_VX),
                       tpt_buf<float>(ctx, TPT_BUF_PART_VY), tpt_buf<uint32_t>(ctx, TPT_BUF_PART_FLAGS),
                       ctx->cell_rec, tpt_buf<uint32_t>(ctx, TPT_BUF_MOVE_LUT), m, slots, (int)ctx->cfg.xres,
                       (int)ctx->cfg.yres, (int)ctx->xcells);
    tpt_pass_end(ctx, PASS_GRAV_APPLY);
    if (ctx->override_armed) {
        // gravity_kernel.rs:987-999, between the build and the snapshot copy.
        tpt_pass_begin(ctx, PASS_GRAV_BUILD);
        hipLaunchKernelGGL(grav_override_apply, cgrid, blk, 0, ctx->stream, mass,
                           dual ? tpt_buf<int32_t>(ctx, TPT_BUF_GRAV_MASS_PREV) : nullptr,
                           tpt_buf<int32_t>(ctx, TPT_BUF_GRAV_MASS_OVERRIDE),
                           tpt_buf<uint32_t>(ctx, TPT_BUF_GRAV_MASS_OVERRIDE_SET), cells);
        tpt_pass_end(ctx, PASS_GRAV_BUILD);
    }
    if (!dual) {
        // prev = current: the snapshot the next tick's solve reads.
        tpt_pass_begin(ctx, PASS_GRAV_BUILD);
        hipLaunchKernelGGL(copy_u32, cgrid, blk, 0, ctx->stream, (const uint32_t *)mass,
                           tpt_buf<uint32_t>(ctx, TPT_BUF_GRAV_MASS_PREV), cells);
        tpt_pass_end(ctx, PASS_GRAV_BUILD);
    }
    return tpt_check_launch("gravity apply/build");
}

int tpt_gravity_override_clear(tpt_ctx *ctx) {
    if (!ctx->override_armed) return TPT_OK;
    const uint32_t cells = ctx->cells;
    const dim3 cgrid((cells + 255) / 256), blk(256);
    hipLaunchKernelGGL(fill_u32, cgrid, blk, 0, ctx->stream, tpt_buf<uint32_t>(ctx, TPT_BUF_GRAV_MASS_OVERRIDE), 0u,
                       cells);
    hipLaunchKernelGGL(fill_u32, cgrid, blk, 0, ctx->stream, tpt_buf<uint32_t>(ctx, TPT_BUF_GRAV_MASS_OVERRIDE_SET),
                       0u, cells);
    ctx->override_armed = false;
    return tpt_check_launch("gravity override clear");
}

int tpt_gravity_free(tpt_ctx *ctx) {
    tpt_fft_state &f = ctx->fft;
    for (float2 *p : {f.work, f.spec, f.kernel, f.twiddle})
        if (p) (void)hipFree(p);
    f = tpt_fft_state();
    return TPT_OK;
}

// ---------------------------------------------------------------------------
// tpt_gravity_solve_noalloc: the gravity solve for callers that cannot
// allocate.  Its FFT buffers live in `scratch`, a static device array each
// GPU holds for the life of the process; the padded grid may be at most
// NOALLOC_MAX_SIDE on each side (canvases up to 2048x2048 px).  Layout:
// twiddles | kernel spectrum | work | spectrum.  Every launch writes each scratch element it later reads, so
// nothing has to start at zero; calls that overlap on one GPU race.

namespace {

constexpr uint32_t NOALLOC_MAX_SIDE = 1024;
constexpr size_t NOALLOC_TW = NOALLOC_MAX_SIDE / 2;
constexpr size_t NOALLOC_GRID = (size_t)NOALLOC_MAX_SIDE * NOALLOC_MAX_SIDE;
constexpr size_t NOALLOC_ROWS = (size_t)NOALLOC_MAX_SIDE * (NOALLOC_MAX_SIDE / 2);  // W x ycells
constexpr size_t NOALLOC_ELEMS = NOALLOC_TW + NOALLOC_GRID + 2 * NOALLOC_ROWS;  // 16.8 MB of float2

}  // namespace

static __device__ __align__(256) float2 scratch[NOALLOC_ELEMS];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float2 *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of the current GPU's copy of `scratch`, looked up once
// per device (host code must never pass the symbol itself to a kernel).
int scratch_for_current_device(float2 **out) {
    int dev = 0;
    TPT_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return tpt_set_error(TPT_EDEVICE, "device %d beyond the %d the scratch table covers", dev,
                             NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return tpt_set_error(TPT_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                                 hipGetErrorString(e));
        g_scratch_ptr[dev] = (float2 *)p;
    }
    *out = g_scratch_ptr[dev];
    return TPT_OK;
}

}  // namespace

extern "C" int tpt_gravity_solve_noalloc(const int32_t *mass, const uint32_t *mask, float *fx, float *fy,
                                         uint32_t xcells, uint32_t ycells, void *stream) {
    if (!mass || !mask || !fx || !fy) return tpt_set_error(TPT_EINVAL, "null grid pointer");
    if (xcells == 0 || ycells == 0) return tpt_set_error(TPT_EINVAL, "empty cell grid");
    const uint32_t W = npow2(xcells * 2), H = npow2(ycells * 2);
    if (W > NOALLOC_MAX_SIDE || H > NOALLOC_MAX_SIDE)
        return tpt_set_error(TPT_EINVAL,
                             "cell grid %ux%u needs a %ux%u FFT grid; the allocation-free solve holds at most "
                             "%ux%u (%zu bytes of scratch)",
                             xcells, ycells, W, H, NOALLOC_MAX_SIDE, NOALLOC_MAX_SIDE, NOALLOC_ELEMS * sizeof(float2));
    float2 *base = nullptr;
    int rc = scratch_for_current_device(&base);
    if (rc) return rc;
    FftDims d;
    d.xc = (int)xcells;
    d.yc = (int)ycells;
    d.W = (int)W;
    d.H = (int)H;
    d.logW = ilog2(W);
    d.logH = ilog2(H);
    const uint32_t ntw = W > H ? W : H;
    d.tw_bits = ilog2(ntw);
    d.logcols = 0;
    const uint32_t tw_len = ntw / 2 ? ntw / 2 : 1;
    const size_t grid = (size_t)W * H, rows = (size_t)W * ycells;
    float2 *tw = base, *kc = tw + NOALLOC_TW, *work = kc + grid, *spec = work + rows;
    hipStream_t st = (hipStream_t)stream;
    build_spectra(tw, kc, d, ntw, tw_len, st);
    launch_solve(mass, mask, fx, fy, nullptr, nullptr, nullptr, nullptr, nullptr, work, spec, kc, tw, d, st,
                 nullptr);
    return tpt_check_launch("allocation-free gravity solve");
}
