/*
 * Oracle: Newtonian gravity -- TEST INFRASTRUCTURE ONLY (see tpt_oracle.h).
 *
 * Restates crates/tpt-sim/src/gravity_gpu/{mod,fft}.rs (the padded radix-2
 * FFT convolution) and crates/tpt-sim/src/gravity_kernel.rs (mass deposit,
 * Loss + force application, the two-grid mass cadence and the host flood fill
 * of the gravity mask).
 *
 * Two solvers:
 *   oracle_gravity_solve      float32, the reference's own arithmetic: bit
 *                             reversal + log2(n) radix-2 DIT passes per axis
 *                             with the twiddle table of fft.rs:242-252
 *                             (cosf/sinf of a float32 angle);
 *   oracle_gravity_solve_f64  the same convolution in double precision with
 *                             double twiddles: the accuracy reference the
 *                             HIP FFT is held to.
 */
#include "tpt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_util.h"

#define M_GRAV_F 6.673e-1f
#define MASS_SCALE 65536.0f

static uint32_t npow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

/* fft.rs:255-263 */
static uint32_t bit_count(uint32_t n) {
    uint32_t bits = 0, v = (n ? n : 1) - 1;
    while (v) {
        v >>= 1;
        bits++;
    }
    return bits;
}

void oracle_fft_twiddle_f32(uint32_t max_len, float *re, float *im) {
    uint32_t half = max_len / 2 ? max_len / 2 : 1;
    const float pi = 3.14159265358979323846f;
    for (uint32_t k = 0; k < half; ++k) {
        float angle = -2.0f * pi * (float)k / (float)max_len;
        re[k] = cosf(angle);
        im[k] = sinf(angle);
    }
}

/* One 1-D transform over `rows` sequences (fft.rs:35-122, mod.rs:285-335),
 * data interleaved re/im. */
void oracle_fft_transform_f32(float *data, uint32_t rows, uint32_t len, uint32_t inner, uint32_t stride,
                              const float *tw_re, const float *tw_im, uint32_t tw_bits, float sign) {
    uint32_t bits = bit_count(len);
#pragma omp parallel for schedule(static)
    for (uint32_t r = 0; r < rows; ++r)
        for (uint32_t i = 0; i < len; ++i) {
            uint32_t src = i, rev = 0;
            for (uint32_t b = 0; b < bits; ++b) {
                rev = rev * 2 + src % 2;
                src /= 2;
            }
            if (i < rev) {
                size_t a = ((size_t)r * inner + (size_t)i * stride) * 2, c = ((size_t)r * inner + (size_t)rev * stride) * 2;
                float t0 = data[a], t1 = data[a + 1];
                data[a] = data[c];
                data[a + 1] = data[c + 1];
                data[c] = t0;
                data[c + 1] = t1;
            }
        }
    for (uint32_t half = 1; half < len; half *= 2) {
        uint32_t shift = tw_bits - bit_count(half * 2);
#pragma omp parallel for schedule(static)
        for (uint32_t r = 0; r < rows; ++r)
            for (uint32_t k = 0; k < len / 2; ++k) {
                uint32_t l2 = half * 2, block = k / half, pos = k % half;
                uint32_t i0 = block * l2 + pos, i1 = i0 + half;
                uint32_t t = pos << shift;
                float wr = tw_re[t], wi = tw_im[t] * sign;
                size_t a = ((size_t)r * inner + (size_t)i0 * stride) * 2, b = ((size_t)r * inner + (size_t)i1 * stride) * 2;
                float xr = data[a], xi = data[a + 1], yr = data[b], yi = data[b + 1];
                float tr = wr * yr - wi * yi;
                float ti = wr * yi + wi * yr;
                data[a] = xr + tr;
                data[a + 1] = xi + ti;
                data[b] = xr - tr;
                data[b + 1] = xi - ti;
            }
    }
}

static void transform_f64(double *data, uint32_t rows, uint32_t len, uint32_t inner, uint32_t stride,
                          const double *tw_re, const double *tw_im, uint32_t tw_bits, double sign) {
    uint32_t bits = bit_count(len);
    for (uint32_t r = 0; r < rows; ++r)
        for (uint32_t i = 0; i < len; ++i) {
            uint32_t src = i, rev = 0;
            for (uint32_t b = 0; b < bits; ++b) {
                rev = rev * 2 + src % 2;
                src /= 2;
            }
            if (i < rev) {
                size_t a = ((size_t)r * inner + (size_t)i * stride) * 2, c = ((size_t)r * inner + (size_t)rev * stride) * 2;
                double t0 = data[a], t1 = data[a + 1];
                data[a] = data[c];
                data[a + 1] = data[c + 1];
                data[c] = t0;
                data[c + 1] = t1;
            }
        }
    for (uint32_t half = 1; half < len; half *= 2) {
        uint32_t shift = tw_bits - bit_count(half * 2);
        for (uint32_t r = 0; r < rows; ++r)
            for (uint32_t k = 0; k < len / 2; ++k) {
                uint32_t l2 = half * 2, block = k / half, pos = k % half;
                uint32_t i0 = block * l2 + pos, i1 = i0 + half;
                uint32_t t = pos << shift;
                double wr = tw_re[t], wi = tw_im[t] * sign;
                size_t a = ((size_t)r * inner + (size_t)i0 * stride) * 2, b = ((size_t)r * inner + (size_t)i1 * stride) * 2;
                double xr = data[a], xi = data[a + 1], yr = data[b], yi = data[b + 1];
                double tr = wr * yr - wi * yi, ti = wr * yi + wi * yr;
                data[a] = xr + tr;
                data[a + 1] = xi + ti;
                data[b] = xr - tr;
                data[b + 1] = xi - ti;
            }
    }
}

/* GravityFft::solve (gravity_gpu/mod.rs:114-278) in float32. */
void oracle_gravity_solve(oracle_state *s) {
    const uint32_t xc = s->xres / 4, yc = s->yres / 4;
    const uint32_t w = npow2(xc * 2), h = npow2(yc * 2), cells = w * h;
    const uint32_t tw_bits = bit_count(w > h ? w : h), tw_len = (1u << tw_bits) / 2 ? (1u << tw_bits) / 2 : 1;
    float *tr = malloc(sizeof(float) * tw_len), *ti = malloc(sizeof(float) * tw_len);
    oracle_fft_twiddle_f32(1u << tw_bits, tr, ti);
    /* The kernel spectra depend on the grid only; like GravityFft (mod.rs:126-133)
     * they are built once per grid size and reused. */
    static float *kx = NULL, *ky = NULL;
    static uint32_t kw = 0, kh = 0;
    float *work = calloc((size_t)cells * 2, sizeof(float));
    float *sx = malloc(sizeof(float) * cells * 2), *sy = malloc(sizeof(float) * cells * 2);
    if (kx == NULL || kw != w || kh != h) {
    free(kx);
    free(ky);
    kx = calloc((size_t)cells * 2, sizeof(float));
    ky = calloc((size_t)cells * 2, sizeof(float));
    kw = w;
    kh = h;
    /* build_kernel_field_kernel (fft.rs:209-236), scale of mod.rs:180 */
    const float scale = -M_GRAV_F / (float)(w * h);
    for (uint32_t id = 0; id < cells; ++id) {
        uint32_t px = id % w, py = id / w;
        float dx = (float)px, dy = (float)py;
        if (px * 2 >= w) dx = (float)px - (float)w;
        if (py * 2 >= h) dy = (float)py - (float)h;
        float vx = 0.0f, vy = 0.0f;
        if (dx != 0.0f || dy != 0.0f) {
            float d = sqrtf(dx * dx + dy * dy);
            float inv = scale / (d * d * d);
            vx = dx * inv;
            vy = dy * inv;
        }
        kx[id * 2] = vx;
        ky[id * 2] = vy;
    }
    oracle_fft_transform_f32(kx, h, w, w, 1, tr, ti, tw_bits, 1.0f);
    oracle_fft_transform_f32(kx, w, h, 1, w, tr, ti, tw_bits, 1.0f);
    oracle_fft_transform_f32(ky, h, w, w, 1, tr, ti, tw_bits, 1.0f);
    oracle_fft_transform_f32(ky, w, h, 1, w, tr, ti, tw_bits, 1.0f);
    }
    /* build_padded_mass_kernel (fft.rs:153-178) */
    const float inv_scale = 1.0f / MASS_SCALE;
    for (uint32_t py = 0; py < yc; ++py)
        for (uint32_t px = 0; px < xc; ++px) {
            uint32_t cell = py * xc + px;
            if (s->grav_mask[cell] != 0) work[((size_t)py * w + px) * 2] = (float)s->grav_mass_prev[cell] * inv_scale;
        }
    oracle_fft_transform_f32(work, h, w, w, 1, tr, ti, tw_bits, 1.0f);
    oracle_fft_transform_f32(work, w, h, 1, w, tr, ti, tw_bits, 1.0f);
    /* complex_mul_kernel (fft.rs:129-140) */
    for (uint32_t i = 0; i < cells; ++i) {
        float wr = work[2 * i], wi = work[2 * i + 1];
        sx[2 * i] = wr * kx[2 * i] - wi * kx[2 * i + 1];
        sx[2 * i + 1] = wr * kx[2 * i + 1] + wi * kx[2 * i];
        sy[2 * i] = wr * ky[2 * i] - wi * ky[2 * i + 1];
        sy[2 * i + 1] = wr * ky[2 * i + 1] + wi * ky[2 * i];
    }
    oracle_fft_transform_f32(sx, h, w, w, 1, tr, ti, tw_bits, -1.0f);
    oracle_fft_transform_f32(sx, w, h, 1, w, tr, ti, tw_bits, -1.0f);
    oracle_fft_transform_f32(sy, h, w, w, 1, tr, ti, tw_bits, -1.0f);
    oracle_fft_transform_f32(sy, w, h, 1, w, tr, ti, tw_bits, -1.0f);
    /* extract_force_kernel (fft.rs:182-200) */
    for (uint32_t py = 0; py < yc; ++py)
        for (uint32_t px = 0; px < xc; ++px) {
            uint32_t id = py * xc + px;
            int on = s->grav_mask[id] != 0;
            s->grav_fx[id] = on ? sx[((size_t)py * w + px) * 2] : 0.0f;
            s->grav_fy[id] = on ? sy[((size_t)py * w + px) * 2] : 0.0f;
        }
    free(tr); free(ti); free(work); free(sx); free(sy);
}

/* The same convolution in double precision; writes fx/fy (xcells*ycells). */
void oracle_gravity_solve_f64(const oracle_state *s, double *fx, double *fy) {
    const uint32_t xc = s->xres / 4, yc = s->yres / 4;
    const uint32_t w = npow2(xc * 2), h = npow2(yc * 2), cells = w * h;
    const uint32_t tw_bits = bit_count(w > h ? w : h), n_tw = 1u << tw_bits;
    const uint32_t tw_len = n_tw / 2 ? n_tw / 2 : 1;
    double *tr = malloc(sizeof(double) * tw_len), *ti = malloc(sizeof(double) * tw_len);
    for (uint32_t k = 0; k < tw_len; ++k) {
        double a = -2.0 * 3.14159265358979323846 * (double)k / (double)n_tw;
        tr[k] = cos(a);
        ti[k] = sin(a);
    }
    double *kx = calloc((size_t)cells * 2, sizeof(double)), *ky = calloc((size_t)cells * 2, sizeof(double));
    double *work = calloc((size_t)cells * 2, sizeof(double));
    const double scale = -(double)M_GRAV_F / (double)(w * h);
    for (uint32_t id = 0; id < cells; ++id) {
        uint32_t px = id % w, py = id / w;
        double dx = px * 2 >= w ? (double)px - (double)w : (double)px;
        double dy = py * 2 >= h ? (double)py - (double)h : (double)py;
        if (dx != 0.0 || dy != 0.0) {
            double d = sqrt(dx * dx + dy * dy);
            kx[id * 2] = dx * scale / (d * d * d);
            ky[id * 2] = dy * scale / (d * d * d);
        }
    }
    transform_f64(kx, h, w, w, 1, tr, ti, tw_bits, 1.0);
    transform_f64(kx, w, h, 1, w, tr, ti, tw_bits, 1.0);
    transform_f64(ky, h, w, w, 1, tr, ti, tw_bits, 1.0);
    transform_f64(ky, w, h, 1, w, tr, ti, tw_bits, 1.0);
    for (uint32_t py = 0; py < yc; ++py)
        for (uint32_t px = 0; px < xc; ++px) {
            uint32_t cell = py * xc + px;
            if (s->grav_mask[cell] != 0) work[((size_t)py * w + px) * 2] = (double)s->grav_mass_prev[cell] / 65536.0;
        }
    transform_f64(work, h, w, w, 1, tr, ti, tw_bits, 1.0);
    transform_f64(work, w, h, 1, w, tr, ti, tw_bits, 1.0);
    for (uint32_t i = 0; i < cells; ++i) {
        double wr = work[2 * i], wi = work[2 * i + 1];
        double ar = wr * kx[2 * i] - wi * kx[2 * i + 1], ai = wr * kx[2 * i + 1] + wi * kx[2 * i];
        double br = wr * ky[2 * i] - wi * ky[2 * i + 1], bi = wr * ky[2 * i + 1] + wi * ky[2 * i];
        kx[2 * i] = ar;
        kx[2 * i + 1] = ai;
        ky[2 * i] = br;
        ky[2 * i + 1] = bi;
    }
    transform_f64(kx, h, w, w, 1, tr, ti, tw_bits, -1.0);
    transform_f64(kx, w, h, 1, w, tr, ti, tw_bits, -1.0);
    transform_f64(ky, h, w, w, 1, tr, ti, tw_bits, -1.0);
    transform_f64(ky, w, h, 1, w, tr, ti, tw_bits, -1.0);
    for (uint32_t py = 0; py < yc; ++py)
        for (uint32_t px = 0; px < xc; ++px) {
            uint32_t id = py * xc + px;
            int on = s->grav_mask[id] != 0;
            fx[id] = on ? kx[((size_t)py * w + px) * 2] : 0.0;
            fy[id] = on ? ky[((size_t)py * w + px) * 2] : 0.0;
        }
    free(tr); free(ti); free(kx); free(ky); free(work);
}

/* Fixed-point rounding of gravity_kernel.rs:336-345 (half away from zero). */
static int32_t quantise(float mass) {
    float scaled = mass * MASS_SCALE;
    if (scaled >= 0.0f) scaled += 0.5f;
    else scaled -= 0.5f;
    return ora_f2i(scaled);
}

/* build_gravity_mass_kernel (gravity_kernel.rs:275-362); grav_mass_mode 1
 * deposits the element weight of every live particle instead (extension). */
void oracle_build_gravity_mass(oracle_state *s) {
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4, yc = yres / 4;
    /* PT_GRVT=39 NBHL=40 NWHL=41 GBMB=42 GPMP=43 */
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0) continue;
        float mass = 0.0f;
        int contributes = 0;
        if (s->grav_mass_mode == 1) {
            if (ty < 256) {
                mass = (float)(int32_t)s->elem_props[ty * 32 + 2];
                contributes = 1;
            }
        } else if (ty == 39) {
            int32_t tmp = ora_clampi(s->part_tmp[i], -100, 100);
            mass = 0.2f * (float)tmp;
            contributes = 1;
        } else if (ty == 40 || ty == 41) {
            float hole = 0.1f;
            if (s->part_tmp[i] != 0) hole = ora_clampf(0.001f * (float)s->part_tmp[i], 0.1f, 51.2f);
            if (ty == 41) hole = -hole;
            mass = hole;
            contributes = 1;
        } else if (ty == 42) {
            int32_t life = s->part_life[i];
            if (life > 20) { mass = 20.0f; contributes = 1; }
            else if (life >= 1) { mass = -80.0f; contributes = 1; }
        } else if (ty == 43) {
            if (s->part_life[i] == 10) { mass = 0.2f * (s->part_temp[i] - 273.15f); contributes = 1; }
        }
        if (!contributes) continue;
        int xi = ora_f2i(s->part_x[i] + 0.5f), yi = ora_f2i(s->part_y[i] + 0.5f);
        if (xi < 0 || yi < 0 || xi >= (int)xres || yi >= (int)yres) continue;
        uint32_t cx = (uint32_t)xi / 4, cy = (uint32_t)yi / 4;
        if (cx >= xc || cy >= yc) continue;
        uint32_t cell = cy * xc + cx;
        if (s->grav_mask[cell] == 0) continue;
        int32_t q = quantise(mass);
        if (q != 0) s->grav_mass[cell] = (int32_t)((uint32_t)s->grav_mass[cell] + (uint32_t)q);
    }
}

/* apply_gravity_kernel (gravity_kernel.rs:434-507), newton_grav scale 1.0 */
void oracle_apply_gravity(oracle_state *s) {
    const uint32_t xres = s->xres, yres = s->yres, xc = xres / 4;
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < s->slot_bound; ++i) {
        uint32_t ty = s->part_ty[i];
        if (ty == 0 || ty >= 256) continue;
        uint32_t flags = s->part_flags[i];
        int eligible = (flags & 0x10u) == 0;
        int xi = ora_f2i(s->part_x[i] + 0.5f), yi = ora_f2i(s->part_y[i] + 0.5f);
        if (xi < 0 || yi < 0 || xi >= (int)xres || yi >= (int)yres) eligible = 0;
        if (xi < 4 || yi < 4 || xi >= (int)xres - 4 || yi >= (int)yres - 4) eligible = 0;
        if (!eligible) continue;
        uint32_t cell = ((uint32_t)yi / 4) * xc + (uint32_t)xi / 4;
        if (s->wall_bmap[cell] == 18u && (s->wall_emap[cell] & 0xFFu) < 8u) continue;
        uint32_t base = ty * 12;
        uint32_t props = s->move_lut[base];
        float loss = ora_bits_f(s->move_lut[base + 7]);
        if (!(ty == 78u && (flags & 0x08u) != 0)) { /* PT_SPNG with FLAG_MOVABLE keeps v */
            s->part_vx[i] *= loss;
            s->part_vy[i] *= loss;
        }
        if ((props & 0x04u) == 0) {
            s->part_vx[i] += 1.0f * s->grav_fx[cell];
            s->part_vy[i] += 1.0f * s->grav_fy[cell];
        }
    }
}

/* step_gravity_gpu (gravity_kernel.rs:826-1016). */
void oracle_step_gravity(oracle_state *s) {
    const uint32_t n = (s->xres / 4) * (s->yres / 4);
    if (!s->newton_grav) {
        memset(s->grav_fx, 0, sizeof(float) * n);
        memset(s->grav_fy, 0, sizeof(float) * n);
        return;
    }
    /* seed_gravity_tool_mass_kernel (gravity_kernel.rs:241-262, launched
     * at :903): a non-zero well replaces the snapshot cell, then is spent. */
    for (uint32_t c = 0; c < n; ++c) {
        const float tool = s->grav_tool_mass[c];
        if (tool != 0.0f) {
            float scaled = tool * 65536.0f;
            if (scaled >= 0.0f) scaled += 0.5f;
            else scaled -= 0.5f;
            s->grav_mass_prev[c] = ora_f2i(scaled);
            s->grav_tool_mass[c] = 0.0f;
        }
    }
    oracle_gravity_solve(s);
    oracle_apply_gravity(s);
    memset(s->grav_mass, 0, sizeof(int32_t) * n);
    oracle_build_gravity_mass(s);
    /* apply_grav_mass_override_kernel (gravity_kernel.rs:379-391, launched
     * at :991): an assigned cell replaces the built mass. */
    for (uint32_t c = 0; c < n; ++c)
        if (s->grav_mass_override_set[c] != 0u) s->grav_mass[c] = s->grav_mass_override[c];
    memcpy(s->grav_mass_prev, s->grav_mass, sizeof(int32_t) * n);
}

/* The modext clear at the start of the element pass
 * (reaction_kernel/modext/mod.rs:53-99): zero the value and flag grids. */
void oracle_clear_grav_override(oracle_state *s) {
    const uint32_t n = (s->xres / 4) * (s->yres / 4);
    memset(s->grav_mass_override, 0, sizeof(int32_t) * n);
    memset(s->grav_mass_override_set, 0, sizeof(uint32_t) * n);
}

/* rebuild_gravity_mask (gravity_kernel.rs:658-710): flood fill from the
 * border through 4-neighbours, not crossing WL_GRAV (14). */
void oracle_gravity_mask(const uint32_t *wall, uint32_t xc, uint32_t yc, uint32_t *mask) {
    size_t cells = (size_t)xc * yc;
    memset(mask, 0, sizeof(uint32_t) * cells);
    uint32_t *stack = malloc(sizeof(uint32_t) * (cells + 4 * (xc + yc) + 16));
    size_t top = 0;
#define SEED(idx)                                                           \
    do {                                                                    \
        uint32_t _i = (idx);                                                \
        if ((wall[_i] & 0xFFu) != 14u && mask[_i] == 0) {                   \
            mask[_i] = 0xFFFFFFFFu;                                         \
            stack[top++] = _i;                                              \
        }                                                                   \
    } while (0)
    for (uint32_t x = 0; x < xc; ++x) {
        SEED(x);
        SEED((yc - 1) * xc + x);
    }
    for (uint32_t y = 1; y + 1 < yc; ++y) {
        SEED(y * xc);
        SEED(y * xc + xc - 1);
    }
    while (top) {
        uint32_t idx = stack[--top], x = idx % xc, y = idx / xc;
        if (x > 0) SEED(idx - 1);
        if (x + 1 < xc) SEED(idx + 1);
        if (y > 0) SEED(idx - xc);
        if (y + 1 < yc) SEED(idx + xc);
    }
#undef SEED
    free(stack);
}

/* Thread count of the OpenMP loops (the CPU baseline times 1 and all). */
#ifdef _OPENMP
#include <omp.h>
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }
#else
void oracle_set_threads(int n) { (void)n; }
int oracle_max_threads(void) { return 1; }
#endif
