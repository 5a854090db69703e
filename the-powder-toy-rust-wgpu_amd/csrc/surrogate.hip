// The learned air surrogate (no counterpart in the reference; it stands in
// for 8 calls of air_kernel::step_air_gpu, air_kernel/mod.rs:120-299).
//
// A 3-level convolutional U-Net on the air grid, 32/64/128 channels, 3x3
// convolutions with GELU, bf16 weights and activations on MFMA with fp32
// accumulation (surrogate_kernels.h).  Input: 6 channels (pressure, vx, vy,
// ambient heat, wall mask, heat-source temperature; two zero channels pad
// to 8); output: the normalised change of pressure, vx, vy and ambient heat
// over 8 plain air ticks.
//
//   E1  8 -> 32 @1      pool      E2  32 -> 64 @1/2   E2b 64 -> 64 @1/2
//   pool                E3  64 -> 128 @1/4            E3b 128 -> 128 @1/4
//   D2a 128 -> 64 @1/4  up + E2b  D2b 64 -> 64 @1/2   D1a 64 -> 32 @1/2
//   up + E1             D1b 32 -> 32 @1               H   32 -> 4 @1 (no GELU)
//
// The object also holds the training set (snapshots of plain-solver
// trajectories, every 8 ticks) and trains itself with its own forward and
// backward kernels and Adam.  A call on a context (tpt_sur_step_air) checks
// the prediction's energy and finiteness and falls back to 8 plain air
// ticks when the check fails.

#include "tpt_internal.h"
#include "surrogate_kernels.h"

#include <math.h>

#include <algorithm>
#include <new>
#include <random>
#include <vector>

using namespace sur;

namespace {

constexpr int N_LAYERS = 10;
constexpr int IN_CH = 8;    // 6 inputs + 2 zero channels
constexpr int OUT_CH = 8;   // 4 outputs + 4 zero channels
constexpr int TICKS_PER_CALL = 8;
constexpr int MAX_ROLLOUT = 8;  // calls a training pair may span (a 64-tick trajectory has 9 snapshots)

struct LayerSpec {
    int cin, cout, level, gelu;
};
constexpr LayerSpec SPEC[N_LAYERS] = {
    {IN_CH, 32, 0, 1},  // E1
    {32, 64, 1, 1},     // E2
    {64, 64, 1, 1},     // E2b
    {64, 128, 2, 1},    // E3
    {128, 128, 2, 1},   // E3b
    {128, 64, 2, 1},    // D2a
    {64, 64, 1, 1},     // D2b
    {64, 32, 1, 1},     // D1a
    {32, 32, 0, 1},     // D1b
    {32, OUT_CH, 0, 0}, // H
};

int ilog2i(int v) {
    int b = 0;
    while ((1 << b) < v) ++b;
    return b;
}
int round_up(int v, int m) { return (v + m - 1) / m * m; }

struct Layer {
    int cin, cout, level, gelu;
    int kpad, kchunks, rows;       // weight matrix rows (cout padded to 64) x kpad
    int kpad_t, kchunks_t, rows_t; // transposed-flipped matrix for the data gradient
    size_t offset;                 // into the flat parameter vector
    bf16_t *w = nullptr, *wt = nullptr;
};

}  // namespace

struct tpt_sur {
    int H = 0, W = 0;               // air grid (cells)
    Layer layer[N_LAYERS];
    size_t n_params = 0;            // flat fp32 parameter vector length
    float *wf = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    bf16_t *wb = nullptr;           // bf16 copy (per-layer views into it)
    bf16_t *wtb = nullptr;          // transposed copies
    float *part = nullptr;          // weight-gradient partial sums
    size_t part_elems = 0;
    int adam_t = 0;
    // activations (NHWC bf16)
    bf16_t *x0lo = nullptr;         // low bf16 half of the input (x0 + x0lo = the fp32 input)
    bf16_t *x0 = nullptr, *a1 = nullptr, *z1 = nullptr, *p1 = nullptr, *a2 = nullptr, *z2 = nullptr,
           *a2b = nullptr, *z2b = nullptr, *p2 = nullptr, *a3 = nullptr, *z3 = nullptr, *a3b = nullptr,
           *z3b = nullptr, *d2a = nullptr, *zd2a = nullptr, *u2 = nullptr, *d2b = nullptr, *zd2b = nullptr,
           *d1a = nullptr, *zd1a = nullptr, *u1 = nullptr, *d1b = nullptr, *zd1b = nullptr;
    float *out = nullptr;           // H x W x OUT_CH fp32
    // gradients
    bf16_t *g_out = nullptr, *g_d1b = nullptr, *g_u1 = nullptr, *g_d1a = nullptr, *g_d2b = nullptr,
           *g_u2 = nullptr, *g_d2a = nullptr, *g_a3b = nullptr, *g_a3 = nullptr, *g_p2 = nullptr,
           *g_a2b = nullptr, *g_a2 = nullptr, *g_p1 = nullptr, *g_a1 = nullptr;
    float *loss_acc = nullptr;      // device accumulators: loss, energy in, energy out, non-finite count
    float *gnorm = nullptr;         // device: sum of squares of the batch gradient
    float *red = nullptr;           // device: per-block partial sums (up to 4 per block)
    float clip = 1.0f;              // global gradient-norm clip (0: none)
    // Training horizon: 1 trains on (snapshot k -> k+1); r > 1 first
    // advances snapshot k r - 1 times by the network itself (no gradient)
    // and trains on (that prediction -> snapshot k+r), so that the network
    // learns to correct the errors of its own rollouts.
    uint32_t rollout = 1;
    float *roll = nullptr;          // the predicted planes of a rollout pair (4 x cells)
    bool fused = true;              // inference folds the glue kernels into the convolutions
    float in_scale[4] = {10.0f, 1.0f, 1.0f, 500.0f};   // input normalisation (state)
    float scale[4] = {1.0f, 1.0f, 1.0f, 10.0f};        // output normalisation (8-tick change)
    float ambient = 295.15f;
    // heat sources of the current scene: temperature per cell, 0 elsewhere
    float *src = nullptr;
    // training set: n_traj x n_snap snapshots of 4 planes (pv, vx, vy, hv),
    // plus each trajectory's wall mask and source map
    float *data = nullptr;
    uint8_t *walls = nullptr;
    float *srcs = nullptr;
    uint32_t n_traj = 0, n_snap = 0;
    // saved air state for the fallback
    float *saved = nullptr;
    float *check_host = nullptr;    // pinned host copy of the check's sums
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
};

namespace {

size_t cells_at(const tpt_sur *s, int level) { return (size_t)(s->H >> level) * (s->W >> level); }

int alloc(tpt_sur *s, void **p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 4) != hipSuccess) {
        *p = nullptr;
        return tpt_set_error(TPT_ENOMEM, "surrogate: %zu bytes could not be allocated", bytes);
    }
    s->allocs.push_back(*p);
    return TPT_OK;
}
template <typename T> int alloc_n(tpt_sur *s, T **p, size_t n) { return alloc(s, (void **)p, n * sizeof(T)); }

dim3 grid1(size_t n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

template <int NT, int KS, bool LO> void set_conv_lds(int bytes) {
    (void)hipFuncSetAttribute((const void *)conv3x3<NT, KS, LO>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}
template <int MT> void set_wgrad_lds(int bytes) {
    (void)hipFuncSetAttribute((const void *)conv3x3_wgrad<MT>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

struct Planes {
    const float *pv, *vx, *vy, *hv;
};

// x0, x0lo: the 8 input channels of every cell as two bf16 halves (high
// and the rounded remainder).
__global__ void sur_prep(Planes p, const uint8_t *__restrict__ wall8, const uint32_t *__restrict__ wall32,
                         const float *__restrict__ src, bf16_t *__restrict__ x0, bf16_t *__restrict__ x0lo, int n,
                         float sp, float sv, float sh, float amb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float wall = wall8 ? (float)(wall8[i] != 0) : (float)(wall32[i] != 0u);
    float v[8];
    prep_channels(p.pv[i], p.vx[i], p.vy[i], p.hv[i], wall, src[i], sp, sv, sh, amb, v);
    float hi[8], lo[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        hi[c] = bf2f(f2bf(v[c]));
        lo[c] = v[c] - hi[c];
    }
    *reinterpret_cast<uint4 *>(x0 + (size_t)i * IN_CH) = pack8(hi);
    *reinterpret_cast<uint4 *>(x0lo + (size_t)i * IN_CH) = pack8(lo);
}

// A conv's staging mode and its extra inputs (STAGE_* in the kernels).
struct StageIn {
    int stage = STAGE_PLAIN;
    const bf16_t *in2 = nullptr;                                       // STAGE_UPADD
    const float *pv = nullptr, *vx = nullptr, *vy = nullptr, *hv = nullptr;  // STAGE_PREP
    const float *src = nullptr;
    const uint32_t *wall = nullptr;
};

// One convolution: forward (w, bias column, optional GELU, pre-activation
// stored) or data gradient (wt, GELU derivative of zref).
void conv(const tpt_sur *s, hipStream_t st, const bf16_t *in, const bf16_t *in_lo, int cin, int cout, int level,
          const bf16_t *w, int kpad, int kchunks, bf16_t *out, bf16_t *zout, const bf16_t *zref, float *out_f32,
          int gelu, int bias_col, const StageIn &sin = StageIn()) {
    ConvArgs a{};
    a.in = in;
    a.in_lo = in_lo;
    a.in2 = sin.in2;
    a.stage = sin.stage;
    a.pv = sin.pv;
    a.vx = sin.vx;
    a.vy = sin.vy;
    a.hv = sin.hv;
    a.src = sin.src;
    a.wall = sin.wall;
    a.sp = s->in_scale[0];
    a.sv = s->in_scale[1];
    a.sh = s->in_scale[3];
    a.amb = s->ambient;
    a.w = w;
    a.out = out;
    a.zout = zout;
    a.zref = zref;
    a.out_f32 = out_f32;
    a.H = s->H >> level;
    a.W = s->W >> level;
    a.cin = cin;
    a.lcin = ilog2i(cin);
    a.cout = cout;
    a.kpad = kpad;
    a.kchunks = kchunks;
    a.gelu = gelu;
    a.bias_col = bias_col;
    // 16 * nt output channels per wave; the K chunks split over ks waves
    // while the grid holds fewer than 2048 waves (8 per CU) and each wave
    // keeps at least two chunks.
    const int nt = cout >= 64 ? 4 : cout >= 32 ? 2 : 1;
    const int segs = (a.W + CONV_SEG - 1) / CONV_SEG, ntiles = (cout + 16 * nt - 1) / (16 * nt);
    int ks = 1;
    const bool lo = in_lo != nullptr || sin.stage == STAGE_PREP;
    while (!lo && ks < 4 && segs * ((a.H + 4 / ks - 1) / (4 / ks)) * ntiles * 4 < 2048 && kchunks >= 4 * ks) ks *= 2;
    const size_t lds = conv_lds_bytes(cin, ks, nt, lo);
    const dim3 g(segs, (a.H + 4 / ks - 1) / (4 / ks), ntiles);
#define TPT_CONV_LAUNCH(NT_, KS_, LO_) \
    hipLaunchKernelGGL((conv3x3<NT_, KS_, LO_>), g, dim3(CONV_THREADS), lds, st, a)
    if (lo) {
        if (nt == 4) TPT_CONV_LAUNCH(4, 1, true);
        else if (nt == 2) TPT_CONV_LAUNCH(2, 1, true);
        else TPT_CONV_LAUNCH(1, 1, true);
    } else if (ks == 1) {
        if (nt == 4) TPT_CONV_LAUNCH(4, 1, false);
        else if (nt == 2) TPT_CONV_LAUNCH(2, 1, false);
        else TPT_CONV_LAUNCH(1, 1, false);
    } else if (ks == 2) {
        if (nt == 4) TPT_CONV_LAUNCH(4, 2, false);
        else if (nt == 2) TPT_CONV_LAUNCH(2, 2, false);
        else TPT_CONV_LAUNCH(1, 2, false);
    } else {
        if (nt == 4) TPT_CONV_LAUNCH(4, 4, false);
        else if (nt == 2) TPT_CONV_LAUNCH(2, 4, false);
        else TPT_CONV_LAUNCH(1, 4, false);
    }
#undef TPT_CONV_LAUNCH
}

// Weight gradient of layer l into s->grad (accumulated when `accumulate`).
void wgrad(tpt_sur *s, hipStream_t st, int l, const bf16_t *x, const bf16_t *dz, int accumulate) {
    const Layer &L = s->layer[l];
    WgradArgs a;
    a.x = x;
    a.dz = dz;
    a.part = s->part;
    a.H = s->H >> L.level;
    a.W = s->W >> L.level;
    a.cin = L.cin;
    a.lcin = ilog2i(L.cin);
    a.cout = L.cout;
    a.cout16 = round_up(L.cout, 16);
    a.kpad = L.kpad;
    const dim3 g(L.kpad / 64, (a.H + WG_ROWS - 1) / WG_ROWS, (a.W + WG_SEG - 1) / WG_SEG);
    const size_t lds = wgrad_lds_bytes(L.cin, a.cout16);
    switch (a.cout16 / 16) {
    case 1: hipLaunchKernelGGL(conv3x3_wgrad<1>, g, dim3(WG_THREADS), lds, st, a); break;
    case 2: hipLaunchKernelGGL(conv3x3_wgrad<2>, g, dim3(WG_THREADS), lds, st, a); break;
    case 4: hipLaunchKernelGGL(conv3x3_wgrad<4>, g, dim3(WG_THREADS), lds, st, a); break;
    default: hipLaunchKernelGGL(conv3x3_wgrad<8>, g, dim3(WG_THREADS), lds, st, a); break;
    }
    const int groups = (int)(g.y * g.z);
    hipLaunchKernelGGL(wgrad_reduce, grid1((size_t)L.cout * L.kpad), dim3(256), 0, st, s->part,
                       s->grad + L.offset, groups, L.cout, a.cout16, L.kpad, accumulate);
}

// `z` receives the pre-activation when training needs it (null otherwise).
void fwd_layer(tpt_sur *s, hipStream_t st, int l, const bf16_t *in, bf16_t *out, bf16_t *z, float *out_f32,
               const StageIn &sin = StageIn()) {
    const Layer &L = s->layer[l];
    conv(s, st, in, l == 0 && sin.stage != STAGE_PREP ? s->x0lo : nullptr, L.cin, L.cout, L.level, L.w, L.kpad,
         L.kchunks, out, z, nullptr, out_f32, L.gelu, 1, sin);
}

// Data gradient through layer l: dx = conv(dz, Wt), times GELU'(zref) when
// given.
void dgrad_layer(tpt_sur *s, hipStream_t st, int l, const bf16_t *dz, bf16_t *dx, const bf16_t *zref) {
    const Layer &L = s->layer[l];
    conv(s, st, dz, nullptr, L.cout, L.cin, L.level, L.wt, L.kpad_t, L.kchunks_t, dx, nullptr, zref, nullptr, 0, 0);
}

// The network on s->x0 into s->out; `train` keeps the pre-activations the
// backward pass needs.  Inference with `s->fused` set folds the pools and
// the upsample-adds into the staging of the convolutions that read them,
// and, given `prep` (the air planes), the input preparation into the first
// layer; the results are bit for bit those of the separate kernels.
void forward(tpt_sur *s, hipStream_t st, bool train, const StageIn *prep = nullptr) {
    const int H = s->H, W = s->W;
    if (!train && s->fused) {
        fwd_layer(s, st, 0, s->x0, s->a1, nullptr, nullptr, prep ? *prep : StageIn());
        StageIn pool;
        pool.stage = STAGE_POOL;
        fwd_layer(s, st, 1, s->a1, s->a2, nullptr, nullptr, pool);
        fwd_layer(s, st, 2, s->a2, s->a2b, nullptr, nullptr);
        fwd_layer(s, st, 3, s->a2b, s->a3, nullptr, nullptr, pool);
        fwd_layer(s, st, 4, s->a3, s->a3b, nullptr, nullptr);
        fwd_layer(s, st, 5, s->a3b, s->d2a, nullptr, nullptr);
        StageIn up;
        up.stage = STAGE_UPADD;
        up.in2 = s->d2a;
        fwd_layer(s, st, 6, s->a2b, s->d2b, nullptr, nullptr, up);
        fwd_layer(s, st, 7, s->d2b, s->d1a, nullptr, nullptr);
        up.in2 = s->d1a;
        fwd_layer(s, st, 8, s->a1, s->d1b, nullptr, nullptr, up);
        fwd_layer(s, st, 9, s->d1b, nullptr, nullptr, s->out);
        return;
    }
    if (prep) {
        const size_t n = cells_at(s, 0);
        hipLaunchKernelGGL(sur_prep, grid1(n), dim3(256), 0, st, Planes{prep->pv, prep->vx, prep->vy, prep->hv}, nullptr,
                           prep->wall, prep->src, s->x0, s->x0lo, (int)n, s->in_scale[0], s->in_scale[1],
                           s->in_scale[3], s->ambient);
    }
    const auto Z = [train](bf16_t *z) { return train ? z : nullptr; };
    fwd_layer(s, st, 0, s->x0, s->a1, Z(s->z1), nullptr);
    hipLaunchKernelGGL(pool2, grid1(cells_at(s, 1) * 4), dim3(256), 0, st, s->a1, s->p1, H, W, 32);
    fwd_layer(s, st, 1, s->p1, s->a2, Z(s->z2), nullptr);
    fwd_layer(s, st, 2, s->a2, s->a2b, Z(s->z2b), nullptr);
    hipLaunchKernelGGL(pool2, grid1(cells_at(s, 2) * 8), dim3(256), 0, st, s->a2b, s->p2, H / 2, W / 2, 64);
    fwd_layer(s, st, 3, s->p2, s->a3, Z(s->z3), nullptr);
    fwd_layer(s, st, 4, s->a3, s->a3b, Z(s->z3b), nullptr);
    fwd_layer(s, st, 5, s->a3b, s->d2a, Z(s->zd2a), nullptr);
    hipLaunchKernelGGL(up2_add, grid1(cells_at(s, 1) * 8), dim3(256), 0, st, s->d2a, s->a2b, s->u2, H / 2, W / 2,
                       64);
    fwd_layer(s, st, 6, s->u2, s->d2b, Z(s->zd2b), nullptr);
    fwd_layer(s, st, 7, s->d2b, s->d1a, Z(s->zd1a), nullptr);
    hipLaunchKernelGGL(up2_add, grid1(cells_at(s, 0) * 4), dim3(256), 0, st, s->d1a, s->a1, s->u1, H, W, 32);
    fwd_layer(s, st, 8, s->u1, s->d1b, Z(s->zd1b), nullptr);
    fwd_layer(s, st, 9, s->d1b, nullptr, nullptr, s->out);
}

// Backward from s->g_out (the loss gradient, H x W x OUT_CH bf16); weight
// gradients accumulate into s->grad when `accumulate`.
void backward(tpt_sur *s, hipStream_t st, int accumulate) {
    const int H = s->H, W = s->W;
    wgrad(s, st, 9, s->d1b, s->g_out, accumulate);
    dgrad_layer(s, st, 9, s->g_out, s->g_d1b, s->zd1b);
    wgrad(s, st, 8, s->u1, s->g_d1b, accumulate);
    dgrad_layer(s, st, 8, s->g_d1b, s->g_u1, nullptr);
    hipLaunchKernelGGL(up2_bwd_gelu, grid1(cells_at(s, 1) * 4), dim3(256), 0, st, s->g_u1, s->zd1a, s->g_d1a, H, W,
                       32);
    wgrad(s, st, 7, s->d2b, s->g_d1a, accumulate);
    dgrad_layer(s, st, 7, s->g_d1a, s->g_d2b, s->zd2b);
    wgrad(s, st, 6, s->u2, s->g_d2b, accumulate);
    dgrad_layer(s, st, 6, s->g_d2b, s->g_u2, nullptr);
    hipLaunchKernelGGL(up2_bwd_gelu, grid1(cells_at(s, 2) * 8), dim3(256), 0, st, s->g_u2, s->zd2a, s->g_d2a,
                       H / 2, W / 2, 64);
    wgrad(s, st, 5, s->a3b, s->g_d2a, accumulate);
    dgrad_layer(s, st, 5, s->g_d2a, s->g_a3b, s->z3b);
    wgrad(s, st, 4, s->a3, s->g_a3b, accumulate);
    dgrad_layer(s, st, 4, s->g_a3b, s->g_a3, s->z3);
    wgrad(s, st, 3, s->p2, s->g_a3, accumulate);
    dgrad_layer(s, st, 3, s->g_a3, s->g_p2, nullptr);
    hipLaunchKernelGGL(pool2_bwd_add_gelu, grid1(cells_at(s, 1) * 8), dim3(256), 0, st, s->g_p2, s->g_u2, s->z2b,
                       s->g_a2b, H / 2, W / 2, 64);
    wgrad(s, st, 2, s->a2, s->g_a2b, accumulate);
    dgrad_layer(s, st, 2, s->g_a2b, s->g_a2, s->z2);
    wgrad(s, st, 1, s->p1, s->g_a2, accumulate);
    dgrad_layer(s, st, 1, s->g_a2, s->g_p1, nullptr);
    hipLaunchKernelGGL(pool2_bwd_add_gelu, grid1(cells_at(s, 0) * 4), dim3(256), 0, st, s->g_p1, s->g_u1, s->z1,
                       s->g_a1, H, W, 32);
    wgrad(s, st, 0, s->x0, s->g_a1, accumulate);
}

void refresh_transposes(tpt_sur *s, hipStream_t st) {
    for (int l = 0; l < N_LAYERS; ++l) {
        const Layer &L = s->layer[l];
        hipLaunchKernelGGL(transpose_flip, grid1((size_t)L.rows_t * L.kpad_t), dim3(256), 0, st, L.w, L.wt, L.cin,
                           L.cout, L.kpad, L.rows_t, L.kpad_t);
    }
}

// ---------------------------------------------------------------------------
// Air-grid glue.

// Loss and its gradient against the next snapshot: target channel c =
// (next_c - cur_c) / scale_c; loss = mean squared error over 4 channels.
__global__ void sur_loss(const float *__restrict__ out, Planes cur, Planes nxt, bf16_t *__restrict__ g_out,
                         float *__restrict__ red, int n, float sp, float sv, float sh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        const float tg[4] = {(nxt.pv[i] - cur.pv[i]) / sp, (nxt.vx[i] - cur.vx[i]) / sv, (nxt.vy[i] - cur.vy[i]) / sv,
                             (nxt.hv[i] - cur.hv[i]) / sh};
        const float inv = 2.0f / (4.0f * (float)n);
        bf16_t *g = g_out + (size_t)i * OUT_CH;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float d = out[(size_t)i * OUT_CH + c] - tg[c];
            l += d * d;
            g[c] = f2bf(inv * d);
        }
#pragma unroll
        for (int c = 4; c < OUT_CH; ++c) g[c] = 0;
    }
    float v[1] = {l / (4.0f * (float)n)};
    block_sum<1>(v, red);
}

// state += prediction * scale; heat sources pinned; the plain step's clamps;
// energy sums before/after and a non-finite count for the check.
__global__ void sur_apply(const float *__restrict__ out, float *__restrict__ pv, float *__restrict__ vx,
                          float *__restrict__ vy, float *__restrict__ hv, const float *__restrict__ src,
                          float *__restrict__ red, float *__restrict__ saved, int n, float sp, float sv, float sh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float e_in = 0.0f, e_out = 0.0f, bad = 0.0f;
    if (i < n) {
        const float p0 = pv[i], u0 = vx[i], v0 = vy[i], h0 = hv[i];
        if (saved) {  // the state before the call, for the fallback
            saved[i] = p0;
            saved[n + i] = u0;
            saved[2 * n + i] = v0;
            saved[3 * n + i] = h0;
        }
        e_in = p0 * p0 + u0 * u0 + v0 * v0;
        const float *o = out + (size_t)i * OUT_CH;
        float p = p0 + o[0] * sp, u = u0 + o[1] * sv, v = v0 + o[2] * sv, h = h0 + o[3] * sh;
        if (!(isfinite(p) && isfinite(u) && isfinite(v) && isfinite(h))) bad = 1.0f;
        p = fminf(fmaxf(p, -256.0f), 256.0f);
        u = fminf(fmaxf(u, -256.0f), 256.0f);
        v = fminf(fmaxf(v, -256.0f), 256.0f);
        h = fminf(fmaxf(h, 0.0f), 9999.0f);
        if (src[i] > 0.0f) h = src[i];
        e_out = p * p + u * u + v * v;
        pv[i] = p;
        vx[i] = u;
        vy[i] = v;
        hv[i] = h;
    }
    if (red) {
        float v[3] = {e_in, e_out, bad};
        block_sum<3>(v, red);
    }
}

// The state a call would leave: cur + prediction * scale with the apply's
// clamps and heat sources, into dst (a training rollout's first step).
__global__ void sur_predict(const float *__restrict__ out, Planes cur, const float *__restrict__ src,
                            float *__restrict__ dst, int n, float sp, float sv, float sh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *o = out + (size_t)i * OUT_CH;
    float p = cur.pv[i] + o[0] * sp, u = cur.vx[i] + o[1] * sv, v = cur.vy[i] + o[2] * sv,
          h = cur.hv[i] + o[3] * sh;
    p = fminf(fmaxf(p, -256.0f), 256.0f);
    u = fminf(fmaxf(u, -256.0f), 256.0f);
    v = fminf(fmaxf(v, -256.0f), 256.0f);
    h = fminf(fmaxf(h, 0.0f), 9999.0f);
    if (src[i] > 0.0f) h = src[i];
    dst[i] = p;
    dst[n + i] = u;
    dst[2 * n + i] = v;
    dst[3 * n + i] = h;
}

__global__ void sur_pin(float *__restrict__ hv, const float *__restrict__ src, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && src[i] > 0.0f) hv[i] = src[i];
}

__global__ void sur_zero_u16(uint16_t *__restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void sur_zero_f(float *__restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// One scene on the context's air grids: Gaussian pressure blobs, zero
// velocity, ambient heat, wall rectangles plus the border ring (WL_WALL),
// the wall masks the air step reads, and the heat-source map.
__global__ void sur_scene_kernel(tpt_sur_scene_desc sc, float *__restrict__ pv, float *__restrict__ vx,
                                 float *__restrict__ vy, float *__restrict__ hv, uint32_t *__restrict__ wall,
                                 uint32_t *__restrict__ emap, uint32_t *__restrict__ ba, uint32_t *__restrict__ bh,
                                 float *__restrict__ src, int xc, int yc, float amb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= xc * yc) return;
    const int x = i % xc, y = i / xc;
    float p = 0.0f;
    for (uint32_t b = 0; b < sc.n_blobs && b < TPT_SUR_MAX_BLOBS; ++b) {
        const float dx = (float)x - sc.blob[b][0], dy = (float)y - sc.blob[b][1], s = sc.blob[b][3];
        p += sc.blob[b][2] * expf(-(dx * dx + dy * dy) / (2.0f * s * s));
    }
    uint32_t w = (x == 0 || y == 0 || x == xc - 1 || y == yc - 1) ? 8u : 0u;
    for (uint32_t r = 0; r < sc.n_rects && r < TPT_SUR_MAX_RECTS; ++r)
        if (x >= (int)sc.rect[r][0] && x < (int)sc.rect[r][2] && y >= (int)sc.rect[r][1] && y < (int)sc.rect[r][3])
            w = 8u;
    float t = 0.0f;
    for (uint32_t k = 0; k < sc.n_sources && k < TPT_SUR_MAX_SOURCES; ++k)
        if (x >= (int)sc.source[k][0] && x < (int)sc.source[k][2] && y >= (int)sc.source[k][1] &&
            y < (int)sc.source[k][3] && w == 0u)
            t = sc.source_temp[k];
    pv[i] = p;
    vx[i] = 0.0f;
    vy[i] = 0.0f;
    hv[i] = t > 0.0f ? t : amb;
    wall[i] = w;
    emap[i] = 0u;
    ba[i] = w ? 1u : 0u;
    bh[i] = w ? 8u : 0u;
    src[i] = t;
}

__global__ void sur_capture(Planes in, float *__restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = in.pv[i];
    dst[n + i] = in.vx[i];
    dst[2 * n + i] = in.vy[i];
    dst[3 * n + i] = in.hv[i];
}

__global__ void sur_restore(const float *__restrict__ srcp, float *__restrict__ pv, float *__restrict__ vx,
                            float *__restrict__ vy, float *__restrict__ hv, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pv[i] = srcp[i];
    vx[i] = srcp[n + i];
    vy[i] = srcp[2 * n + i];
    hv[i] = srcp[3 * n + i];
}

__global__ void sur_wall_mask(const uint32_t *__restrict__ ba, uint8_t *__restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = ba[i] ? 1u : 0u;
}

Planes snap_planes(const tpt_sur *s, uint32_t traj, uint32_t snap) {
    const size_t n = cells_at(s, 0);
    const float *b = s->data + ((size_t)traj * s->n_snap + snap) * 4 * n;
    return Planes{b, b + n, b + 2 * n, b + 3 * n};
}

Planes ctx_planes(tpt_ctx *ctx) {
    return Planes{tpt_buf<float>(ctx, TPT_BUF_AIR_PV), tpt_buf<float>(ctx, TPT_BUF_AIR_VX),
                  tpt_buf<float>(ctx, TPT_BUF_AIR_VY), tpt_buf<float>(ctx, TPT_BUF_AIR_HV)};
}

int check_ctx(const tpt_sur *s, const tpt_ctx *ctx) {
    if (!s || !ctx) return tpt_set_error(TPT_EINVAL, "null surrogate or context");
    if ((int)ctx->xcells != s->W || (int)ctx->ycells != s->H)
        return tpt_set_error(TPT_EINVAL, "context grid %ux%u differs from the surrogate's %dx%d", ctx->xcells,
                             ctx->ycells, s->W, s->H);
    return TPT_OK;
}

// Plain air ticks with the scene's heat sources held at their temperature.
int plain_ticks(tpt_sur *s, tpt_ctx *ctx, uint32_t ticks) {
    const int n = (int)cells_at(s, 0);
    for (uint32_t t = 0; t < ticks; ++t) {
        int rc = tpt_air_step(ctx);
        if (rc) return rc;
        hipLaunchKernelGGL(sur_pin, grid1(n), dim3(256), 0, ctx->stream, tpt_buf<float>(ctx, TPT_BUF_AIR_HV),
                           s->src, n);
    }
    return tpt_check_launch("surrogate plain ticks");
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" int tpt_sur_create(uint32_t xcells, uint32_t ycells, tpt_sur **out) {
    if (!out) return tpt_set_error(TPT_EINVAL, "null output");
    *out = nullptr;
    if (xcells < 8 || ycells < 8 || xcells % 4 || ycells % 4 || xcells > TPT_MAX_DIM / 4 || ycells > TPT_MAX_DIM / 4)
        return tpt_set_error(TPT_EINVAL, "surrogate grid %ux%u: sides must be multiples of 4 in [8, %u]", xcells,
                             ycells, TPT_MAX_DIM / 4);
    tpt_sur *s = new (std::nothrow) tpt_sur();
    if (!s) return tpt_set_error(TPT_ENOMEM, "surrogate object");
    s->H = (int)ycells;
    s->W = (int)xcells;
    size_t off = 0, wt_elems = 0, part = 0;
    for (int l = 0; l < N_LAYERS; ++l) {
        Layer &L = s->layer[l];
        L.cin = SPEC[l].cin;
        L.cout = SPEC[l].cout;
        L.level = SPEC[l].level;
        L.gelu = SPEC[l].gelu;
        L.kpad = round_up(9 * L.cin + 1, 64);
        L.kchunks = round_up(9 * L.cin + 1, 32) / 32;
        L.rows = round_up(L.cout, 64);
        L.kpad_t = round_up(9 * L.cout, 64);
        L.kchunks_t = round_up(9 * L.cout, 32) / 32;
        L.rows_t = round_up(L.cin, 64);
        L.offset = off;
        off += (size_t)L.rows * L.kpad;
        wt_elems += (size_t)L.rows_t * L.kpad_t;
        const int h = s->H >> L.level, w = s->W >> L.level;
        const size_t groups = (size_t)((h + WG_ROWS - 1) / WG_ROWS) * ((w + WG_SEG - 1) / WG_SEG);
        const size_t need = groups * round_up(L.cout, 16) * L.kpad;
        if (need > part) part = need;
    }
    s->n_params = off;
    s->part_elems = part;
    int rc = TPT_OK;
    const size_t n0 = cells_at(s, 0), n1 = cells_at(s, 1), n2 = cells_at(s, 2);
    if (!rc) rc = alloc_n(s, &s->wf, off);
    if (!rc) rc = alloc_n(s, &s->grad, off);
    if (!rc) rc = alloc_n(s, &s->m1, off);
    if (!rc) rc = alloc_n(s, &s->m2, off);
    if (!rc) rc = alloc_n(s, &s->wb, off);
    if (!rc) rc = alloc_n(s, &s->wtb, wt_elems);
    if (!rc) rc = alloc_n(s, &s->part, part);
    struct A {
        bf16_t **p;
        size_t n;
    } acts[] = {
        {&s->x0, n0 * IN_CH}, {&s->x0lo, n0 * IN_CH}, {&s->a1, n0 * 32}, {&s->z1, n0 * 32}, {&s->p1, n1 * 32}, {&s->a2, n1 * 64},
        {&s->z2, n1 * 64}, {&s->a2b, n1 * 64}, {&s->z2b, n1 * 64}, {&s->p2, n2 * 64}, {&s->a3, n2 * 128},
        {&s->z3, n2 * 128}, {&s->a3b, n2 * 128}, {&s->z3b, n2 * 128}, {&s->d2a, n2 * 64}, {&s->zd2a, n2 * 64},
        {&s->u2, n1 * 64}, {&s->d2b, n1 * 64}, {&s->zd2b, n1 * 64}, {&s->d1a, n1 * 32}, {&s->zd1a, n1 * 32},
        {&s->u1, n0 * 32}, {&s->d1b, n0 * 32}, {&s->zd1b, n0 * 32}, {&s->g_out, n0 * OUT_CH},
        {&s->g_d1b, n0 * 32}, {&s->g_u1, n0 * 32}, {&s->g_d1a, n1 * 32}, {&s->g_d2b, n1 * 64},
        {&s->g_u2, n1 * 64}, {&s->g_d2a, n2 * 64}, {&s->g_a3b, n2 * 128}, {&s->g_a3, n2 * 128},
        {&s->g_p2, n2 * 64}, {&s->g_a2b, n1 * 64}, {&s->g_a2, n1 * 64}, {&s->g_p1, n1 * 32}, {&s->g_a1, n0 * 32},
    };
    for (auto &a : acts)
        if (!rc) rc = alloc_n(s, a.p, a.n);
    if (!rc) rc = alloc_n(s, &s->out, n0 * OUT_CH);
    if (!rc) rc = alloc_n(s, &s->loss_acc, 4);
    if (!rc) rc = alloc_n(s, &s->gnorm, 1);
    if (!rc) rc = alloc_n(s, &s->red, 4 * std::max<size_t>(grid1(n0).x, grid1(s->n_params).x));
    if (!rc) rc = alloc_n(s, &s->src, n0);
    if (!rc) rc = alloc_n(s, &s->saved, 4 * n0);
    if (!rc) rc = alloc_n(s, &s->roll, 4 * n0);
    if (!rc && hipHostMalloc((void **)&s->check_host, 4 * sizeof(float), hipHostMallocDefault) != hipSuccess) {
        s->check_host = nullptr;
        rc = tpt_set_error(TPT_ENOMEM, "surrogate: pinned check buffer");
    }
    if (!rc && hipStreamCreate(&s->stream) != hipSuccess) rc = tpt_set_error(TPT_EDEVICE, "surrogate stream");
    if (rc) {
        tpt_sur_destroy(s);
        return rc;
    }
    size_t wt_off = 0;
    for (int l = 0; l < N_LAYERS; ++l) {
        Layer &L = s->layer[l];
        L.w = s->wb + L.offset;
        L.wt = s->wtb + wt_off;
        wt_off += (size_t)L.rows_t * L.kpad_t;
    }
    set_conv_lds<1, 1, false>((int)conv_lds_bytes(128, 1, 1));
    set_conv_lds<2, 1, false>((int)conv_lds_bytes(128, 1, 2));
    set_conv_lds<4, 1, false>((int)conv_lds_bytes(128, 1, 4));
    set_conv_lds<1, 2, false>((int)conv_lds_bytes(128, 2, 1));
    set_conv_lds<2, 2, false>((int)conv_lds_bytes(128, 2, 2));
    set_conv_lds<4, 2, false>((int)conv_lds_bytes(128, 2, 4));
    set_conv_lds<1, 4, false>((int)conv_lds_bytes(128, 4, 1));
    set_conv_lds<2, 4, false>((int)conv_lds_bytes(128, 4, 2));
    set_conv_lds<4, 4, false>((int)conv_lds_bytes(128, 4, 4));
    set_conv_lds<1, 1, true>((int)conv_lds_bytes(IN_CH, 1, 1, true));
    set_conv_lds<2, 1, true>((int)conv_lds_bytes(IN_CH, 1, 2, true));
    set_conv_lds<4, 1, true>((int)conv_lds_bytes(IN_CH, 1, 4, true));
    set_wgrad_lds<1>((int)wgrad_lds_bytes(128, 128));
    set_wgrad_lds<2>((int)wgrad_lds_bytes(128, 128));
    set_wgrad_lds<4>((int)wgrad_lds_bytes(128, 128));
    set_wgrad_lds<8>((int)wgrad_lds_bytes(128, 128));
    hipLaunchKernelGGL(sur_zero_f, grid1(n0), dim3(256), 0, s->stream, s->src, n0);
    *out = s;
    return tpt_sur_init_params(s, 1);
}

extern "C" int tpt_sur_destroy(tpt_sur *s) {
    if (!s) return TPT_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void *p : s->allocs) (void)hipFree(p);
    if (s->check_host) (void)hipHostFree(s->check_host);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return TPT_OK;
}

extern "C" int tpt_sur_param_count(tpt_sur *s, size_t *n) {
    if (!s || !n) return tpt_set_error(TPT_EINVAL, "null argument");
    *n = s->n_params;
    return TPT_OK;
}

// He-normal weights (std sqrt(2 / (9 cin))), zero bias and padding; the
// output layer starts at a tenth of that so the first predictions are small.
extern "C" int tpt_sur_init_params(tpt_sur *s, uint64_t seed) {
    if (!s) return tpt_set_error(TPT_EINVAL, "null surrogate");
    std::vector<float> h(s->n_params, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < N_LAYERS; ++l) {
        const Layer &L = s->layer[l];
        const int real_out = l == N_LAYERS - 1 ? 4 : L.cout;
        float sd = sqrtf(2.0f / (9.0f * (float)L.cin));
        if (l == N_LAYERS - 1) sd *= 0.1f;
        const int real_in = l == 0 ? 6 : L.cin;
        std::normal_distribution<float> nd(0.0f, sd);
        for (int co = 0; co < real_out; ++co)
            for (int tap = 0; tap < 9; ++tap)
                for (int ci = 0; ci < real_in; ++ci) h[L.offset + (size_t)co * L.kpad + tap * L.cin + ci] = nd(rng);
    }
    return tpt_sur_set_params(s, h.data(), h.size());
}

extern "C" int tpt_sur_set_params(tpt_sur *s, const float *params, size_t n) {
    if (!s || !params) return tpt_set_error(TPT_EINVAL, "null argument");
    if (n != s->n_params) return tpt_set_error(TPT_EINVAL, "%zu parameters given, the network has %zu", n, s->n_params);
    TPT_HIP_CHECK(hipMemcpyAsync(s->wf, params, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(f32_to_bf16, grid1(n), dim3(256), 0, s->stream, s->wf, s->wb, (int)n);
    hipLaunchKernelGGL(sur_zero_f, grid1(n), dim3(256), 0, s->stream, s->m1, n);
    hipLaunchKernelGGL(sur_zero_f, grid1(n), dim3(256), 0, s->stream, s->m2, n);
    hipLaunchKernelGGL(sur_zero_f, grid1(n), dim3(256), 0, s->stream, s->grad, n);
    s->adam_t = 0;
    refresh_transposes(s, s->stream);
    TPT_HIP_CHECK(hipStreamSynchronize(s->stream));
    return tpt_check_launch("surrogate parameters");
}

extern "C" int tpt_sur_get_params(tpt_sur *s, float *params, size_t n) {
    if (!s || !params) return tpt_set_error(TPT_EINVAL, "null argument");
    if (n != s->n_params) return tpt_set_error(TPT_EINVAL, "%zu parameters asked, the network has %zu", n, s->n_params);
    TPT_HIP_CHECK(hipMemcpyAsync(params, s->wf, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    TPT_HIP_CHECK(hipStreamSynchronize(s->stream));
    return TPT_OK;
}

extern "C" int tpt_sur_get_grads(tpt_sur *s, float *grads, size_t n) {
    if (!s || !grads) return tpt_set_error(TPT_EINVAL, "null argument");
    if (n != s->n_params) return tpt_set_error(TPT_EINVAL, "%zu gradients asked, the network has %zu", n, s->n_params);
    TPT_HIP_CHECK(hipMemcpyAsync(grads, s->grad, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    TPT_HIP_CHECK(hipStreamSynchronize(s->stream));
    return TPT_OK;
}

extern "C" int tpt_sur_set_scales(tpt_sur *s, const float in_scales[4], const float out_scales[4], float ambient) {
    if (!s || !in_scales || !out_scales) return tpt_set_error(TPT_EINVAL, "null argument");
    for (int c = 0; c < 4; ++c)
        if (!(in_scales[c] > 0.0f) || !(out_scales[c] > 0.0f))
            return tpt_set_error(TPT_EINVAL, "scale %d must be positive", c);
    for (int c = 0; c < 4; ++c) {
        s->in_scale[c] = in_scales[c];
        s->scale[c] = out_scales[c];
    }
    s->ambient = ambient;
    return TPT_OK;
}

// Raw forward / backward on host buffers (tests): x0 is H x W x 8 bf16 bits
// (an input that bf16 holds exactly: the low half is zero),
// out H x W x 8 fp32; g_out (H x W x 8 bf16 bits) seeds the backward and the
// weight gradients replace s->grad.
extern "C" int tpt_sur_forward_raw(tpt_sur *s, const uint16_t *x0, float *out) {
    if (!s || !x0 || !out) return tpt_set_error(TPT_EINVAL, "null argument");
    const size_t n = cells_at(s, 0);
    TPT_HIP_CHECK(hipMemcpyAsync(s->x0, x0, n * IN_CH * 2, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(sur_zero_u16, grid1(n * IN_CH), dim3(256), 0, s->stream, s->x0lo, n * IN_CH);
    forward(s, s->stream, true);
    TPT_HIP_CHECK(hipMemcpyAsync(out, s->out, n * OUT_CH * 4, hipMemcpyDeviceToHost, s->stream));
    TPT_HIP_CHECK(hipStreamSynchronize(s->stream));
    return tpt_check_launch("surrogate forward");
}

extern "C" int tpt_sur_backward_raw(tpt_sur *s, const uint16_t *g_out) {
    if (!s || !g_out) return tpt_set_error(TPT_EINVAL, "null argument");
    const size_t n = cells_at(s, 0);
    TPT_HIP_CHECK(hipMemcpyAsync(s->g_out, g_out, n * OUT_CH * 2, hipMemcpyHostToDevice, s->stream));
    backward(s, s->stream, 0);
    TPT_HIP_CHECK(hipStreamSynchronize(s->stream));
    return tpt_check_launch("surrogate backward");
}

// Scene set-up on a context of the surrogate's grid size.
extern "C" int tpt_sur_scene(tpt_sur *s, tpt_ctx *ctx, const tpt_sur_scene_desc *sc) {
    int rc = check_ctx(s, ctx);
    if (rc) return rc;
    if (!sc) return tpt_set_error(TPT_EINVAL, "null scene");
    const int n = (int)cells_at(s, 0);
    hipLaunchKernelGGL(sur_scene_kernel, grid1(n), dim3(256), 0, ctx->stream, *sc, tpt_buf<float>(ctx, TPT_BUF_AIR_PV),
                       tpt_buf<float>(ctx, TPT_BUF_AIR_VX), tpt_buf<float>(ctx, TPT_BUF_AIR_VY),
                       tpt_buf<float>(ctx, TPT_BUF_AIR_HV), tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP),
                       tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP), tpt_buf<uint32_t>(ctx, TPT_BUF_BLOCK_AIR),
                       tpt_buf<uint32_t>(ctx, TPT_BUF_BLOCK_AIRH), s->src, s->W, s->H, s->ambient);
    ctx->maps_fresh = ctx->rec_cover = ctx->rebuild_fresh = false;
    ctx->has_wire = ctx->has_stasis = false;
    ctx->epoch++;
    return tpt_check_launch("surrogate scene");
}

extern "C" int tpt_sur_plain_ticks(tpt_sur *s, tpt_ctx *ctx, uint32_t ticks) {
    int rc = check_ctx(s, ctx);
    return rc ? rc : plain_ticks(s, ctx, ticks);
}

extern "C" int tpt_sur_dataset_alloc(tpt_sur *s, uint32_t n_traj, uint32_t n_snap) {
    if (!s || n_traj == 0 || n_snap < 2) return tpt_set_error(TPT_EINVAL, "dataset needs a trajectory of 2+ snapshots");
    if (s->data) return tpt_set_error(TPT_ESTATE, "dataset already allocated");
    const size_t n = cells_at(s, 0);
    int rc = alloc_n(s, &s->data, (size_t)n_traj * n_snap * 4 * n);
    if (!rc) rc = alloc_n(s, &s->walls, (size_t)n_traj * n);
    if (!rc) rc = alloc_n(s, &s->srcs, (size_t)n_traj * n);
    if (rc) return rc;
    s->n_traj = n_traj;
    s->n_snap = n_snap;
    return TPT_OK;
}

// Snapshot `snap` of trajectory `traj` from the context's air grids; the
// first snapshot also records the trajectory's walls and sources.
extern "C" int tpt_sur_dataset_capture(tpt_sur *s, tpt_ctx *ctx, uint32_t traj, uint32_t snap) {
    int rc = check_ctx(s, ctx);
    if (rc) return rc;
    if (!s->data || traj >= s->n_traj || snap >= s->n_snap) return tpt_set_error(TPT_EINVAL, "snapshot out of range");
    const int n = (int)cells_at(s, 0);
    float *dst = s->data + ((size_t)traj * s->n_snap + snap) * 4 * n;
    hipLaunchKernelGGL(sur_capture, grid1(n), dim3(256), 0, ctx->stream, ctx_planes(ctx), dst, n);
    if (snap == 0) {
        hipLaunchKernelGGL(sur_wall_mask, grid1(n), dim3(256), 0, ctx->stream,
                           tpt_buf<uint32_t>(ctx, TPT_BUF_BLOCK_AIR), s->walls + (size_t)traj * n, n);
        TPT_HIP_CHECK(hipMemcpyAsync(s->srcs + (size_t)traj * n, s->src, n * sizeof(float), hipMemcpyDeviceToDevice,
                                     ctx->stream));
    }
    return tpt_check_launch("surrogate capture");
}

// Training on dataset pairs (snapshot t -> t + 1 of a trajectory):
// pairs[2i] = trajectory, pairs[2i + 1] = snapshot.  Gradients are summed
// over `accum` pairs, then Adam steps with the mean.  *mean_loss receives the
// average loss over the pairs.
extern "C" int tpt_sur_train(tpt_sur *s, const uint32_t *pairs, uint32_t n_pairs, uint32_t accum, float lr,
                             float *mean_loss) {
    if (!s || !pairs || !n_pairs || !accum) return tpt_set_error(TPT_EINVAL, "bad training arguments");
    if (!s->data) return tpt_set_error(TPT_ESTATE, "no dataset");
    for (uint32_t i = 0; i < n_pairs; ++i)
        if (pairs[2 * i] >= s->n_traj || pairs[2 * i + 1] + s->rollout >= s->n_snap)
            return tpt_set_error(TPT_EINVAL, "pair %u out of range for a %u-step rollout", i, s->rollout);
    const int n = (int)cells_at(s, 0);
    hipStream_t st = s->stream;
    hipLaunchKernelGGL(sur_zero_f, dim3(1), dim3(64), 0, st, s->loss_acc, 4);
    uint32_t in_batch = 0;
    for (uint32_t i = 0; i < n_pairs; ++i) {
        const uint32_t tr = pairs[2 * i], sn = pairs[2 * i + 1];
        Planes cur = snap_planes(s, tr, sn);
        const Planes nxt = snap_planes(s, tr, sn + s->rollout);
        const uint8_t *walls = s->walls + (size_t)tr * n;
        const float *srcs = s->srcs + (size_t)tr * n;
        for (uint32_t r = 1; r < s->rollout; ++r) {
            // the network's own steps, without gradient (each reads the
            // previous prediction and overwrites it cell by cell)
            hipLaunchKernelGGL(sur_prep, grid1(n), dim3(256), 0, st, cur, walls, nullptr, srcs, s->x0, s->x0lo, n,
                               s->in_scale[0], s->in_scale[1], s->in_scale[3], s->ambient);
            forward(s, st, false);
            hipLaunchKernelGGL(sur_predict, grid1(n), dim3(256), 0, st, s->out, cur, srcs, s->roll, n, s->scale[0],
                               s->scale[1], s->scale[3]);
            cur = Planes{s->roll, s->roll + n, s->roll + 2 * (size_t)n, s->roll + 3 * (size_t)n};
        }
        hipLaunchKernelGGL(sur_prep, grid1(n), dim3(256), 0, st, cur, walls, nullptr, srcs, s->x0, s->x0lo, n,
                           s->in_scale[0], s->in_scale[1], s->in_scale[3], s->ambient);
        forward(s, st, true);
        hipLaunchKernelGGL(sur_loss, grid1(n), dim3(256), 0, st, s->out, cur, nxt, s->g_out, s->red, n,
                           s->scale[0], s->scale[1], s->scale[3]);
        hipLaunchKernelGGL(sum_blocks, dim3(1), dim3(256), 0, st, s->red, (int)grid1(n).x, 1, s->loss_acc, 1);
        backward(s, st, in_batch > 0 ? 1 : 0);
        if (++in_batch == accum || i + 1 == n_pairs) {
            s->adam_t += 1;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.0f - powf(b1, (float)s->adam_t), c2 = 1.0f - powf(b2, (float)s->adam_t);
            hipLaunchKernelGGL(sur_zero_f, dim3(1), dim3(64), 0, st, s->gnorm, 1);
            hipLaunchKernelGGL(grad_sumsq, grid1(s->n_params), dim3(256), 0, st, s->grad, (int)s->n_params, s->red);
            hipLaunchKernelGGL(sum_blocks, dim3(1), dim3(256), 0, st, s->red, (int)grid1(s->n_params).x, 1,
                               s->gnorm, 1);
            hipLaunchKernelGGL(adam_step, grid1(s->n_params), dim3(256), 0, st, s->wf, s->wb, s->grad, s->m1, s->m2,
                               (int)s->n_params, lr, b1, b2, 1e-8f, c1, c2, 1.0f / (float)in_batch, s->gnorm,
                               s->clip);
            refresh_transposes(s, st);
            in_batch = 0;
        }
    }
    float acc[4];
    TPT_HIP_CHECK(hipMemcpyAsync(acc, s->loss_acc, sizeof acc, hipMemcpyDeviceToHost, st));
    TPT_HIP_CHECK(hipStreamSynchronize(st));
    if (mean_loss) *mean_loss = acc[0] / (float)n_pairs;
    return tpt_check_launch("surrogate training");
}

// Global L2 norm bound of the averaged batch gradient (0 turns clipping off).
extern "C" int tpt_sur_set_rollout(tpt_sur *s, uint32_t steps) {
    if (!s || steps < 1 || steps > MAX_ROLLOUT)
        return tpt_set_error(TPT_EINVAL, "rollout steps must be 1 .. %d", MAX_ROLLOUT);
    s->rollout = steps;
    return TPT_OK;
}

extern "C" int tpt_sur_set_clip(tpt_sur *s, float clip) {
    if (!s || !(clip >= 0.0f)) return tpt_set_error(TPT_EINVAL, "bad clip");
    s->clip = clip;
    return TPT_OK;
}

// Inference with the glue kernels folded into the convolutions (1, the
// default) or as separate kernels (0); both give the same bits.
extern "C" int tpt_sur_set_fused(tpt_sur *s, int fused) {
    if (!s) return tpt_set_error(TPT_EINVAL, "null surrogate");
    s->fused = fused != 0;
    return TPT_OK;
}

// Loss of the current parameters on dataset pairs, without training.
extern "C" int tpt_sur_eval_loss(tpt_sur *s, const uint32_t *pairs, uint32_t n_pairs, float *mean_loss) {
    if (!s || !pairs || !n_pairs || !mean_loss) return tpt_set_error(TPT_EINVAL, "bad arguments");
    if (!s->data) return tpt_set_error(TPT_ESTATE, "no dataset");
    const int n = (int)cells_at(s, 0);
    hipStream_t st = s->stream;
    hipLaunchKernelGGL(sur_zero_f, dim3(1), dim3(64), 0, st, s->loss_acc, 4);
    for (uint32_t i = 0; i < n_pairs; ++i) {
        const uint32_t tr = pairs[2 * i], sn = pairs[2 * i + 1];
        if (tr >= s->n_traj || sn + 1 >= s->n_snap) return tpt_set_error(TPT_EINVAL, "pair %u out of range", i);
        const Planes cur = snap_planes(s, tr, sn), nxt = snap_planes(s, tr, sn + 1);
        hipLaunchKernelGGL(sur_prep, grid1(n), dim3(256), 0, st, cur, s->walls + (size_t)tr * n, nullptr,
                           s->srcs + (size_t)tr * n, s->x0, s->x0lo, n, s->in_scale[0], s->in_scale[1], s->in_scale[3], s->ambient);
        forward(s, st, false);
        hipLaunchKernelGGL(sur_loss, grid1(n), dim3(256), 0, st, s->out, cur, nxt, s->g_out, s->red, n,
                           s->scale[0], s->scale[1], s->scale[3]);
        hipLaunchKernelGGL(sum_blocks, dim3(1), dim3(256), 0, st, s->red, (int)grid1(n).x, 1, s->loss_acc, 1);
    }
    float acc[4];
    TPT_HIP_CHECK(hipMemcpyAsync(acc, s->loss_acc, sizeof acc, hipMemcpyDeviceToHost, st));
    TPT_HIP_CHECK(hipStreamSynchronize(st));
    *mean_loss = acc[0] / (float)n_pairs;
    return tpt_check_launch("surrogate evaluation");
}

// 8 ticks of the context's air grids by the network.  With `check` set, the
// prediction is refused when any value is not finite or when the energy
// sum(p^2 + vx^2 + vy^2) grows beyond `energy_ratio` times its value at the
// call's start; the call then restores its input and runs 8 plain air ticks.
// *fell_back (nullable) reports which ran.
extern "C" int tpt_sur_step_air(tpt_sur *s, tpt_ctx *ctx, int check, float energy_ratio, int *fell_back) {
    int rc = check_ctx(s, ctx);
    if (rc) return rc;
    const int n = (int)cells_at(s, 0);
    hipStream_t st = ctx->stream;
    Planes p = ctx_planes(ctx);
    StageIn prep;
    prep.stage = STAGE_PREP;
    prep.pv = p.pv;
    prep.vx = p.vx;
    prep.vy = p.vy;
    prep.hv = p.hv;
    prep.src = s->src;
    prep.wall = tpt_buf<uint32_t>(ctx, TPT_BUF_BLOCK_AIR);
    forward(s, st, false, &prep);
    hipLaunchKernelGGL(sur_apply, grid1(n), dim3(256), 0, st, s->out, tpt_buf<float>(ctx, TPT_BUF_AIR_PV),
                       tpt_buf<float>(ctx, TPT_BUF_AIR_VX), tpt_buf<float>(ctx, TPT_BUF_AIR_VY),
                       tpt_buf<float>(ctx, TPT_BUF_AIR_HV), s->src, check ? s->red : nullptr,
                       check ? s->saved : nullptr, n, s->scale[0], s->scale[1], s->scale[3]);
    if (check)
        hipLaunchKernelGGL(sum_blocks, dim3(1), dim3(256), 0, st, s->red, (int)grid1(n).x, 3, s->loss_acc + 1, 0);
    if ((rc = tpt_check_launch("surrogate step"))) return rc;
    ctx->maps_fresh = ctx->rec_cover = false;
    int fb = 0;
    if (check) {
        float *acc = s->check_host;  // pinned: the copy needs no staging
        TPT_HIP_CHECK(hipMemcpyAsync(acc, s->loss_acc, 4 * sizeof(float), hipMemcpyDeviceToHost, st));
        TPT_HIP_CHECK(hipStreamSynchronize(st));
        if (acc[3] != 0.0f || !(acc[2] <= energy_ratio * acc[1] + 1e-3f * (float)n)) {
            fb = 1;
            hipLaunchKernelGGL(sur_restore, grid1(n), dim3(256), 0, st, s->saved,
                               tpt_buf<float>(ctx, TPT_BUF_AIR_PV), tpt_buf<float>(ctx, TPT_BUF_AIR_VX),
                               tpt_buf<float>(ctx, TPT_BUF_AIR_VY), tpt_buf<float>(ctx, TPT_BUF_AIR_HV), n);
            if ((rc = plain_ticks(s, ctx, TICKS_PER_CALL))) return rc;
        }
    }
    if (fell_back) *fell_back = fb;
    return TPT_OK;
}

// The last call's energy sums (before, after) and non-finite count.
extern "C" int tpt_sur_last_check(tpt_sur *s, float out[3]) {
    if (!s || !out) return tpt_set_error(TPT_EINVAL, "null argument");
    float acc[4];
    TPT_HIP_CHECK(hipMemcpy(acc, s->loss_acc, sizeof acc, hipMemcpyDeviceToHost));
    out[0] = acc[1];
    out[1] = acc[2];
    out[2] = acc[3];
    return TPT_OK;
}
