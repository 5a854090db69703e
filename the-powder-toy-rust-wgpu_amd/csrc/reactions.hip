// The part of reaction_kernel::step_reactions_gpu
// (crates/tpt-sim/src/reaction_kernel/mod.rs:1123-1237) this library runs:
// the modext clear of the gravity mass-override channel
// (reaction_kernel/modext/mod.rs:53-99) and the life pass
// (life_and_local_reaction_kernel, reaction_kernel/life.rs:47-205): fire,
// plasma and TRON burn-in lifetimes, fire burning out to smoke or steam,
// plasma to noble gas or steam, sparks settling back into their conductor,
// the element kill at zero life, and the stasis flag.  Each slot reads and
// writes only its own state, so one thread per slot reproduces the
// reference exactly.

#include "tpt_internal.h"

namespace {

constexpr int BLK = 256;
constexpr uint32_t T_WATR = 2, T_FIRE = 4, T_METL = 11, T_SPRK = 12, T_SMKE = 16, T_SLTW = 19, T_WTRV = 21,
                   T_PLSM = 28, T_PSCN = 50, T_NSCN = 51, T_ETRD = 55, T_SWCH = 60, T_INWR = 63, T_NBLE = 97,
                   T_HFLM = 100, T_TRON = 151, T_RSST = 191;
constexpr uint32_t PROP_LIFE_KILL = 0x8000u, FLAG_STASIS = 0x20u;

// hash_u32 (reaction_kernel/shared.rs:1015-1023), MurmurHash3's finaliser.
__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    uint32_t h = x;
    h ^= h >> 16;
    h *= 2246822507u;
    h ^= h >> 13;
    h *= 3266489909u;
    h ^= h >> 16;
    return h;
}

struct LifeArgs {
    uint32_t *ty;
    int32_t *life, *tmp;
    uint32_t *ctype, *flags;
    float *temp;
    const float *x, *y;
    const uint32_t *wall, *emap, *props, *counters;
    uint32_t slots;
    int xres, yres, xc, yc;
};

__global__ __launch_bounds__(BLK) void rx_life(LifeArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.slots) return;
    const uint32_t ty = a.ty[i];
    if (ty == 0u) return;
    const uint32_t seed = a.counters[TPT_CTR_SEED];
    const int xi = tpt_f2i(a.x[i] + 0.5f), yi = tpt_f2i(a.y[i] + 0.5f);
    bool frozen = false;
    if (xi >= 0 && yi >= 0 && xi < a.xres && yi < a.yres) {
        const int cell = (yi / TPT_CELL) * a.xc + xi / TPT_CELL;
        frozen = a.wall[cell] == 18u && (a.emap[cell] & 0xFFu) < 8u;
    }
    const uint32_t flags = a.flags[i];
    if (frozen) {
        if (!(flags & FLAG_STASIS)) a.flags[i] = flags | FLAG_STASIS;
        return;
    }
    if (flags & FLAG_STASIS) a.flags[i] = flags & ~FLAG_STASIS;
    const uint32_t props = ty < 256u ? a.props[ty * 32u + 1u] : 0u;
    const int32_t life0 = a.life[i];
    int32_t life = life0;
    uint32_t new_ty = ty;
    bool dead = false;
    if (ty == T_FIRE && life == 0) life = 180 + (int32_t)(hash_u32(i ^ seed) % 80u);
    else if ((ty == T_PLSM || ty == T_HFLM) && life == 0) life = 50 + (int32_t)(hash_u32(i ^ seed ^ 0xA5A5u) % 150u);
    else if (ty == T_TRON && life == 0) life = 25 + (int32_t)(hash_u32(i ^ seed ^ 0x3C3Cu) % 25u);
    if (ty == T_FIRE) {
        if (life <= 1) {
            if ((a.tmp[i] & 3) == 3) {
                new_ty = T_WTRV;
                life = 0;
                a.ctype[i] = T_FIRE;
            } else if (a.temp[i] < 625.0f) {
                new_ty = T_SMKE;
                life = 250 + (int32_t)(hash_u32(i ^ seed ^ 0x5ACE123u) % 20u);
                a.tmp[i] = 0;
            } else {
                dead = true;
            }
        }
    } else if (ty == T_PLSM) {
        if (life <= 1) {
            const uint32_t ct = a.ctype[i];
            if (ct == T_NBLE) {
                new_ty = T_NBLE;
                life = 0;
                a.ctype[i] = 0u;
            } else if ((a.tmp[i] & 3) == 3) {
                new_ty = T_WTRV;
                life = 0;
                a.ctype[i] = T_FIRE;
            } else {
                dead = true;
            }
        }
    } else if (ty == T_SPRK) {
        if (life <= 0) {
            uint32_t ct = a.ctype[i];
            if (ct == T_WATR || ct == T_SLTW || ct == T_PSCN || ct == T_NSCN || ct == T_ETRD || ct == T_INWR)
                a.temp[i] = 293.15f;
            if (ct == 0u || ct >= 256u) ct = T_METL;
            a.ctype[i] = 0u;
            life = 4;
            if (ct == T_WATR) life = 64;
            else if (ct == T_SLTW) life = 54;
            else if (ct == T_SWCH) life = 14;
            else if (ct == T_RSST) dead = true;
            if (!dead) new_ty = ct;
        }
    }
    if (new_ty == ty && (props & PROP_LIFE_KILL) && life <= 0) dead = true;
    if (dead) {
        a.ty[i] = 0u;
    } else {
        if (new_ty != ty) a.ty[i] = new_ty;
        if (life != life0) a.life[i] = life;
    }
}

}  // namespace

int tpt_reactions_step(tpt_ctx *ctx) {
    int rc = tpt_gravity_override_clear(ctx);
    if (rc) return rc;
    const uint32_t slots = ctx->slot_bound;
    LifeArgs a{tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY),    tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE),
               tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP),    tpt_buf<uint32_t>(ctx, TPT_BUF_PART_CTYPE),
               tpt_buf<uint32_t>(ctx, TPT_BUF_PART_FLAGS), tpt_buf<float>(ctx, TPT_BUF_PART_TEMP),
               tpt_buf<float>(ctx, TPT_BUF_PART_X),        tpt_buf<float>(ctx, TPT_BUF_PART_Y),
               tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP),  tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP),
               tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS), ctx->counters,
               slots,                                      (int)ctx->cfg.xres,
               (int)ctx->cfg.yres,                         (int)ctx->xcells,
               (int)ctx->ycells};
    tpt_pass_begin(ctx, PASS_REACTION_LIFE);
    hipLaunchKernelGGL(rx_life, dim3((slots + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream, a);
    tpt_pass_end(ctx, PASS_REACTION_LIFE);
    return tpt_check_launch("rx_life");
}
