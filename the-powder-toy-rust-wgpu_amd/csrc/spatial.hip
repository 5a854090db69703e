// RecalcFreeParticles for gfx950: SimBuffers::recalc_free_particles_device
// (crates/tpt-sim/src/buffers/spatial.rs:371-472).
//
//   recalc_clear     zero pmap / photons / pmap_count (16-byte stores)
//   recalc_scatter   life decay + kills, map scatter, stack count.  The
//                    reference's last-writer store on a shared pixel becomes
//                    atomicMax on the packed word, which is the highest slot,
//                    the serial order's last writer.  INVIS/FILT (which only
//                    take an empty pixel) are queued instead.
//   recalc_invis     the queued INVIS/FILT particles take their pixel only if
//                    no other matter holds it; among several the lowest slot
//                    wins (the first serial writer), by a CAS-min loop.
//   stacking_mark / stacking_apply   CheckStacking (spatial.rs:190-275).

#include "tpt_internal.h"

namespace {

constexpr int BLK = 256;

__device__ __forceinline__ uint32_t lowbias_hash(uint32_t seed) {
    uint32_t h = seed ^ 0x9E3779B9u;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    return h ^ (h >> 16);
}

__global__ void recalc_clear(uint4 *__restrict__ pmap, uint4 *__restrict__ photons, uint4 *__restrict__ count,
                             uint32_t n4, uint32_t *__restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) counters[0] = 0u;
    if (i >= n4) return;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    pmap[i] = z;
    photons[i] = z;
    count[i] = z;
}

__global__ __launch_bounds__(BLK) void recalc_scatter(
    uint32_t *__restrict__ ty_a, int32_t *__restrict__ life_a, const float *__restrict__ px_a,
    const float *__restrict__ py_a, const uint32_t *__restrict__ elem_props, const uint32_t *__restrict__ wall_bmap,
    const uint32_t *__restrict__ wall_emap, uint32_t *__restrict__ pmap, uint32_t *__restrict__ photons,
    uint32_t *__restrict__ pmap_count, uint32_t *__restrict__ invis_list, uint32_t *__restrict__ counters,
    uint32_t slots, int xres, int yres, int xc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    uint32_t ty = ty_a[i];
    if (ty == 0u || ty >= (uint32_t)TPT_PT_NUM) return;
    const int x = tpt_f2i(px_a[i] + 0.5f), y = tpt_f2i(py_a[i] + 0.5f);
    const bool in_bounds = x >= 0 && y >= 0 && x < xres && y < yres;
    bool stasis = false;
    if (in_bounds) {
        const int cell = (y / TPT_CELL) * xc + x / TPT_CELL;
        stasis = (wall_bmap[cell] & 0xFFu) == 18u && (wall_emap[cell] & 0xFFu) < 8u;
    }
    const uint32_t props = elem_props[ty * 32u + 1u];
    const int32_t life = life_a[i];
    if (life > 0 && (props & 0x4000u) && (!in_bounds || !stasis)) {
        const int32_t next = life - 1;
        life_a[i] = next;
        if (next <= 0 && (props & (0x10000u | 0x8000u))) {
            ty_a[i] = 0u;
            ty = 0u;
        }
    } else if (life <= 0 && (props & 0x8000u) && (!in_bounds || !stasis)) {
        ty_a[i] = 0u;
        ty = 0u;
    }
    if (ty == 0u || !in_bounds) return;
    const uint32_t idx = (uint32_t)y * (uint32_t)xres + (uint32_t)x;
    const uint32_t packed = ((i + 1u) << 9) | (ty & 0x1FFu);
    if (props & 0x10u) {
        atomicMax(&photons[idx], packed);
        return;
    }
    if (ty == 124u || ty == 134u) {  // INVIS / FILT
        invis_list[atomicAdd(&counters[0], 1u)] = i;
    } else {
        atomicMax(&pmap[idx], packed);
    }
    if (ty != 96u && ty != 153u && ty != 159u && ty != 28u) atomicAdd(&pmap_count[idx], 1u);
}

__global__ __launch_bounds__(BLK) void recalc_invis(const uint32_t *__restrict__ ty_a, const float *__restrict__ px_a,
                                                    const float *__restrict__ py_a, uint32_t *__restrict__ pmap,
                                                    const uint32_t *__restrict__ invis_list,
                                                    const uint32_t *__restrict__ counters, int xres) {
    const uint32_t n = counters[0];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t i = invis_list[k];
        const uint32_t ty = ty_a[i];
        const int x = tpt_f2i(px_a[i] + 0.5f), y = tpt_f2i(py_a[i] + 0.5f);
        const uint32_t idx = (uint32_t)y * (uint32_t)xres + (uint32_t)x;
        const uint32_t packed = ((i + 1u) << 9) | (ty & 0x1FFu);
        uint32_t cur = pmap[idx];
        for (;;) {
            const uint32_t t = cur & 0x1FFu;
            const bool other_matter = cur != 0u && t != 124u && t != 134u;
            if (other_matter || (cur != 0u && cur <= packed)) break;
            const uint32_t prev = atomicCAS(&pmap[idx], cur, packed);
            if (prev == cur) break;
            cur = prev;
        }
    }
}

__global__ void stacking_mark(uint32_t *__restrict__ pmap_count, const uint32_t *__restrict__ wall_bmap, uint32_t n,
                              uint32_t xres, uint32_t xc, uint32_t npart, uint32_t seed) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const uint32_t count = pmap_count[idx];
    if (count <= 5u) return;
    const uint32_t x = idx % xres, y = idx / xres;
    const uint32_t wall = wall_bmap[(y / TPT_CELL) * xc + x / TPT_CELL] & 0xFFu;
    if (wall == 12u) {
        if (count > 1500u) pmap_count[idx] = count + npart;
    } else {
        const uint32_t roll = lowbias_hash(seed ^ (idx * 0x2545F491u)) % 1600u;
        if (count > 1500u || roll <= count + 100u) pmap_count[idx] = count + npart;
    }
}

__global__ __launch_bounds__(BLK) void stacking_apply(uint32_t *__restrict__ ty_a, const float *__restrict__ px_a,
                                                      const float *__restrict__ py_a, float *__restrict__ temp_a,
                                                      int32_t *__restrict__ tmp_a, uint32_t *__restrict__ pmap,
                                                      uint32_t *__restrict__ pmap_count,
                                                      const uint32_t *__restrict__ elem_props, uint32_t slots, int xres,
                                                      int yres, uint32_t npart) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint32_t ty = ty_a[i];
    if (ty == 0u || ty >= (uint32_t)TPT_PT_NUM) return;
    const int x = tpt_f2i(px_a[i] + 0.5f), y = tpt_f2i(py_a[i] + 0.5f);
    if (x < 0 || y < 0 || x >= xres || y >= yres) return;
    if (elem_props[ty * 32u + 1u] & 0x10u) return;
    const uint32_t idx = (uint32_t)y * (uint32_t)xres + (uint32_t)x;
    const uint32_t count = __hip_atomic_load(&pmap_count[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (count < npart) return;
    const uint32_t packed = pmap[idx];
    const uint32_t owner = packed ? (packed >> 9) - 1u : 0xFFFFFFFFu;
    if (count > npart && owner == i) {
        ty_a[i] = 40u;  // PT_NBHL
        temp_a[i] = 9999.0f;
        uint32_t strength = count - npart;
        if (strength > 51200u) strength = 51200u;
        tmp_a[i] = (int32_t)strength;
        __hip_atomic_store(&pmap_count[idx], npart, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pmap[idx] = ((i + 1u) << 9) | 40u;
    } else {
        ty_a[i] = 0u;
    }
}

}  // namespace

int tpt_recalc_step(tpt_ctx *ctx, uint32_t seed, int stacking) {
    const uint32_t P = ctx->pixels, slots = ctx->slot_bound;
    const int xres = (int)ctx->cfg.xres, yres = (int)ctx->cfg.yres;
    uint32_t *pmap = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP), *photons = tpt_buf<uint32_t>(ctx, TPT_BUF_PHOTONS),
             *count = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP_COUNT);
    tpt_pass_begin(ctx, PASS_RECALC_CLEAR);
    // P is a multiple of 16 (xres and yres are multiples of 4).
    const uint32_t n4 = P / 4;
    hipLaunchKernelGGL(recalc_clear, dim3((n4 + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream, (uint4 *)pmap,
                       (uint4 *)photons, (uint4 *)count, n4, ctx->counters);
    tpt_pass_end(ctx, PASS_RECALC_CLEAR);
    tpt_pass_begin(ctx, PASS_RECALC_SCATTER);
    hipLaunchKernelGGL(recalc_scatter, dim3((slots + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE),
                       tpt_buf<float>(ctx, TPT_BUF_PART_X), tpt_buf<float>(ctx, TPT_BUF_PART_Y),
                       tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS), tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP),
                       tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP), pmap, photons, count, ctx->slot_scratch_u,
                       ctx->counters, slots, xres, yres, (int)ctx->xcells);
    hipLaunchKernelGGL(recalc_invis, dim3(64), dim3(BLK), 0, ctx->stream, tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY),
                       tpt_buf<float>(ctx, TPT_BUF_PART_X), tpt_buf<float>(ctx, TPT_BUF_PART_Y), pmap,
                       ctx->slot_scratch_u, ctx->counters, xres);
    tpt_pass_end(ctx, PASS_RECALC_SCATTER);
    int rc = tpt_check_launch("recalc_free_particles");
    if (rc || !stacking) return rc;
    const uint32_t npart = ctx->cfg.xres * ctx->cfg.yres;
    tpt_pass_begin(ctx, PASS_STACKING);
    hipLaunchKernelGGL(stacking_mark, dim3((P + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream, count,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP), P, ctx->cfg.xres, ctx->xcells, npart, seed);
    hipLaunchKernelGGL(stacking_apply, dim3((slots + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<float>(ctx, TPT_BUF_PART_X),
                       tpt_buf<float>(ctx, TPT_BUF_PART_Y), tpt_buf<float>(ctx, TPT_BUF_PART_TEMP),
                       tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP), pmap, count, tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS),
                       slots, xres, yres, npart);
    tpt_pass_end(ctx, PASS_STACKING);
    return tpt_check_launch("check_stacking");
}
