// RecalcFreeParticles for gfx950: SimBuffers::recalc_free_particles_device
// (crates/tpt-sim/src/buffers/spatial.rs:371-472).
//
//   recalc_clear     zero the pixel records / photons / pmap_count (16-byte
//                    stores); pmap itself is rewritten whole by recalc_extract
//   recalc_scatter   life decay + kills, map scatter, stack count.  The
//                    reference's last-writer store on a shared pixel must
//                    leave the highest slot (the serial order's last writer)
//                    on the pixel record ((slot+1)<<9 | flag<<8 | type) << 32
//                    | temp bits, whose high word orders by slot.  Every slot
//                    first stores its record with a plain store (any one of
//                    a shared pixel's writers survives); recalc_check lists
//                    the slots that find another slot's record on their
//                    pixel, and recalc_resolve runs only those through a
//                    64-bit atomicMax, whose old value tells who lost and
//                    is counted and listed.  The surviving plain store acts
//                    as the first atomic.  (A device-scope atomic executes
//                    past the L2 of its XCD; one per particle cost more
//                    than the plain stores and the check's loads together.)
//                    INVIS/FILT (which only take an empty pixel) are queued
//                    instead.
//   recalc_invis     the queued INVIS/FILT particles take their pixel only if
//                    no other matter holds it; among several the lowest slot
//                    wins (the first serial writer), by a CAS-min loop.
//   recalc_extract   pmap = record high words without the flag; pmap_count
//                    gains the owner's own +1 (losers counted themselves).
//   stacking_mark / stacking_apply   CheckStacking (spatial.rs:190-275).
//
// The pixel record is library-internal: it lets the conduction pass run as
// a pixel-major stencil.  Its flag bit says "this owner is handled through
// the off-map list": every live slot whose rounded or truncated pixel is on
// the canvas ends up either as the unflagged owner of that pixel or in the
// list, never both.  The list holds energy particles, INVIS/FILT, particles
// that lose their pixel to a higher slot, and particles whose rounded and
// truncated pixels differ; per-slot consumers re-check their own bounds.
//
// The movement pass rebuilds the same records, list and a private count map
// after its commit (tpt_move_rebuild_maps).  Until the next
// RecalcFreeParticles only air and gravity run, which move no particle, so
// that recalc is incremental (recalc_decay / recalc_relist /
// recalc_invis_list): life decay kills, a killed owner clears its record and
// every listed particle is offered its pixel again; pmap_count is the
// private map minus the killed.

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

// Appends `v` to a list for every lane where `pred` holds, one atomic per
// wave.  Every lane of the wave must call it.
__device__ __forceinline__ void wave_append(uint32_t *__restrict__ list, uint32_t *__restrict__ count, uint32_t v,
                                            bool pred) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = __ffsll((unsigned long long)m) - 1u;
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (pred) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
}

__device__ __forceinline__ bool counted_type(uint32_t t) { return t != 96u && t != 153u && t != 159u && t != 28u; }

__global__ void recalc_clear(uint4 *__restrict__ rec, uint4 *__restrict__ photons, uint4 *__restrict__ count,
                             uint32_t n4, uint32_t *__restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        counters[TPT_CTR_INVIS] = 0u;
        counters[TPT_CTR_OFFMAP] = 0u;
        counters[TPT_CTR_CONFLICT] = 0u;
    }
    if (i >= n4) return;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    rec[2 * i] = z;
    rec[2 * i + 1] = z;
    photons[i] = z;
    count[i] = z;
}

__global__ __launch_bounds__(BLK) void recalc_scatter(
    uint32_t *__restrict__ ty_a, int32_t *__restrict__ life_a, const float *__restrict__ px_a,
    const float *__restrict__ py_a, const float *__restrict__ temp_a, const uint32_t *__restrict__ elem_props,
    const uint32_t *__restrict__ wall_bmap, const uint32_t *__restrict__ wall_emap, unsigned long long *__restrict__ rec,
    uint32_t *__restrict__ photons, uint32_t *__restrict__ pmap_count, uint32_t *__restrict__ invis_list,
    uint32_t *__restrict__ offmap, uint32_t *__restrict__ counters, uint32_t slots, int xres, int yres, int xc,
    uint32_t decay) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool app_invis = false, app_self = false;
    uint32_t ty = i < slots ? ty_a[i] : 0u;
    if (ty != 0u && ty < (uint32_t)TPT_PT_NUM) {
        const float fx = px_a[i], fy = py_a[i];
        const int x = tpt_f2i(fx + 0.5f), y = tpt_f2i(fy + 0.5f);
        const bool in_bounds = x >= 0 && y >= 0 && x < xres && y < yres;
        const uint32_t props = elem_props[ty * 32u + 1u];
        const int32_t life = decay ? life_a[i] : 0;
        // Life decay and kills, unless the particle sits in an unpowered
        // stasis wall cell; the wall grids are read only when it matters.
        const bool dec = decay && life > 0 && (props & 0x4000u);
        const bool kill = decay && !dec && life <= 0 && (props & 0x8000u);
        bool stasis = false;
        if ((dec || kill) && in_bounds) {
            const int cell = (y / TPT_CELL) * xc + x / TPT_CELL;
            stasis = (wall_bmap[cell] & 0xFFu) == 18u && (wall_emap[cell] & 0xFFu) < 8u;
        }
        if (dec && !stasis) {
            const int32_t next = life - 1;
            life_a[i] = next;
            if (next <= 0 && (props & (0x10000u | 0x8000u))) {
                ty_a[i] = 0u;
                ty = 0u;
            }
        } else if (kill && !stasis) {
            ty_a[i] = 0u;
            ty = 0u;
        }
        // The pixel the per-slot thermal kernels use (truncated position).
        const int ox = tpt_f2i(fx), oy = tpt_f2i(fy);
        const bool own_in = ox >= 0 && oy >= 0 && ox < xres && oy < yres;
        const bool same_px = ox == x && oy == y;
        if (ty != 0u && !in_bounds) {
            app_self = own_in;
        } else if (ty != 0u) {
            const uint32_t idx = (uint32_t)y * (uint32_t)xres + (uint32_t)x;
            const uint32_t packed = ((i + 1u) << 9) | (ty & 0x1FFu);
            if (props & 0x10u) {
                photons[idx] = packed;
                app_self = true;
            } else if (ty == 124u || ty == 134u) {  // INVIS / FILT
                app_invis = true;
                app_self = true;
                atomicAdd(&pmap_count[idx], 1u);
            } else {
                const uint32_t key = packed | (same_px ? 0u : 0x100u);
                rec[idx] = ((unsigned long long)key << 32) | (unsigned long long)__float_as_uint(temp_a[i]);
                if (!same_px) app_self = true;
            }
        }
    }
    wave_append(invis_list, &counters[TPT_CTR_INVIS], i, app_invis);
    wave_append(offmap, &counters[TPT_CTR_OFFMAP], i, app_self);
}

// Where slot i's scatter store went: kind 0 nowhere (empty, off the canvas,
// INVIS/FILT), 1 the photon map, 2 the pixel record; `key` is the stored
// word (the record's high word).
struct MapStore {
    uint32_t kind, idx, key;
};

__device__ __forceinline__ MapStore map_store(uint32_t i, uint32_t ty, float fx, float fy,
                                              const uint32_t *__restrict__ elem_props, int xres, int yres) {
    MapStore m{0u, 0u, 0u};
    if (ty == 0u || ty >= (uint32_t)TPT_PT_NUM || ty == 124u || ty == 134u) return m;
    const int x = tpt_f2i(fx + 0.5f), y = tpt_f2i(fy + 0.5f);
    if (x < 0 || y < 0 || x >= xres || y >= yres) return m;
    m.idx = (uint32_t)y * (uint32_t)xres + (uint32_t)x;
    m.key = ((i + 1u) << 9) | ty;
    if (elem_props[ty * 32u + 1u] & 0x10u) {
        m.kind = 1u;
    } else {
        m.kind = 2u;
        if (tpt_f2i(fx) != x || tpt_f2i(fy) != y) m.key |= 0x100u;
    }
    return m;
}

// Lists the slots whose scatter store was overwritten by another slot's.
__global__ __launch_bounds__(BLK) void recalc_check(const uint32_t *__restrict__ ty_a, const float *__restrict__ px_a,
                                                    const float *__restrict__ py_a,
                                                    const uint32_t *__restrict__ elem_props,
                                                    const unsigned long long *__restrict__ rec,
                                                    const uint32_t *__restrict__ photons,
                                                    uint32_t *__restrict__ conflicts, uint32_t *__restrict__ counters,
                                                    uint32_t slots, int xres, int yres) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool clash = false;
    const uint32_t ty = i < slots ? ty_a[i] : 0u;
    if (ty != 0u) {
        const MapStore m = map_store(i, ty, px_a[i], py_a[i], elem_props, xres, yres);
        if (m.kind == 1u) clash = photons[m.idx] != m.key;
        else if (m.kind == 2u) clash = (uint32_t)(rec[m.idx] >> 32) != m.key;
    }
    wave_append(conflicts, &counters[TPT_CTR_CONFLICT], i, clash);
}

// The listed slots take their pixel by atomicMax; whoever the old value
// shows lost (this slot, or the holder it displaced) is counted on the
// pixel and listed off the map.
__global__ __launch_bounds__(BLK) void recalc_resolve(const uint32_t *__restrict__ ty_a,
                                                      const float *__restrict__ px_a,
                                                      const float *__restrict__ py_a,
                                                      const float *__restrict__ temp_a,
                                                      const uint32_t *__restrict__ elem_props,
                                                      unsigned long long *__restrict__ rec,
                                                      uint32_t *__restrict__ photons,
                                                      uint32_t *__restrict__ pmap_count,
                                                      const uint32_t *__restrict__ conflicts,
                                                      uint32_t *__restrict__ offmap, uint32_t *__restrict__ counters,
                                                      int xres, int yres) {
    const uint32_t n = counters[TPT_CTR_CONFLICT];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t i = conflicts[k];
        const uint32_t ty = ty_a[i];
        const MapStore m = map_store(i, ty, px_a[i], py_a[i], elem_props, xres, yres);
        if (m.kind == 1u) {
            atomicMax(&photons[m.idx], m.key);
            continue;
        }
        const unsigned long long v =
            ((unsigned long long)m.key << 32) | (unsigned long long)__float_as_uint(temp_a[i]);
        const uint32_t okey = (uint32_t)(atomicMax(&rec[m.idx], v) >> 32);
        if (okey > m.key) {
            // Lost to a higher slot: count self, list self (a flagged slot
            // is listed already).
            if (counted_type(ty)) atomicAdd(&pmap_count[m.idx], 1u);
            if ((m.key & 0x100u) == 0u) offmap[atomicAdd(&counters[TPT_CTR_OFFMAP], 1u)] = i;
        } else if (okey != 0u) {
            // Displaced the previous holder: count it, list it.
            if (counted_type(okey & 0xFFu)) atomicAdd(&pmap_count[m.idx], 1u);
            if ((okey & 0x100u) == 0u) offmap[atomicAdd(&counters[TPT_CTR_OFFMAP], 1u)] = (okey >> 9) - 1u;
        }
    }
}

__global__ __launch_bounds__(BLK) void recalc_invis(const uint32_t *__restrict__ ty_a, const float *__restrict__ px_a,
                                                    const float *__restrict__ py_a, const float *__restrict__ temp_a,
                                                    unsigned long long *__restrict__ rec,
                                                    const uint32_t *__restrict__ invis_list,
                                                    const uint32_t *__restrict__ counters, int xres) {
    const uint32_t n = counters[TPT_CTR_INVIS];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t i = invis_list[k];
        const uint32_t ty = ty_a[i];
        const int x = tpt_f2i(px_a[i] + 0.5f), y = tpt_f2i(py_a[i] + 0.5f);
        const uint32_t idx = (uint32_t)y * (uint32_t)xres + (uint32_t)x;
        const uint32_t packed = ((i + 1u) << 9) | (ty & 0x1FFu);
        // INVIS/FILT owners are always handled through the off-map list.
        const unsigned long long v = ((unsigned long long)(packed | 0x100u) << 32) |
                                     (unsigned long long)__float_as_uint(temp_a[i]);
        unsigned long long cur = rec[idx];
        for (;;) {
            const uint32_t ck = (uint32_t)(cur >> 32);
            const uint32_t t = ck & 0xFFu;
            const bool other_matter = ck != 0u && t != 124u && t != 134u;
            if (other_matter || (ck != 0u && (ck & ~0x100u) <= packed)) break;
            const unsigned long long prev = atomicCAS(&rec[idx], cur, v);
            if (prev == cur) break;
            cur = prev;
        }
    }
}

// pmap and the owners' own pmap_count contribution from the pixel records,
// four pixels per thread.
// With `count_src` (the incremental recalc) pmap_count is that private
// count map, already net of the killed, instead of gaining the owners.
__global__ void recalc_extract(const ulonglong2 *__restrict__ rec, uint4 *__restrict__ pmap, uint4 *__restrict__ count,
                               const uint4 *__restrict__ count_src, uint32_t n4) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    if (count_src) count[i] = count_src[i];
    const ulonglong2 a = rec[2 * i], b = rec[2 * i + 1];
    const uint32_t k[4] = {(uint32_t)(a.x >> 32), (uint32_t)(a.y >> 32), (uint32_t)(b.x >> 32), (uint32_t)(b.y >> 32)};
    uint32_t add[4];
    bool any = false;
    for (int j = 0; j < 4; ++j) {
        const uint32_t t = k[j] & 0xFFu;
        add[j] = (k[j] != 0u && t != 124u && t != 134u && counted_type(t)) ? 1u : 0u;
        any |= add[j] != 0u;
    }
    pmap[i] = make_uint4(k[0] & ~0x100u, k[1] & ~0x100u, k[2] & ~0x100u, k[3] & ~0x100u);
    if (any && count && !count_src) {
        uint4 c = count[i];
        c.x += add[0];
        c.y += add[1];
        c.z += add[2];
        c.w += add[3];
        count[i] = c;
    }
}

// Incremental RecalcFreeParticles, step 1: life decay and kills (the
// decay rule of recalc_scatter); a killed particle leaves its pixel record
// or photons entry if it held it, and the count of its pixel.
__global__ __launch_bounds__(BLK) void recalc_decay(
    uint32_t *__restrict__ ty_a, int32_t *__restrict__ life_a, const float *__restrict__ px_a,
    const float *__restrict__ py_a, const uint32_t *__restrict__ elem_props, const uint32_t *__restrict__ wall_bmap,
    const uint32_t *__restrict__ wall_emap, unsigned long long *__restrict__ rec, uint32_t *__restrict__ photons,
    uint32_t *__restrict__ pmap_count, uint32_t slots, int xres, int yres, int xc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint32_t ty = ty_a[i];
    if (ty == 0u || ty >= (uint32_t)TPT_PT_NUM) return;
    const uint32_t props = elem_props[ty * 32u + 1u];
    if ((props & (0x4000u | 0x8000u)) == 0u) return;
    // Neither rule can apply (a LIFE_DEC particle with no life left and no
    // LIFE_KILL, the common case): nothing to do, whatever the stasis test.
    const int32_t life = life_a[i];
    if (!(life > 0 && (props & 0x4000u)) && !(life <= 0 && (props & 0x8000u))) return;
    const int x = tpt_f2i(px_a[i] + 0.5f), y = tpt_f2i(py_a[i] + 0.5f);
    const bool in_bounds = x >= 0 && y >= 0 && x < xres && y < yres;
    bool stasis = false;
    if (in_bounds) {
        const int cell = (y / TPT_CELL) * xc + x / TPT_CELL;
        stasis = (wall_bmap[cell] & 0xFFu) == 18u && (wall_emap[cell] & 0xFFu) < 8u;
    }
    bool killed = false;
    if (life > 0 && (props & 0x4000u) && (!in_bounds || !stasis)) {
        const int32_t next = life - 1;
        life_a[i] = next;
        killed = next <= 0 && (props & (0x10000u | 0x8000u));
    } else if (life <= 0 && (props & 0x8000u) && (!in_bounds || !stasis)) {
        killed = true;
    }
    if (!killed) return;
    ty_a[i] = 0u;
    if (!in_bounds) return;
    const uint32_t idx = (uint32_t)y * (uint32_t)xres + (uint32_t)x;
    if (props & 0x10u) {
        if ((photons[idx] >> 9) == i + 1u) photons[idx] = 0u;
        return;
    }
    if ((uint32_t)(rec[idx] >> 41) == i + 1u) rec[idx] = 0ull;
    if (counted_type(ty)) atomicSub(&pmap_count[idx], 1u);
}

// Step 2: every live listed particle is offered its pixel again (a no-op
// unless the holder was killed); matter takes the record with the flag set,
// since it stays on the list.  INVIS/FILT follow in recalc_invis_list.
__global__ __launch_bounds__(BLK) void recalc_relist(const uint32_t *__restrict__ ty_a, const float *__restrict__ px_a,
                                                     const float *__restrict__ py_a, const float *__restrict__ temp_a,
                                                     const uint32_t *__restrict__ elem_props,
                                                     unsigned long long *__restrict__ rec,
                                                     uint32_t *__restrict__ photons, const uint32_t *__restrict__ list,
                                                     const uint32_t *__restrict__ counters, uint32_t slots, int xres,
                                                     int yres) {
    const uint32_t n = counters[TPT_CTR_OFFMAP];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t i = list[k];
        if (i >= slots) continue;
        const uint32_t ty = ty_a[i];
        if (ty == 0u || ty >= (uint32_t)TPT_PT_NUM || ty == 124u || ty == 134u) continue;
        const int x = tpt_f2i(px_a[i] + 0.5f), y = tpt_f2i(py_a[i] + 0.5f);
        if (x < 0 || y < 0 || x >= xres || y >= yres) continue;
        const uint32_t idx = (uint32_t)y * (uint32_t)xres + (uint32_t)x;
        const uint32_t packed = ((i + 1u) << 9) | ty;
        if (elem_props[ty * 32u + 1u] & 0x10u)
            atomicMax(&photons[idx], packed);
        else
            atomicMax(&rec[idx], ((unsigned long long)(packed | 0x100u) << 32) |
                                     (unsigned long long)__float_as_uint(temp_a[i]));
    }
}

// Step 3: the listed INVIS/FILT run the placement of recalc_invis again.
__global__ __launch_bounds__(BLK) void recalc_invis_list(const uint32_t *__restrict__ ty_a,
                                                         const float *__restrict__ px_a,
                                                         const float *__restrict__ py_a,
                                                         const float *__restrict__ temp_a,
                                                         unsigned long long *__restrict__ rec,
                                                         const uint32_t *__restrict__ list,
                                                         const uint32_t *__restrict__ counters, uint32_t slots,
                                                         int xres, int yres) {
    const uint32_t n = counters[TPT_CTR_OFFMAP];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t i = list[k];
        if (i >= slots) continue;
        const uint32_t ty = ty_a[i];
        if (ty != 124u && ty != 134u) continue;
        const int x = tpt_f2i(px_a[i] + 0.5f), y = tpt_f2i(py_a[i] + 0.5f);
        if (x < 0 || y < 0 || x >= xres || y >= yres) continue;
        const uint32_t idx = (uint32_t)y * (uint32_t)xres + (uint32_t)x;
        const uint32_t packed = ((i + 1u) << 9) | ty;
        const unsigned long long v = ((unsigned long long)(packed | 0x100u) << 32) |
                                     (unsigned long long)__float_as_uint(temp_a[i]);
        unsigned long long cur = rec[idx];
        for (;;) {
            const uint32_t ck = (uint32_t)(cur >> 32);
            const uint32_t t = ck & 0xFFu;
            const bool other_matter = ck != 0u && t != 124u && t != 134u;
            if (other_matter || (ck != 0u && (ck & ~0x100u) <= packed)) break;
            const unsigned long long prev = atomicCAS(&rec[idx], cur, v);
            if (prev == cur) break;
            cur = prev;
        }
    }
}

// CheckStacking runs when the device frame says so (every 10th frame), so
// that a captured tick carries it; counters[TPT_CTR_STACK] is the flag.
__global__ void stacking_mark(uint32_t *__restrict__ pmap_count, const uint32_t *__restrict__ wall_bmap, uint32_t n,
                              uint32_t xres, uint32_t xc, uint32_t npart, const uint32_t *__restrict__ counters) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n || !counters[TPT_CTR_STACK]) return;
    const uint32_t seed = counters[TPT_CTR_SEED];
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
                                                      unsigned long long *__restrict__ rec,
                                                      uint32_t *__restrict__ pmap_count,
                                                      const uint32_t *__restrict__ elem_props, uint32_t slots, int xres,
                                                      int yres, uint32_t npart, const uint32_t *__restrict__ counters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots || !counters[TPT_CTR_STACK]) return;
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
        const uint32_t flag = (uint32_t)(rec[idx] >> 32) & 0x100u;
        rec[idx] = ((unsigned long long)(((i + 1u) << 9) | flag | 40u) << 32) |
                   (unsigned long long)__float_as_uint(9999.0f);
    } else {
        ty_a[i] = 0u;
    }
}

}  // namespace

// Full map build: clear, (decay +) scatter, INVIS/FILT placement, pmap
// extraction; counts into `count`.
static int map_build(tpt_ctx *ctx, int decay, uint32_t *count, int pass_clear, int pass_scatter) {
    const uint32_t P = ctx->pixels, slots = ctx->slot_bound;
    const int xres = (int)ctx->cfg.xres, yres = (int)ctx->cfg.yres;
    uint32_t *pmap = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP), *photons = tpt_buf<uint32_t>(ctx, TPT_BUF_PHOTONS);
    unsigned long long *rec = (unsigned long long *)ctx->pix_rec;
    // P is a multiple of 16 (xres and yres are multiples of 4).
    const uint32_t n4 = P / 4;
    tpt_pass_begin(ctx, pass_clear);
    hipLaunchKernelGGL(recalc_clear, dim3((n4 + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream, (uint4 *)rec,
                       (uint4 *)photons, (uint4 *)count, n4, ctx->counters);
    if (pass_scatter != pass_clear) {  // one timing slot for both halves otherwise
        tpt_pass_end(ctx, pass_clear);
        tpt_pass_begin(ctx, pass_scatter);
    }
    hipLaunchKernelGGL(recalc_scatter, dim3((slots + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE),
                       tpt_buf<float>(ctx, TPT_BUF_PART_X), tpt_buf<float>(ctx, TPT_BUF_PART_Y),
                       tpt_buf<float>(ctx, TPT_BUF_PART_TEMP), tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS),
                       tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP), tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP), rec,
                       photons, count, ctx->slot_scratch_u, ctx->offmap, ctx->counters, slots, xres, yres,
                       (int)ctx->xcells, decay ? 1u : 0u);
    hipLaunchKernelGGL(recalc_check, dim3((slots + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<float>(ctx, TPT_BUF_PART_X),
                       tpt_buf<float>(ctx, TPT_BUF_PART_Y), tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS),
                       (const unsigned long long *)rec, photons, ctx->map_conflicts, ctx->counters, slots, xres, yres);
    hipLaunchKernelGGL(recalc_resolve, dim3(64), dim3(BLK), 0, ctx->stream, tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY),
                       tpt_buf<float>(ctx, TPT_BUF_PART_X), tpt_buf<float>(ctx, TPT_BUF_PART_Y),
                       tpt_buf<float>(ctx, TPT_BUF_PART_TEMP), tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS), rec, photons,
                       count, ctx->map_conflicts, ctx->offmap, ctx->counters, xres, yres);
    hipLaunchKernelGGL(recalc_invis, dim3(64), dim3(BLK), 0, ctx->stream, tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY),
                       tpt_buf<float>(ctx, TPT_BUF_PART_X), tpt_buf<float>(ctx, TPT_BUF_PART_Y),
                       tpt_buf<float>(ctx, TPT_BUF_PART_TEMP), rec, ctx->slot_scratch_u, ctx->counters, xres);
    hipLaunchKernelGGL(recalc_extract, dim3((n4 + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream,
                       (const ulonglong2 *)rec, (uint4 *)pmap, (uint4 *)count, (const uint4 *)nullptr, n4);
    tpt_pass_end(ctx, pass_scatter);
    return TPT_OK;
}

// Incremental form (see the file comment), valid while rebuild_fresh.
static int map_update(tpt_ctx *ctx) {
    const uint32_t P = ctx->pixels, slots = ctx->slot_bound;
    const int xres = (int)ctx->cfg.xres, yres = (int)ctx->cfg.yres;
    uint32_t *pmap = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP), *photons = tpt_buf<uint32_t>(ctx, TPT_BUF_PHOTONS),
             *count = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP_COUNT);
    unsigned long long *rec = (unsigned long long *)ctx->pix_rec;
    const uint32_t *ty = tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY);
    const float *px = tpt_buf<float>(ctx, TPT_BUF_PART_X), *py = tpt_buf<float>(ctx, TPT_BUF_PART_Y),
                *temp = tpt_buf<float>(ctx, TPT_BUF_PART_TEMP);
    const uint32_t *props = tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS);
    const uint32_t n4 = P / 4;
    // The kills come off the movement rebuild's private count map, which
    // the extract then publishes as pmap_count.
    uint32_t *count_mv = ctx->pix_count_mv;
    tpt_pass_begin(ctx, PASS_RECALC_SCATTER);
    hipLaunchKernelGGL(recalc_decay, dim3((slots + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream,
                       tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<int32_t>(ctx, TPT_BUF_PART_LIFE), px, py,
                       props, tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP), tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_EMAP), rec,
                       photons, count_mv, slots, xres, yres, (int)ctx->xcells);
    hipLaunchKernelGGL(recalc_relist, dim3(256), dim3(BLK), 0, ctx->stream, ty, px, py, temp, props, rec, photons,
                       ctx->offmap, ctx->counters, slots, xres, yres);
    hipLaunchKernelGGL(recalc_invis_list, dim3(64), dim3(BLK), 0, ctx->stream, ty, px, py, temp, rec, ctx->offmap,
                       ctx->counters, slots, xres, yres);
    hipLaunchKernelGGL(recalc_extract, dim3((n4 + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream,
                       (const ulonglong2 *)rec, (uint4 *)pmap, (uint4 *)count, (const uint4 *)count_mv, n4);
    tpt_pass_end(ctx, PASS_RECALC_SCATTER);
    return TPT_OK;
}

int tpt_move_rebuild_maps(tpt_ctx *ctx) {
    int rc = map_build(ctx, 0, ctx->pix_count_mv, PASS_MOVE_REBUILD, PASS_MOVE_REBUILD);
    if (rc == TPT_OK) rc = tpt_check_launch("map rebuild");
    ctx->rebuild_fresh = rc == TPT_OK;
    return rc;
}

// `stacking`: CheckStacking this call (the device flag must agree); while a
// tick is captured the stacking kernels are always recorded and gate
// themselves on the device flag.
int tpt_recalc_step(tpt_ctx *ctx, int stacking) {
    const uint32_t P = ctx->pixels, slots = ctx->slot_bound;
    const int xres = (int)ctx->cfg.xres, yres = (int)ctx->cfg.yres;
    uint32_t *pmap = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP), *count = tpt_buf<uint32_t>(ctx, TPT_BUF_PMAP_COUNT);
    unsigned long long *rec = (unsigned long long *)ctx->pix_rec;
    const bool incremental = ctx->rebuild_fresh;
    ctx->maps_fresh = ctx->rec_cover = ctx->rebuild_fresh = false;
    int rc = incremental ? map_update(ctx) : map_build(ctx, 1, count, PASS_RECALC_CLEAR, PASS_RECALC_SCATTER);
    if (rc == TPT_OK) rc = tpt_check_launch("recalc_free_particles");
    if (rc) return rc;
    if (stacking || ctx->capturing) {
        const uint32_t npart = ctx->cfg.xres * ctx->cfg.yres;
        tpt_pass_begin(ctx, PASS_STACKING);
        hipLaunchKernelGGL(stacking_mark, dim3((P + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream, count,
                           tpt_buf<uint32_t>(ctx, TPT_BUF_WALL_BMAP), P, ctx->cfg.xres, ctx->xcells, npart,
                           ctx->counters);
        hipLaunchKernelGGL(stacking_apply, dim3((slots + BLK - 1) / BLK), dim3(BLK), 0, ctx->stream,
                           tpt_buf<uint32_t>(ctx, TPT_BUF_PART_TY), tpt_buf<float>(ctx, TPT_BUF_PART_X),
                           tpt_buf<float>(ctx, TPT_BUF_PART_Y), tpt_buf<float>(ctx, TPT_BUF_PART_TEMP),
                           tpt_buf<int32_t>(ctx, TPT_BUF_PART_TMP), pmap, rec, count,
                           tpt_buf<uint32_t>(ctx, TPT_BUF_ELEM_PROPS), slots, xres, yres, npart, ctx->counters);
        tpt_pass_end(ctx, PASS_STACKING);
        rc = tpt_check_launch("check_stacking");
        if (rc) return rc;
    }
    ctx->maps_fresh = ctx->rec_cover = true;
    return TPT_OK;
}
