/* Shared helpers of the CPU oracle -- TEST INFRASTRUCTURE ONLY. */
#ifndef ORACLE_UTIL_H
#define ORACLE_UTIL_H

#include <stdint.h>
#include <string.h>

/* Truncation toward zero, saturating, NaN -> 0 (i32::cast_from on f32). */
static inline int ora_f2i(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int)0x80000000u;
    return (int)f;
}

static inline uint32_t ora_f2u(float f) {
    if (!(f == f) || f <= 0.0f) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

static inline float ora_clampf(float v, float lo, float hi) {
    float out = v;
    if (out < lo) out = lo;
    if (out > hi) out = hi;
    return out;
}

static inline int32_t ora_clampi(int32_t v, int32_t lo, int32_t hi) {
    int32_t out = v;
    if (out < lo) out = lo;
    if (out > hi) out = hi;
    return out;
}

static inline float ora_bits_f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static inline uint32_t ora_f_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

#endif
