/*
 * xo_seedhash.h — generator state of one packet of a batch instance.
 *
 * The batched and served engines key every packet's MWC generator on
 * (instance seed, packet index) instead of on the lane that happens to run
 * it, so a result is the same whatever else shares the launch. The
 * multiplier a is a row of the safe-prime table (rows 1..499999, as the
 * reference's lanes use) picked by one SplitMix64 hash of the key; the
 * state x = (c << 32) | lo is a uniformly drawn valid state (lo in
 * [1, 2^32-2], c in [0, a-2]) from a second hash. SplitMix64's finalizer is
 * the published mixer of G. Steele, D. Lea and C. Flood (OOPSLA 2014).
 * Plain integer arithmetic: host and device compute identical states.
 */
#ifndef XO_SEEDHASH_H
#define XO_SEEDHASH_H

#include <stdint.h>

#include "xo_math.h" /* XO_HD */

#define XO_GOLDEN64 0x9E3779B97F4A7C15ull
#define XO_SEED_TABLE_ROWS 499999u

XO_HD uint64_t xo_mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}

/* table[i] is the multiplier of seed-table row i + 1. */
XO_HD void xo_packet_state(uint64_t seed, uint64_t p, const uint32_t *table,
                           uint64_t *x, uint32_t *a) {
    uint64_t z = seed + (p + 1u) * XO_GOLDEN64;
    uint64_t h1 = xo_mix64(z);
    uint64_t h2 = xo_mix64(z + XO_GOLDEN64);
    /* row = floor(hi32(h1) * rows / 2^32): one multiply (Lemire's
     * multiply-shift range reduction) instead of a 64-bit remainder */
    uint32_t m = table[(uint32_t)(((h1 >> 32) * (uint64_t)XO_SEED_TABLE_ROWS) >> 32)];
    uint32_t lo = (uint32_t)h2;
    if (lo == 0u || lo == 0xFFFFFFFFu) lo = 1u;
    uint32_t c = (uint32_t)(((uint64_t)(uint32_t)(h2 >> 32) * (uint64_t)(m - 1u)) >> 32);
    *x = ((uint64_t)c << 32) | lo;
    *a = m;
}

#endif /* XO_SEEDHASH_H */
