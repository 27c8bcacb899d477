/*
 * GF(2^128) arithmetic, restating vendor/field/src/gf128/portable.rs:11-108.
 * Test infrastructure only (see oracle.h).
 *
 * Field: F2[X]/(X^128 + X^7 + X^2 + X + 1). Element = (lo, hi) 64-bit words,
 * bit i of lo = coefficient of X^i, bit i of hi = coefficient of X^(64+i)
 * (vendor/field/src/gf128.rs, src/bitz/codec.rs:4-21 for the 16-byte wire form).
 */
#include "oracle.h"

#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

static int g_force_portable = 0;
void or_gf_force_portable(int on) { g_force_portable = on; }

/* portable.rs:11-40: four residue classes counted by integer multiplication,
 * parity kept by masking; b's top nibble restored by a fixed bit schedule. */
static void clmul64_portable(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
    const uint64_t M = 0x1111111111111111ull;
    const unsigned __int128 W = ((unsigned __int128)M << 64) | M;
    uint64_t low_b = b & 0x0fffffffffffffffull;
    unsigned __int128 a0 = a & M, a1 = a & (M << 1), a2 = a & (M << 2), a3 = a & (M << 3);
    unsigned __int128 b0 = low_b & M, b1 = low_b & (M << 1), b2 = low_b & (M << 2),
                      b3 = low_b & (M << 3);
    unsigned __int128 z0 = (a0 * b0) ^ (a1 * b3) ^ (a2 * b2) ^ (a3 * b1);
    unsigned __int128 z1 = (a0 * b1) ^ (a1 * b0) ^ (a2 * b3) ^ (a3 * b2);
    unsigned __int128 z2 = (a0 * b2) ^ (a1 * b1) ^ (a2 * b0) ^ (a3 * b3);
    unsigned __int128 z3 = (a0 * b3) ^ (a1 * b2) ^ (a2 * b1) ^ (a3 * b0);
    unsigned __int128 p = (z0 & W) | (z1 & (W << 1)) | (z2 & (W << 2)) | (z3 & (W << 3));
    for (int i = 60; i < 64; i++) {
        if ((b >> i) & 1) p ^= ((unsigned __int128)a) << i;
    }
    *lo = (uint64_t)p;
    *hi = (uint64_t)(p >> 64);
}

#if defined(__x86_64__)
__attribute__((target("pclmul,sse4.1"))) static void clmul64_pclmul(uint64_t a, uint64_t b,
                                                                     uint64_t* lo, uint64_t* hi) {
    __m128i va = _mm_set_epi64x(0, (long long)a);
    __m128i vb = _mm_set_epi64x(0, (long long)b);
    __m128i r = _mm_clmulepi64_si128(va, vb, 0x00);
    *lo = (uint64_t)_mm_cvtsi128_si64(r);
    *hi = (uint64_t)_mm_extract_epi64(r, 1);
}
static int have_pclmul(void) {
    static int cached = -1;
    if (cached < 0) cached = __builtin_cpu_supports("pclmul") ? 1 : 0;
    return cached;
}
#else
static int have_pclmul(void) { return 0; }
#endif

int or_gf_uses_pclmul(void) { return have_pclmul() && !g_force_portable; }

static inline void clmul64(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
#if defined(__x86_64__)
    if (have_pclmul() && !g_force_portable) {
        clmul64_pclmul(a, b, lo, hi);
        return;
    }
#endif
    clmul64_portable(a, b, lo, hi);
}

/* portable.rs:58-64: Karatsuba 128x128 -> 256 */
static void clmul128(gf128_t a, gf128_t b, uint64_t r[4]) {
    uint64_t l0, l1, h0, h1, m0, m1;
    clmul64(a.lo, b.lo, &l0, &l1);
    clmul64(a.hi, b.hi, &h0, &h1);
    clmul64(a.lo ^ a.hi, b.lo ^ b.hi, &m0, &m1);
    r[0] = l0;
    r[1] = l1 ^ m0 ^ l0 ^ h0;
    r[2] = h0 ^ m1 ^ l1 ^ h1;
    r[3] = h1;
}

/* portable.rs:67-87: substitute X^128 = X^7 + X^2 + X + 1 twice */
static gf128_t reduce256(const uint64_t r[4]) {
    uint64_t a0 = r[2] ^ (r[2] << 1) ^ (r[2] << 2) ^ (r[2] << 7);
    uint64_t a1 = (r[2] >> 63) ^ (r[2] >> 62) ^ (r[2] >> 57);
    uint64_t b0 = r[3] ^ (r[3] << 1) ^ (r[3] << 2) ^ (r[3] << 7);
    uint64_t b1 = (r[3] >> 63) ^ (r[3] >> 62) ^ (r[3] >> 57);
    uint64_t w0 = r[0] ^ a0;
    uint64_t w1 = r[1] ^ a1 ^ b0;
    gf128_t out;
    out.lo = w0 ^ b1 ^ (b1 << 1) ^ (b1 << 2) ^ (b1 << 7);
    out.hi = w1;
    return out;
}

gf128_t or_gf_mul(gf128_t a, gf128_t b) {
    uint64_t r[4];
    clmul128(a, b, r);
    return reduce256(r);
}

/* portable.rs:91-95, 114-130: spread bits, then reduce */
static uint64_t spread32(uint32_t w) {
    uint64_t x = w;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

gf128_t or_gf_square(gf128_t a) {
    uint64_t r[4] = {spread32((uint32_t)a.lo), spread32((uint32_t)(a.lo >> 32)),
                     spread32((uint32_t)a.hi), spread32((uint32_t)(a.hi >> 32))};
    return reduce256(r);
}

/* inverse_or_zero (vendor/field/src/gf128.rs:127): a^(2^128-2), 0 -> 0.
 * 2^128 - 2 = sum_{i=1}^{127} 2^i, so a^-1 = prod_{i=1}^{127} a^(2^i). */
gf128_t or_gf_inv(gf128_t a) {
    gf128_t acc = {1, 0};
    gf128_t sq = a;
    for (int i = 1; i < 128; i++) {
        sq = or_gf_square(sq);
        acc = or_gf_mul(acc, sq);
    }
    if (a.lo == 0 && a.hi == 0) acc.lo = 0;
    return acc;
}

gf128_t or_gf_pow(gf128_t a, uint64_t e_lo, uint64_t e_hi) {
    gf128_t acc = {1, 0};
    for (int i = 127; i >= 0; i--) {
        acc = or_gf_square(acc);
        uint64_t bit = i >= 64 ? (e_hi >> (i - 64)) & 1 : (e_lo >> i) & 1;
        if (bit) acc = or_gf_mul(acc, a);
    }
    return acc;
}

void or_gf_mul_arrays(const gf128_t* a, const gf128_t* b, gf128_t* out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = or_gf_mul(a[i], b[i]);
}
