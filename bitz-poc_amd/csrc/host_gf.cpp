// Host GF(2^128) loops (host_gf.h). The 128 x 128 -> 256-bit carry-less
// product is schoolbook over 64-bit halves (four PCLMULQDQ, or the integer
// residue-class product below); the 256-bit result reduces modulo
// x^128 + x^7 + x^2 + x + 1 by folding the top two words down with
// x^128 = x^7 + x^2 + x + 1. Sums accumulate unreduced and reduce once:
// the reduction is F2-linear.
#include "host_gf.h"

#include <immintrin.h>

namespace bitz {
namespace hostgf {
namespace {

struct W {
    uint64_t w[4];
};

// 64 x 64 -> 128 carry-less product with integer multiplies: bits of each
// operand split into four residue classes mod 4. With the top nibble of b
// set aside (added back by shifts), a product of two classes sums at most 15
// terms into each result bit of its class, so carries stay inside the 4-bit
// slot, and the class mask keeps the sum's low bit: the XOR.
inline void clmul64_int(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
    const uint64_t M = 0x1111111111111111ull;
    const unsigned __int128 Mw = ((unsigned __int128)M << 64) | M;
    const uint64_t bl = b & 0x0fffffffffffffffull;
    unsigned __int128 a0 = a & M, a1 = a & (M << 1), a2 = a & (M << 2), a3 = a & (M << 3);
    unsigned __int128 b0 = bl & M, b1 = bl & (M << 1), b2 = bl & (M << 2), b3 = bl & (M << 3);
    unsigned __int128 z0 = (a0 * b0) ^ (a1 * b3) ^ (a2 * b2) ^ (a3 * b1);
    unsigned __int128 z1 = (a0 * b1) ^ (a1 * b0) ^ (a2 * b3) ^ (a3 * b2);
    unsigned __int128 z2 = (a0 * b2) ^ (a1 * b1) ^ (a2 * b0) ^ (a3 * b3);
    unsigned __int128 z3 = (a0 * b3) ^ (a1 * b2) ^ (a2 * b1) ^ (a3 * b0);
    unsigned __int128 p = (z0 & Mw) | (z1 & (Mw << 1)) | (z2 & (Mw << 2)) | (z3 & (Mw << 3));
    for (int i = 60; i < 64; i++)
        if ((b >> i) & 1) p ^= ((unsigned __int128)a) << i;
    *lo = (uint64_t)p;
    *hi = (uint64_t)(p >> 64);
}

template <bool kClmul>
__attribute__((target("pclmul,sse4.1"), always_inline)) inline W wide(E a, E b) {
    W r;
    if constexpr (kClmul) {
        const __m128i A = _mm_set_epi64x((long long)a.hi, (long long)a.lo);
        const __m128i B = _mm_set_epi64x((long long)b.hi, (long long)b.lo);
        const __m128i lo = _mm_clmulepi64_si128(A, B, 0x00);
        const __m128i hi = _mm_clmulepi64_si128(A, B, 0x11);
        const __m128i mid = _mm_xor_si128(_mm_clmulepi64_si128(A, B, 0x01), _mm_clmulepi64_si128(A, B, 0x10));
        r.w[0] = (uint64_t)_mm_cvtsi128_si64(lo);
        r.w[1] = (uint64_t)_mm_extract_epi64(lo, 1) ^ (uint64_t)_mm_cvtsi128_si64(mid);
        r.w[2] = (uint64_t)_mm_cvtsi128_si64(hi) ^ (uint64_t)_mm_extract_epi64(mid, 1);
        r.w[3] = (uint64_t)_mm_extract_epi64(hi, 1);
    } else {
        uint64_t l0, l1, h0, h1, m0, m1, n0, n1;
        clmul64_int(a.lo, b.lo, &l0, &l1);
        clmul64_int(a.hi, b.hi, &h0, &h1);
        clmul64_int(a.lo, b.hi, &m0, &m1);
        clmul64_int(a.hi, b.lo, &n0, &n1);
        r.w[0] = l0;
        r.w[1] = l1 ^ m0 ^ n0;
        r.w[2] = h0 ^ m1 ^ n1;
        r.w[3] = h1;
    }
    return r;
}

inline E reduce(const W& r) {
    const uint64_t r2 = r.w[2], r3 = r.w[3];
    const uint64_t a0 = r2 ^ (r2 << 1) ^ (r2 << 2) ^ (r2 << 7);
    const uint64_t a1 = (r2 >> 63) ^ (r2 >> 62) ^ (r2 >> 57);
    const uint64_t b0 = r3 ^ (r3 << 1) ^ (r3 << 2) ^ (r3 << 7);
    const uint64_t b1 = (r3 >> 63) ^ (r3 >> 62) ^ (r3 >> 57);
    return E{r.w[0] ^ a0 ^ b1 ^ (b1 << 1) ^ (b1 << 2) ^ (b1 << 7), r.w[1] ^ a1 ^ b0};
}

inline void acc(W& a, const W& b) {
    for (int i = 0; i < 4; i++) a.w[i] ^= b.w[i];
}
inline E add(E a, E b) { return E{a.lo ^ b.lo, a.hi ^ b.hi}; }

template <bool kClmul>
__attribute__((target("pclmul,sse4.1"), always_inline)) inline E mul_t(E a, E b) {
    return reduce(wide<kClmul>(a, b));
}

template <bool kClmul>
__attribute__((target("pclmul,sse4.1"))) void eq_table_t(const E* r, int n, E* out) {
    out[0] = E{1, 0};
    for (int i = 0; i < n; i++) {
        const uint64_t len = 1ull << i;
        for (uint64_t x = 0; x < len; x++) {
            E hi = mul_t<kClmul>(out[x], r[i]);
            out[x + len] = hi;
            out[x] = add(out[x], hi);
        }
    }
}

template <bool kClmul>
__attribute__((target("pclmul,sse4.1"))) void round_sums_t(const E* l, const E* r, uint64_t h, const E* w,
                                                            int send_one, E out[2]) {
    W end{{0, 0, 0, 0}}, inf{{0, 0, 0, 0}};
    for (uint64_t x = 0; x < h; x++) {
        const E l0 = l[x], l1 = l[x + h], r0 = r[x], r1 = r[x + h];
        const E le = send_one ? l1 : l0, re = send_one ? r1 : r0;
        acc(end, wide<kClmul>(mul_t<kClmul>(w[x], le), re));
        acc(inf, wide<kClmul>(mul_t<kClmul>(w[x], add(l0, l1)), add(r0, r1)));
    }
    out[0] = reduce(end);
    out[1] = reduce(inf);
}

template <bool kClmul>
__attribute__((target("pclmul,sse4.1"))) void fold_t(E* v, uint64_t h, E rho) {
    for (uint64_t j = 0; j < h; j++) v[j] = add(v[j], mul_t<kClmul>(rho, add(v[j], v[j + h])));
}

template <bool kClmul>
__attribute__((target("pclmul,sse4.1"))) E mul_call(E a, E b) {
    return mul_t<kClmul>(a, b);
}

template <bool kClmul>
__attribute__((target("pclmul,sse4.1"))) void inner_round_t(const E* m, const E* w, uint64_t pairs, E out[3]) {
    W a0{{0, 0, 0, 0}}, a1{{0, 0, 0, 0}}, q{{0, 0, 0, 0}};
    for (uint64_t k = 0; k < pairs; k++) {
        const E m0 = m[2 * k], m1 = m[2 * k + 1], w0 = w[2 * k], w1 = w[2 * k + 1];
        acc(a0, wide<kClmul>(m0, w0));
        acc(a1, wide<kClmul>(m1, w1));
        acc(q, wide<kClmul>(add(m0, m1), add(w0, w1)));
    }
    out[0] = reduce(a0);
    out[1] = reduce(a1);
    out[2] = reduce(q);
}

template <bool kClmul>
__attribute__((target("pclmul,sse4.1"))) void fold_lsb_t(E* v, uint64_t half, E r) {
    for (uint64_t i = 0; i < half; i++) {
        const E v0 = v[2 * i], v1 = v[2 * i + 1];
        v[i] = add(v0, mul_t<kClmul>(r, add(v0, v1)));
    }
}

bool detect() {
    __builtin_cpu_init();
    return __builtin_cpu_supports("pclmul") && __builtin_cpu_supports("sse4.1");
}

}  // namespace

bool using_clmul() {
    static const bool on = detect();
    return on;
}

E mul(E a, E b) { return using_clmul() ? mul_call<true>(a, b) : mul_call<false>(a, b); }

void eq_table(const E* r, int n, E* out) {
    if (using_clmul())
        eq_table_t<true>(r, n, out);
    else
        eq_table_t<false>(r, n, out);
}

void gkr_round_sums(const E* l, const E* r, uint64_t h, const E* w, int send_one, E out[2]) {
    if (using_clmul())
        round_sums_t<true>(l, r, h, w, send_one, out);
    else
        round_sums_t<false>(l, r, h, w, send_one, out);
}

void drop_top(E* v, uint64_t h) {
    for (uint64_t j = 0; j < h; j++) v[j] = add(v[j], v[j + h]);
}

void inner_round(const E* m, const E* w, uint64_t pairs, E out[3]) {
    if (using_clmul())
        inner_round_t<true>(m, w, pairs, out);
    else
        inner_round_t<false>(m, w, pairs, out);
}

void fold_lsb(E* v, uint64_t half, E r) {
    if (using_clmul())
        fold_lsb_t<true>(v, half, r);
    else
        fold_lsb_t<false>(v, half, r);
}

void fold(E* v, uint64_t h, E rho) {
    if (using_clmul())
        fold_t<true>(v, h, rho);
    else
        fold_t<false>(v, h, rho);
}

}  // namespace hostgf
}  // namespace bitz
