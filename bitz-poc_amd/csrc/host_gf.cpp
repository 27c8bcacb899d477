// Host GF(2^128) loops (host_gf.h), with two implementations of the field
// operations behind one set of loops:
//  - ClmulOps: elements stay in SSE registers; a product is four PCLMULQDQ
//    (schoolbook over 64-bit halves), and a 256-bit value hi x^128 + lo
//    reduces as lo + hi * (x^7 + x^2 + x + 1) with three more (hi's halves
//    times 0x87, then the 7 bits that spill past x^128 once more);
//  - IntOps: the same results from integer multiplies (residue-class
//    carry-less products) and shift-based reduction, for CPUs without
//    PCLMULQDQ.
// Sums accumulate unreduced and reduce once: the reduction is F2-linear.
#include "host_gf.h"

#include <immintrin.h>

namespace bitz {
namespace hostgf {
namespace {

#define HG_TARGET __attribute__((target("pclmul,sse4.1")))
#define HG_INLINE __attribute__((target("pclmul,sse4.1"), always_inline)) inline

struct ClmulOps {
    typedef __m128i V;
    struct W {
        __m128i lo, hi;
    };
    static HG_INLINE V load(const E* p) { return _mm_loadu_si128((const __m128i*)p); }
    static HG_INLINE void store(E* p, V v) { _mm_storeu_si128((__m128i*)p, v); }
    static HG_INLINE V from(E e) { return _mm_set_epi64x((long long)e.hi, (long long)e.lo); }
    static HG_INLINE V one() { return _mm_set_epi64x(0, 1); }
    static HG_INLINE V add(V a, V b) { return _mm_xor_si128(a, b); }
    static HG_INLINE W zero_w() { return W{_mm_setzero_si128(), _mm_setzero_si128()}; }
    static HG_INLINE W wide(V a, V b) {
        const __m128i p0 = _mm_clmulepi64_si128(a, b, 0x00);
        const __m128i p3 = _mm_clmulepi64_si128(a, b, 0x11);
        const __m128i m = _mm_xor_si128(_mm_clmulepi64_si128(a, b, 0x01), _mm_clmulepi64_si128(a, b, 0x10));
        return W{_mm_xor_si128(p0, _mm_slli_si128(m, 8)), _mm_xor_si128(p3, _mm_srli_si128(m, 8))};
    }
    static HG_INLINE void acc(W& a, const W& b) {
        a.lo = _mm_xor_si128(a.lo, b.lo);
        a.hi = _mm_xor_si128(a.hi, b.hi);
    }
    static HG_INLINE V reduce(const W& r) {
        const __m128i k = _mm_set_epi64x(0, 0x87);
        const __m128i t0 = _mm_clmulepi64_si128(r.hi, k, 0x00);  // hi.lo * 0x87, < 2^71
        const __m128i t1 = _mm_clmulepi64_si128(r.hi, k, 0x01);  // hi.hi * 0x87, weight 2^64
        const __m128i v = _mm_xor_si128(r.lo, _mm_xor_si128(t0, _mm_slli_si128(t1, 8)));
        const __m128i over = _mm_srli_si128(t1, 8);  // 7 bits at weight 2^128
        return _mm_xor_si128(v, _mm_clmulepi64_si128(over, k, 0x00));
    }
    static HG_INLINE V mul(V a, V b) { return reduce(wide(a, b)); }
    static HG_INLINE E get(V v) {
        E e;
        _mm_storeu_si128((__m128i*)&e, v);
        return e;
    }
};

struct IntOps {
    typedef E V;
    struct W {
        uint64_t w[4];
    };
    // 64 x 64 -> 128 carry-less product with integer multiplies: bits of each
    // operand split into four residue classes mod 4. With the top nibble of
    // b set aside (added back by shifts), a product of two classes sums at
    // most 15 terms into each result bit of its class, so carries stay inside
    // the 4-bit slot, and the class mask keeps the sum's low bit: the XOR.
    static inline void clmul64(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
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
    static inline V load(const E* p) { return *p; }
    static inline void store(E* p, V v) { *p = v; }
    static inline V from(E e) { return e; }
    static inline V one() { return E{1, 0}; }
    static inline V add(V a, V b) { return E{a.lo ^ b.lo, a.hi ^ b.hi}; }
    static inline W zero_w() { return W{{0, 0, 0, 0}}; }
    static inline W wide(V a, V b) {
        uint64_t l0, l1, h0, h1, m0, m1, n0, n1;
        clmul64(a.lo, b.lo, &l0, &l1);
        clmul64(a.hi, b.hi, &h0, &h1);
        clmul64(a.lo, b.hi, &m0, &m1);
        clmul64(a.hi, b.lo, &n0, &n1);
        return W{{l0, l1 ^ m0 ^ n0, h0 ^ m1 ^ n1, h1}};
    }
    static inline void acc(W& a, const W& b) {
        for (int i = 0; i < 4; i++) a.w[i] ^= b.w[i];
    }
    static inline V reduce(const W& r) {
        const uint64_t r2 = r.w[2], r3 = r.w[3];
        const uint64_t a0 = r2 ^ (r2 << 1) ^ (r2 << 2) ^ (r2 << 7);
        const uint64_t a1 = (r2 >> 63) ^ (r2 >> 62) ^ (r2 >> 57);
        const uint64_t b0 = r3 ^ (r3 << 1) ^ (r3 << 2) ^ (r3 << 7);
        const uint64_t b1 = (r3 >> 63) ^ (r3 >> 62) ^ (r3 >> 57);
        return E{r.w[0] ^ a0 ^ b1 ^ (b1 << 1) ^ (b1 << 2) ^ (b1 << 7), r.w[1] ^ a1 ^ b0};
    }
    static inline V mul(V a, V b) { return reduce(wide(a, b)); }
    static inline E get(V v) { return v; }
};

// Four elements per 512-bit register (VPCLMULQDQ works on each 128-bit
// lane): the same arithmetic as ClmulOps, four slots per instruction.
#define HG_TARGET512 __attribute__((target("pclmul,sse4.1,avx512f,avx512bw,vpclmulqdq")))
#define HG_INLINE512 __attribute__((target("pclmul,sse4.1,avx512f,avx512bw,vpclmulqdq"), always_inline)) inline
struct ZmmOps {
    typedef __m512i V;
    struct W {
        __m512i lo, hi;
    };
    static HG_INLINE512 V load(const E* p) { return _mm512_loadu_si512((const void*)p); }
    static HG_INLINE512 void store(E* p, V v) { _mm512_storeu_si512((void*)p, v); }
    static HG_INLINE512 V bcast(E e) {
        return _mm512_broadcast_i32x4(_mm_set_epi64x((long long)e.hi, (long long)e.lo));
    }
    static HG_INLINE512 V add(V a, V b) { return _mm512_xor_si512(a, b); }
    static HG_INLINE512 W zero_w() { return W{_mm512_setzero_si512(), _mm512_setzero_si512()}; }
    static HG_INLINE512 W wide(V a, V b) {
        const __m512i p0 = _mm512_clmulepi64_epi128(a, b, 0x00);
        const __m512i p3 = _mm512_clmulepi64_epi128(a, b, 0x11);
        const __m512i m = _mm512_xor_si512(_mm512_clmulepi64_epi128(a, b, 0x01), _mm512_clmulepi64_epi128(a, b, 0x10));
        return W{_mm512_xor_si512(p0, _mm512_bslli_epi128(m, 8)), _mm512_xor_si512(p3, _mm512_bsrli_epi128(m, 8))};
    }
    static HG_INLINE512 void acc(W& a, const W& b) {
        a.lo = _mm512_xor_si512(a.lo, b.lo);
        a.hi = _mm512_xor_si512(a.hi, b.hi);
    }
    static HG_INLINE512 V reduce(const W& r) {
        const __m512i k = _mm512_broadcast_i32x4(_mm_set_epi64x(0, 0x87));
        const __m512i t0 = _mm512_clmulepi64_epi128(r.hi, k, 0x00);
        const __m512i t1 = _mm512_clmulepi64_epi128(r.hi, k, 0x01);
        const __m512i v = _mm512_xor_si512(r.lo, _mm512_xor_si512(t0, _mm512_bslli_epi128(t1, 8)));
        return _mm512_xor_si512(v, _mm512_clmulepi64_epi128(_mm512_bsrli_epi128(t1, 8), k, 0x00));
    }
    static HG_INLINE512 V mul(V a, V b) { return reduce(wide(a, b)); }
    // the XOR of the four lanes, reduced
    static HG_INLINE512 E fold4(const W& w) {
        const __m128i lo = _mm_xor_si128(
            _mm_xor_si128(_mm512_extracti32x4_epi32(w.lo, 0), _mm512_extracti32x4_epi32(w.lo, 1)),
            _mm_xor_si128(_mm512_extracti32x4_epi32(w.lo, 2), _mm512_extracti32x4_epi32(w.lo, 3)));
        const __m128i hi = _mm_xor_si128(
            _mm_xor_si128(_mm512_extracti32x4_epi32(w.hi, 0), _mm512_extracti32x4_epi32(w.hi, 1)),
            _mm_xor_si128(_mm512_extracti32x4_epi32(w.hi, 2), _mm512_extracti32x4_epi32(w.hi, 3)));
        return ClmulOps::get(ClmulOps::reduce(ClmulOps::W{lo, hi}));
    }
};

// Loops of four slots at a time (h a multiple of 4).
HG_TARGET512 void round_sums_zmm(const E* l, const E* r, uint64_t h, const E* w, int send_one, E out[2]) {
    typedef ZmmOps O;
    O::W end = O::zero_w(), inf = O::zero_w();
    for (uint64_t x = 0; x < h; x += 4) {
        const O::V l0 = O::load(&l[x]), l1 = O::load(&l[x + h]), r0 = O::load(&r[x]), r1 = O::load(&r[x + h]),
                   wx = O::load(&w[x]);
        O::acc(end, O::wide(O::mul(wx, send_one ? l1 : l0), send_one ? r1 : r0));
        O::acc(inf, O::wide(O::mul(wx, O::add(l0, l1)), O::add(r0, r1)));
    }
    out[0] = O::fold4(end);
    out[1] = O::fold4(inf);
}

HG_TARGET512 void fold_zmm(E* v, uint64_t h, E rho) {
    typedef ZmmOps O;
    const O::V rr = O::bcast(rho);
    for (uint64_t j = 0; j < h; j += 4) {
        const O::V a = O::load(&v[j]);
        O::store(&v[j], O::add(a, O::mul(rr, O::add(a, O::load(&v[j + h])))));
    }
}

HG_TARGET512 void eq_stage_zmm(E* out, uint64_t len, E ri) {
    typedef ZmmOps O;
    const O::V rr = O::bcast(ri);
    for (uint64_t x = 0; x < len; x += 4) {
        const O::V p = O::load(&out[x]);
        const O::V hi = O::mul(p, rr);
        O::store(&out[x + len], hi);
        O::store(&out[x], O::add(p, hi));
    }
}

// pairs (m[2k], m[2k+1]): two registers hold four pairs; lanes 0, 2 of the
// even/odd split are gathered with a permute
HG_TARGET512 void inner_round_zmm(const E* m, const E* w, uint64_t pairs, E out[3]) {
    typedef ZmmOps O;
    O::W a0 = O::zero_w(), a1 = O::zero_w(), q = O::zero_w();
    const __m512i ev = _mm512_set_epi64(13, 12, 9, 8, 5, 4, 1, 0);   // elements 0, 2, 4, 6 of 8
    const __m512i od = _mm512_set_epi64(15, 14, 11, 10, 7, 6, 3, 2);  // elements 1, 3, 5, 7
    for (uint64_t k = 0; k < pairs; k += 4) {
        const __m512i ma = O::load(&m[2 * k]), mb = O::load(&m[2 * k + 4]);
        const __m512i wa = O::load(&w[2 * k]), wb = O::load(&w[2 * k + 4]);
        const __m512i m0 = _mm512_permutex2var_epi64(ma, ev, mb), m1 = _mm512_permutex2var_epi64(ma, od, mb);
        const __m512i w0 = _mm512_permutex2var_epi64(wa, ev, wb), w1 = _mm512_permutex2var_epi64(wa, od, wb);
        O::acc(a0, O::wide(m0, w0));
        O::acc(a1, O::wide(m1, w1));
        O::acc(q, O::wide(O::add(m0, m1), O::add(w0, w1)));
    }
    out[0] = O::fold4(a0);
    out[1] = O::fold4(a1);
    out[2] = O::fold4(q);
}

HG_TARGET512 void fold_lsb_zmm(E* v, uint64_t half, E r) {
    typedef ZmmOps O;
    const O::V rr = O::bcast(r);
    const __m512i ev = _mm512_set_epi64(13, 12, 9, 8, 5, 4, 1, 0);
    const __m512i od = _mm512_set_epi64(15, 14, 11, 10, 7, 6, 3, 2);
    // out index i reads 2i, 2i + 1: four outputs per eight inputs, written
    // after both registers are read (in place is safe, i < 2i for i > 0)
    for (uint64_t i = 0; i < half; i += 4) {
        const __m512i a = O::load(&v[2 * i]), b = O::load(&v[2 * i + 4]);
        const __m512i v0 = _mm512_permutex2var_epi64(a, ev, b), v1 = _mm512_permutex2var_epi64(a, od, b);
        O::store(&v[i], O::add(v0, O::mul(rr, O::add(v0, v1))));
    }
}

template <bool kClmul>
struct OpsOf {
    typedef IntOps T;
};
template <>
struct OpsOf<true> {
    typedef ClmulOps T;
};

template <bool kClmul>
HG_TARGET E mul_call(E a, E b) {
    typedef typename OpsOf<kClmul>::T O;
    return O::get(O::mul(O::from(a), O::from(b)));
}

template <bool kClmul>
HG_TARGET void eq_table_t(const E* r, int n, E* out) {
    typedef typename OpsOf<kClmul>::T O;
    O::store(&out[0], O::one());
    for (int i = 0; i < n; i++) {
        const uint64_t len = 1ull << i;
        const typename O::V ri = O::load(&r[i]);
        for (uint64_t x = 0; x < len; x++) {
            const typename O::V p = O::load(&out[x]);
            const typename O::V hi = O::mul(p, ri);
            O::store(&out[x + len], hi);
            O::store(&out[x], O::add(p, hi));
        }
    }
}

template <bool kClmul>
HG_TARGET void round_sums_t(const E* l, const E* r, uint64_t h, const E* w, int send_one, E out[2]) {
    typedef typename OpsOf<kClmul>::T O;
    typename O::W end = O::zero_w(), inf = O::zero_w();
    for (uint64_t x = 0; x < h; x++) {
        const typename O::V l0 = O::load(&l[x]), l1 = O::load(&l[x + h]), r0 = O::load(&r[x]),
                            r1 = O::load(&r[x + h]), wx = O::load(&w[x]);
        O::acc(end, O::wide(O::mul(wx, send_one ? l1 : l0), send_one ? r1 : r0));
        O::acc(inf, O::wide(O::mul(wx, O::add(l0, l1)), O::add(r0, r1)));
    }
    out[0] = O::get(O::reduce(end));
    out[1] = O::get(O::reduce(inf));
}

template <bool kClmul>
HG_TARGET void fold_t(E* v, uint64_t h, E rho) {
    typedef typename OpsOf<kClmul>::T O;
    const typename O::V rr = O::from(rho);
    for (uint64_t j = 0; j < h; j++) {
        const typename O::V a = O::load(&v[j]);
        O::store(&v[j], O::add(a, O::mul(rr, O::add(a, O::load(&v[j + h])))));
    }
}

template <bool kClmul>
HG_TARGET void inner_round_t(const E* m, const E* w, uint64_t pairs, E out[3]) {
    typedef typename OpsOf<kClmul>::T O;
    typename O::W a0 = O::zero_w(), a1 = O::zero_w(), q = O::zero_w();
    for (uint64_t k = 0; k < pairs; k++) {
        const typename O::V m0 = O::load(&m[2 * k]), m1 = O::load(&m[2 * k + 1]), w0 = O::load(&w[2 * k]),
                            w1 = O::load(&w[2 * k + 1]);
        O::acc(a0, O::wide(m0, w0));
        O::acc(a1, O::wide(m1, w1));
        O::acc(q, O::wide(O::add(m0, m1), O::add(w0, w1)));
    }
    out[0] = O::get(O::reduce(a0));
    out[1] = O::get(O::reduce(a1));
    out[2] = O::get(O::reduce(q));
}

template <bool kClmul>
HG_TARGET void fold_lsb_t(E* v, uint64_t half, E r) {
    typedef typename OpsOf<kClmul>::T O;
    const typename O::V rr = O::from(r);
    for (uint64_t i = 0; i < half; i++) {
        const typename O::V v0 = O::load(&v[2 * i]), v1 = O::load(&v[2 * i + 1]);
        O::store(&v[i], O::add(v0, O::mul(rr, O::add(v0, v1))));
    }
}

bool detect() {
    __builtin_cpu_init();
    return __builtin_cpu_supports("pclmul") && __builtin_cpu_supports("sse4.1");
}

bool detect512() {
    __builtin_cpu_init();
    return detect() && __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
           __builtin_cpu_supports("vpclmulqdq");
}

}  // namespace

bool using_clmul() {
    static const bool on = detect();
    return on;
}

bool using_vpclmul() {
    static const bool on = detect512();
    return on;
}

E mul(E a, E b) { return using_clmul() ? mul_call<true>(a, b) : mul_call<false>(a, b); }

void eq_table(const E* r, int n, E* out) {
    if (using_vpclmul() && n > 2) {
        eq_table_t<true>(r, 2, out);
        for (int i = 2; i < n; i++) eq_stage_zmm(out, 1ull << i, r[i]);
    } else if (using_clmul()) {
        eq_table_t<true>(r, n, out);
    } else {
        eq_table_t<false>(r, n, out);
    }
}

void gkr_round_sums(const E* l, const E* r, uint64_t h, const E* w, int send_one, E out[2]) {
    if (using_vpclmul() && h % 4 == 0)
        round_sums_zmm(l, r, h, w, send_one, out);
    else if (using_clmul())
        round_sums_t<true>(l, r, h, w, send_one, out);
    else
        round_sums_t<false>(l, r, h, w, send_one, out);
}

void drop_top(E* v, uint64_t h) {
    for (uint64_t j = 0; j < h; j++) v[j] = E{v[j].lo ^ v[j + h].lo, v[j].hi ^ v[j + h].hi};
}

void inner_round(const E* m, const E* w, uint64_t pairs, E out[3]) {
    if (using_vpclmul() && pairs % 4 == 0)
        inner_round_zmm(m, w, pairs, out);
    else if (using_clmul())
        inner_round_t<true>(m, w, pairs, out);
    else
        inner_round_t<false>(m, w, pairs, out);
}

void fold_lsb(E* v, uint64_t half, E r) {
    if (using_vpclmul() && half % 4 == 0)
        fold_lsb_zmm(v, half, r);
    else if (using_clmul())
        fold_lsb_t<true>(v, half, r);
    else
        fold_lsb_t<false>(v, half, r);
}

void fold(E* v, uint64_t h, E rho) {
    if (using_vpclmul() && h % 4 == 0)
        fold_zmm(v, h, rho);
    else if (using_clmul())
        fold_t<true>(v, h, rho);
    else
        fold_t<false>(v, h, rho);
}

}  // namespace hostgf
}  // namespace bitz
