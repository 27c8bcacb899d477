// Integers modulo an odd q < 2^126 on the host: the standalone instance's
// prime field (its eq tables and claim) and the prime search's
// Miller-Rabin rounds. Plain C++ (no device code).
#pragma once
#include <stdint.h>
#include <string.h>

namespace bitz {

typedef unsigned __int128 u128;

inline u128 addmod(u128 a, u128 b, u128 q) {
    u128 s = a + b;  // q < 2^126: no overflow
    return s >= q ? s - q : s;
}
inline u128 mulmod_serial(u128 a, u128 b, u128 q) {
    u128 r = 0;
    a %= q;
    for (int i = 127; i >= 0; i--) {
        r = addmod(r, r, q);
        if ((b >> i) & 1) r = addmod(r, a, q);
    }
    return r;
}

// Montgomery arithmetic modulo an odd n < 2^126 with R = 2^128.
struct ModN {
    u128 n = 0, ninv = 0, r2 = 0;  // ninv = -n^-1 mod R, r2 = R^2 mod n
    void set(u128 n_) {
        n = n_;
        u128 x = n;  // inverse of n modulo 2^3; each Newton step doubles the bits
        for (int i = 0; i < 7; i++) x *= 2 - n * x;
        ninv = (u128)0 - x;
        u128 r = (~(u128)0 % n + 1) % n;  // R mod n
        r2 = mulmod_serial(r, r, n);
    }
    static void wide(u128 a, u128 b, u128* hi, u128* lo) {
        uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64), b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
        u128 p00 = (u128)a0 * b0, p01 = (u128)a0 * b1, p10 = (u128)a1 * b0, p11 = (u128)a1 * b1;
        u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
        *lo = (uint64_t)p00 | (mid << 64);
        *hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
    }
    // (hi R + lo) R^-1 mod n for hi R + lo < R n
    u128 redc(u128 hi, u128 lo) const {
        u128 m = lo * ninv, mh, ml;
        wide(m, n, &mh, &ml);
        u128 c = (lo + ml) < lo ? 1 : 0;
        u128 r = hi + mh + c;
        return r >= n ? r - n : r;
    }
    u128 mul(u128 a, u128 b) const {
        if (a >= n) a %= n;
        if (b >= n) b %= n;
        u128 hi, lo;
        wide(a, b, &hi, &lo);
        u128 t = redc(hi, lo);  // a b R^-1
        wide(t, r2, &hi, &lo);
        return redc(hi, lo);
    }
    // a b R^-1 mod n, for a b < R n (e.g. any a < R with b < n): one
    // product and one reduction, no division
    u128 mont(u128 a, u128 b) const {
        u128 hi, lo;
        wide(a, b, &hi, &lo);
        return redc(hi, lo);
    }
    u128 to_mont(u128 a) const { return mont(a % n, r2); }  // a R mod n
};

// a b mod q (q odd, q < 2^126); the Montgomery constants of the last
// modulus are kept per thread.
inline u128 mulmod(u128 a, u128 b, u128 q) {
    static thread_local ModN ctx;
    if (ctx.n != q) ctx.set(q);
    return ctx.mul(a, b);
}
inline u128 powmod(u128 a, u128 e, u128 q) {
    u128 r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = mulmod(r, a, q);
        a = mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}
inline u128 u128_from(const uint8_t* p) {
    uint64_t lo, hi;
    memcpy(&lo, p, 8);
    memcpy(&hi, p + 8, 8);
    return ((u128)hi << 64) | lo;
}
inline void u128_put(u128 v, uint8_t* p) {
    uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
    memcpy(p, &lo, 8);
    memcpy(p + 8, &hi, 8);
}

}  // namespace bitz
