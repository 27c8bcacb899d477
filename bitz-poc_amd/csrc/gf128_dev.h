// GF(2^128) = F2[X]/(X^128 + X^7 + X^2 + X + 1) on CDNA4 integer ALUs.
//
// Element layout: four little-endian 32-bit words w[0..3] (w[0] = bits 0..31),
// i.e. the 16 wire bytes lo_le || hi_le of vendor/field/src/gf128.rs.
//
// CDNA4 has no carry-less multiply. Two multiplication paths:
//  * gf_mul: variable x variable. 32x32 carry-less products are computed with
//    ordinary integer multiplies on four residue classes of bits spaced four
//    apart (a class of a 32-bit word has at most 8 set bits, so each partial
//    count fits its 4-bit slot and its low bit is the GF(2) coefficient), then
//    two levels of Karatsuba (9 products) and the two-step reduction by
//    X^128 = X^7 + X^2 + X + 1.
//  * gf_mul_tab: fixed multiplier c. A 4-bit window table T[i][n] =
//    (n * X^(4i)) * c (32 windows x 16 entries x 16 B = 8 KiB) lives in LDS;
//    v*c = XOR over the 32 nibbles of v of T[i][nibble_i(v)].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct gf128 {
    uint32_t w[4];
};

__device__ __forceinline__ gf128 gf_zero() { return gf128{{0u, 0u, 0u, 0u}}; }
__device__ __forceinline__ gf128 gf_one() { return gf128{{1u, 0u, 0u, 0u}}; }

__device__ __forceinline__ gf128 gf_add(gf128 a, gf128 b) {
    return gf128{{a.w[0] ^ b.w[0], a.w[1] ^ b.w[1], a.w[2] ^ b.w[2], a.w[3] ^ b.w[3]}};
}

__device__ __forceinline__ bool gf_is_zero(gf128 a) {
    return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) == 0u;
}

// Field elements cross the kernel-argument boundary as two u64 scalars
// (vector types by value do not keep one layout on host and device).
__device__ __forceinline__ gf128 gf_from_u64(uint64_t lo, uint64_t hi) {
    return gf128{{(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32)}};
}

__device__ __forceinline__ gf128 gf_load(const uint4* p) {
    uint4 v = *p;
    return gf128{{v.x, v.y, v.z, v.w}};
}
__device__ __forceinline__ void gf_store(uint4* p, gf128 a) {
    *p = make_uint4(a.w[0], a.w[1], a.w[2], a.w[3]);
}
// Streaming store (nontemporal hint) for outputs far larger than the caches
// that the next pass reads from HBM anyway.
__device__ __forceinline__ void gf_store_nt(uint4* p, gf128 a) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    u32x4 v = {a.w[0], a.w[1], a.w[2], a.w[3]};
    __builtin_nontemporal_store(v, (u32x4*)p);
}

// 3-input XOR (v_bitop3_b32, truth table 0x96), per 32-bit half for u64
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ uint64_t xor3(uint64_t a, uint64_t b, uint64_t c) {
    return (uint64_t)xor3((uint32_t)a, (uint32_t)b, (uint32_t)c) |
           ((uint64_t)xor3((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32)) << 32);
}

// bitwise m ? a : b (v_bitop3_b32; src0/1/2 weigh 0xF0/0xCC/0xAA in the table)
__device__ __forceinline__ uint32_t bitsel(uint32_t m, uint32_t a, uint32_t b) {
    return __builtin_amdgcn_bitop3_b32(m, a, b, 0xCA);
}

// 32x32 -> 64 carry-less product via four residue classes.
__device__ __forceinline__ uint64_t clmul32(uint32_t a, uint32_t b) {
    const uint32_t m = 0x11111111u;
    uint32_t a0 = a & m, a1 = a & (m << 1), a2 = a & (m << 2), a3 = a & (m << 3);
    uint32_t b0 = b & m, b1 = b & (m << 1), b2 = b & (m << 2), b3 = b & (m << 3);
    // Class product (i, j) counts, at output slot 4k + ((i+j) mod 4), the
    // pairs u + v = k (i + j < 4: at most min(k+1, 15-k)) or u + v = k - 1
    // (i + j >= 4: at most min(k, 16-k)); one of each kind adds to at most 15
    // per 4-bit slot, so such pairs share one multiply-add (v_mad_u64_u32's
    // addend) instead of an XOR.
    uint64_t z0 = xor3((uint64_t)a0 * b0 + (uint64_t)a1 * b3, (uint64_t)a2 * b2, (uint64_t)a3 * b1);
    uint64_t z1 = ((uint64_t)a0 * b1 + (uint64_t)a2 * b3) ^ ((uint64_t)a1 * b0 + (uint64_t)a3 * b2);
    uint64_t z2 = xor3((uint64_t)a0 * b2 + (uint64_t)a3 * b3, (uint64_t)a1 * b1, (uint64_t)a2 * b0);
    uint64_t z3 = xor3((uint64_t)a0 * b3, (uint64_t)a1 * b2, (uint64_t)a2 * b1) ^ ((uint64_t)a3 * b0);
    // the classes' bit positions are disjoint and cover the word: select
    const uint32_t M = 0x11111111u;
    uint32_t lo = bitsel(M, (uint32_t)z0, bitsel(M << 1, (uint32_t)z1, bitsel(M << 2, (uint32_t)z2, (uint32_t)z3)));
    uint32_t hi = bitsel(M, (uint32_t)(z0 >> 32),
                         bitsel(M << 1, (uint32_t)(z1 >> 32), bitsel(M << 2, (uint32_t)(z2 >> 32), (uint32_t)(z3 >> 32))));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// 64x64 -> 128 (Karatsuba over 32-bit halves); out[0] low, out[1] high.
__device__ __forceinline__ void clmul64(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
    uint32_t al = (uint32_t)a, ah = (uint32_t)(a >> 32);
    uint32_t bl = (uint32_t)b, bh = (uint32_t)(b >> 32);
    uint64_t l = clmul32(al, bl);
    uint64_t h = clmul32(ah, bh);
    uint64_t m = xor3(clmul32(al ^ ah, bl ^ bh), l, h);
    *lo = l ^ (m << 32);
    *hi = h ^ (m >> 32);
}

// Unreduced 256-bit product r[0..3] (64-bit words, least significant first).
struct gf256 {
    uint64_t r[4];
};

__device__ __forceinline__ gf256 gf_mul_wide(gf128 a, gf128 b) {
    uint64_t a0 = (uint64_t)a.w[0] | ((uint64_t)a.w[1] << 32);
    uint64_t a1 = (uint64_t)a.w[2] | ((uint64_t)a.w[3] << 32);
    uint64_t b0 = (uint64_t)b.w[0] | ((uint64_t)b.w[1] << 32);
    uint64_t b1 = (uint64_t)b.w[2] | ((uint64_t)b.w[3] << 32);
    uint64_t l0, l1, h0, h1, m0, m1;
    clmul64(a0, b0, &l0, &l1);
    clmul64(a1, b1, &h0, &h1);
    clmul64(a0 ^ a1, b0 ^ b1, &m0, &m1);
    gf256 o;
    o.r[0] = l0;
    o.r[1] = xor3(l1, m0, l0) ^ h0;
    o.r[2] = xor3(h0, m1, l1) ^ h1;
    o.r[3] = h1;
    return o;
}

__device__ __forceinline__ gf256 gf256_zero() { return gf256{{0, 0, 0, 0}}; }
__device__ __forceinline__ void gf256_acc(gf256* acc, gf256 x) {
    acc->r[0] ^= x.r[0];
    acc->r[1] ^= x.r[1];
    acc->r[2] ^= x.r[2];
    acc->r[3] ^= x.r[3];
}

// Reduce modulo X^128 + X^7 + X^2 + X + 1: fold the two high words down as
// r2*g + (r3*g)*X^64 with g = X^7+X^2+X+1; the <= 7 bits that spill above
// X^128 are folded once more.
__device__ __forceinline__ gf128 gf_reduce(gf256 x) {
    uint64_t r2 = x.r[2], r3 = x.r[3];
    uint64_t a0 = xor3(r2, r2 << 1, r2 << 2) ^ (r2 << 7);
    uint64_t a1 = xor3(r2 >> 63, r2 >> 62, r2 >> 57);
    uint64_t b0 = xor3(r3, r3 << 1, r3 << 2) ^ (r3 << 7);
    uint64_t b1 = xor3(r3 >> 63, r3 >> 62, r3 >> 57);
    uint64_t w0 = xor3(x.r[0], a0, b1) ^ xor3(b1 << 1, b1 << 2, b1 << 7);
    uint64_t w1 = xor3(x.r[1], a1, b0);
    return gf128{{(uint32_t)w0, (uint32_t)(w0 >> 32), (uint32_t)w1, (uint32_t)(w1 >> 32)}};
}

__device__ __forceinline__ gf128 gf_mul(gf128 a, gf128 b) { return gf_reduce(gf_mul_wide(a, b)); }

// Squaring spreads the bits (cross terms vanish in characteristic two).
__device__ __forceinline__ uint64_t gf_spread32(uint32_t w) {
    uint64_t x = w;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}
__device__ __forceinline__ gf128 gf_square(gf128 a) {
    gf256 s{{gf_spread32(a.w[0]), gf_spread32(a.w[1]), gf_spread32(a.w[2]), gf_spread32(a.w[3])}};
    return gf_reduce(s);
}

// a * X (shift left by one with reduction).
__device__ __forceinline__ gf128 gf_mul_x(gf128 a) {
    uint32_t carry = a.w[3] >> 31;
    gf128 o;
    o.w[3] = (a.w[3] << 1) | (a.w[2] >> 31);
    o.w[2] = (a.w[2] << 1) | (a.w[1] >> 31);
    o.w[1] = (a.w[1] << 1) | (a.w[0] >> 31);
    o.w[0] = (a.w[0] << 1) ^ (carry * 0x87u);
    return o;
}

// a * X^4: shift by a nibble; the spilled nibble t (bits 124..127) returns
// as t * (X^7 + X^2 + X + 1), at most 11 bits, into the low word.
__device__ __forceinline__ gf128 gf_mul_x4(gf128 a) {
    const uint32_t t = a.w[3] >> 28;
    gf128 o;
    o.w[3] = (a.w[3] << 4) | (a.w[2] >> 28);
    o.w[2] = (a.w[2] << 4) | (a.w[1] >> 28);
    o.w[1] = (a.w[1] << 4) | (a.w[0] >> 28);
    o.w[0] = (a.w[0] << 4) ^ t ^ (t << 1) ^ (t << 2) ^ (t << 7);
    return o;
}

// ---- fixed-multiplier 4-bit window tables (8 KiB each) ----
// Layout: tab[i*16 + n] = (n * X^(4i)) * c, i in [0,32), n in [0,16).
// Built cooperatively by `nthreads` threads of a block; caller syncs.
__device__ __forceinline__ void gf_table_build(uint4* tab, gf128 c, int tid, int nthreads) {
    // Each of the 32 windows needs c * X^(4i) .. c * X^(4i+3); thread handles
    // whole windows (32 windows; threads beyond 32 idle here).
    for (int i = tid; i < 32; i += nthreads) {
        gf128 b0 = c;
        for (int k = 0; k < i; k++) b0 = gf_mul_x4(b0);
        gf128 b1 = gf_mul_x(b0), b2 = gf_mul_x(b1), b3 = gf_mul_x(b2);
        for (int n = 0; n < 16; n++) {
            gf128 acc = gf_zero();
            if (n & 1) acc = gf_add(acc, b0);
            if (n & 2) acc = gf_add(acc, b1);
            if (n & 4) acc = gf_add(acc, b2);
            if (n & 8) acc = gf_add(acc, b3);
            gf_store(&tab[i * 16 + n], acc);
        }
    }
}

// gf_mul_tab one input word at a time (the word loop kept rolled): at most
// eight lookups in flight, for kernels that run many table products on
// register-resident data, where hoisting every lookup would spill.
__device__ __forceinline__ gf128 gf_mul_tab_lean(const uint4* tab, gf128 v) {
    uint32_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;
#pragma unroll 1
    for (int word = 0; word < 4; word++) {
        // a uniform select, not an indexed register array (that would move
        // v to memory)
        const uint32_t w = word == 0 ? v.w[0] : word == 1 ? v.w[1] : word == 2 ? v.w[2] : v.w[3];
        const uint4* tw = tab + word * 128;
#pragma unroll
        for (int nib = 0; nib < 8; nib += 2) {
            uint4 t = tw[nib * 16 + ((w >> (4 * nib)) & 15u)];
            uint4 u = tw[(nib + 1) * 16 + ((w >> (4 * nib + 4)) & 15u)];
            x0 = __builtin_amdgcn_bitop3_b32(x0, t.x, u.x, 0x96);
            x1 = __builtin_amdgcn_bitop3_b32(x1, t.y, u.y, 0x96);
            x2 = __builtin_amdgcn_bitop3_b32(x2, t.z, u.z, 0x96);
            x3 = __builtin_amdgcn_bitop3_b32(x3, t.w, u.w, 0x96);
        }
    }
    return gf128{{x0, x1, x2, x3}};
}

// Two windows per step, so each accumulator word takes one 3-input XOR
// (v_bitop3_b32) per pair of table entries.
__device__ __forceinline__ gf128 gf_mul_tab(const uint4* tab, gf128 v) {
    uint32_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;
#pragma unroll
    for (int word = 0; word < 4; word++) {
        uint32_t w = v.w[word];
#pragma unroll
        for (int nib = 0; nib < 8; nib += 2) {
            uint4 t = tab[(word * 8 + nib) * 16 + ((w >> (4 * nib)) & 15u)];
            uint4 u = tab[(word * 8 + nib + 1) * 16 + ((w >> (4 * nib + 4)) & 15u)];
            x0 = __builtin_amdgcn_bitop3_b32(x0, t.x, u.x, 0x96);
            x1 = __builtin_amdgcn_bitop3_b32(x1, t.y, u.y, 0x96);
            x2 = __builtin_amdgcn_bitop3_b32(x2, t.z, u.z, 0x96);
            x3 = __builtin_amdgcn_bitop3_b32(x3, t.w, u.w, 0x96);
        }
    }
    return gf128{{x0, x1, x2, x3}};
}
