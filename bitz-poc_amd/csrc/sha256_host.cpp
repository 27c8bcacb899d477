// SHA-256 compression on the host (declarations in sha256_host.h).
#include "sha256_host.h"

#include <cpuid.h>
#include <immintrin.h>
#include <string.h>

namespace bitz {
namespace chain {

const uint32_t kSha256K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

namespace {
inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

bool detect_shani() {
    unsigned a, b, c, d;
    // leaf 7 sub-leaf 0: EBX bit 29 = SHA extensions; leaf 1: ECX bit 19 = SSE4.1
    if (!__get_cpuid_count(7, 0, &a, &b, &c, &d) || !((b >> 29) & 1)) return false;
    if (!__get_cpuid(1, &a, &b, &c, &d) || !((c >> 19) & 1)) return false;
    return true;
}
}  // namespace

bool using_shani() {
    static const bool v = detect_shani();
    return v;
}

void compress_portable(const uint32_t state[8], const uint32_t block[16], uint32_t out[8]) {
    uint32_t w[64];
    for (int i = 0; i < 16; i++) w[i] = block[i];
    for (int i = 16; i < 64; i++) {
        const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = state[0], b = state[1], c = state[2], d = state[3], e = state[4], f = state[5], g = state[6],
             h = state[7];
    for (int i = 0; i < 64; i++) {
        const uint32_t t1 = h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + kSha256K[i] + w[i];
        const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    const uint32_t wk[8] = {a, b, c, d, e, f, g, h};
    for (int i = 0; i < 8; i++) out[i] = state[i] + wk[i];
}

// The SHA extensions keep the working state as (A, B, E, F) and (C, D, G, H)
// in the 32-bit lanes 3..0 of two registers; sha256rnds2 runs two rounds
// from the low two lanes of (message word + constant), sha256msg1/msg2 form
// the schedule's sigma0 and sigma1 terms four words at a time.
__attribute__((target("sha,sse4.1"))) void compress_shani(const uint32_t state[8], const uint32_t block[16],
                                                          uint32_t out[8]) {
    __m128i t = _mm_loadu_si128((const __m128i*)&state[0]);   // a b c d (lanes 0..3)
    __m128i s1 = _mm_loadu_si128((const __m128i*)&state[4]);  // e f g h
    t = _mm_shuffle_epi32(t, 0xB1);                            // b a d c
    s1 = _mm_shuffle_epi32(s1, 0x1B);                          // h g f e
    __m128i s0 = _mm_alignr_epi8(t, s1, 8);                    // f e b a: (A B E F) in lanes 3..0
    s1 = _mm_blend_epi16(s1, t, 0xF0);                         // h g d c: (C D G H) in lanes 3..0
    const __m128i abef = s0, cdgh = s1;
    __m128i w[4];
    for (int g = 0; g < 4; g++) w[g] = _mm_loadu_si128((const __m128i*)&block[4 * g]);
    for (int g = 0; g < 16; g++) {
        if (g >= 4) {
            // words 4g..4g+3: W[t-16] + sigma0(W[t-15]) + W[t-7] + sigma1(W[t-2])
            __m128i x = _mm_sha256msg1_epu32(w[g & 3], w[(g + 1) & 3]);
            x = _mm_add_epi32(x, _mm_alignr_epi8(w[(g + 3) & 3], w[(g + 2) & 3], 4));
            w[g & 3] = _mm_sha256msg2_epu32(x, w[(g + 3) & 3]);
        }
        const __m128i k = _mm_add_epi32(w[g & 3], _mm_loadu_si128((const __m128i*)&kSha256K[4 * g]));
        s1 = _mm_sha256rnds2_epu32(s1, s0, k);
        s0 = _mm_sha256rnds2_epu32(s0, s1, _mm_shuffle_epi32(k, 0x0E));
    }
    s0 = _mm_add_epi32(s0, abef);
    s1 = _mm_add_epi32(s1, cdgh);
    t = _mm_shuffle_epi32(s0, 0x1B);          // a b e f
    s1 = _mm_shuffle_epi32(s1, 0xB1);         // g h c d
    s0 = _mm_blend_epi16(t, s1, 0xF0);        // a b c d
    s1 = _mm_alignr_epi8(s1, t, 8);           // e f g h
    _mm_storeu_si128((__m128i*)&out[0], s0);
    _mm_storeu_si128((__m128i*)&out[4], s1);
}

void compress(const uint32_t state[8], const uint32_t block[16], uint32_t out[8]) {
    if (using_shani())
        compress_shani(state, block, out);
    else
        compress_portable(state, block, out);
}

}  // namespace chain
}  // namespace bitz
