// BLAKE3 compression on CDNA4 (public BLAKE3 specification), register
// resident: 16 state words + 16 message words per thread, rotations map to
// v_alignbit_b32. Used for Merkle leaves (non-root chunk chaining values,
// counter 0) and parents (non-root PARENT compressions), matching the
// reference's hazmat calls in vendor/flock-mod/crates/flock-core/src/merkle.rs:192-202,
// and for the proof-of-work scan.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b3 {

enum : uint32_t { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

__device__ __constant__ static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u,
                                                      0xA54FF53Au, 0x510E527Fu, 0x9B05688Cu,
                                                      0x1F83D9ABu, 0x5BE0CD19u};

__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

#define B3_G(a, b, c, d, mx, my)      \
    do {                              \
        s[a] = s[a] + s[b] + (mx);    \
        s[d] = rotr(s[d] ^ s[a], 16); \
        s[c] = s[c] + s[d];           \
        s[b] = rotr(s[b] ^ s[c], 12); \
        s[a] = s[a] + s[b] + (my);    \
        s[d] = rotr(s[d] ^ s[a], 8);  \
        s[c] = s[c] + s[d];           \
        s[b] = rotr(s[b] ^ s[c], 7);  \
    } while (0)

// Message schedule: word index used by round r at slot i is sigma_r(i), with
// sigma_0 = identity and sigma_{r+1} = sigma_r o PERM.
__device__ __forceinline__ void round_fn(uint32_t* s, const uint32_t* m) {
    B3_G(0, 4, 8, 12, m[0], m[1]);
    B3_G(1, 5, 9, 13, m[2], m[3]);
    B3_G(2, 6, 10, 14, m[4], m[5]);
    B3_G(3, 7, 11, 15, m[6], m[7]);
    B3_G(0, 5, 10, 15, m[8], m[9]);
    B3_G(1, 6, 11, 12, m[10], m[11]);
    B3_G(2, 7, 8, 13, m[12], m[13]);
    B3_G(3, 4, 9, 14, m[14], m[15]);
}

__device__ __forceinline__ void permute(uint32_t* m) {
    uint32_t t[16];
    t[0] = m[2];
    t[1] = m[6];
    t[2] = m[3];
    t[3] = m[10];
    t[4] = m[7];
    t[5] = m[0];
    t[6] = m[4];
    t[7] = m[13];
    t[8] = m[1];
    t[9] = m[11];
    t[10] = m[12];
    t[11] = m[5];
    t[12] = m[9];
    t[13] = m[14];
    t[14] = m[15];
    t[15] = m[8];
#pragma unroll
    for (int i = 0; i < 16; i++) m[i] = t[i];
}

// Compress in place: cv (8 words) <- first 8 output words (chaining value).
// If out_hi != nullptr it also receives output words 8..15 (XOF/root use).
__device__ __forceinline__ void compress(uint32_t* cv, const uint32_t* block_words, uint64_t counter,
                                         uint32_t block_len, uint32_t flags,
                                         uint32_t* out_hi = nullptr) {
    uint32_t s[16], m[16];
#pragma unroll
    for (int i = 0; i < 16; i++) m[i] = block_words[i];
#pragma unroll
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    s[8] = IV[0];
    s[9] = IV[1];
    s[10] = IV[2];
    s[11] = IV[3];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
#pragma unroll
    for (int r = 0; r < 7; r++) {
        round_fn(s, m);
        if (r < 6) permute(m);
    }
    if (out_hi) {
#pragma unroll
        for (int i = 0; i < 8; i++) out_hi[i] = s[i + 8] ^ cv[i];
    }
#pragma unroll
    for (int i = 0; i < 8; i++) cv[i] = s[i] ^ s[i + 8];
}

#undef B3_G

__device__ __forceinline__ void iv_init(uint32_t* cv) {
#pragma unroll
    for (int i = 0; i < 8; i++) cv[i] = IV[i];
}

}  // namespace b3
