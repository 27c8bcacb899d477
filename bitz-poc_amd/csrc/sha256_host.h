// The SHA-256 compression function on the host (FIPS 180-4 6.2.2), for the
// chain's statement (every chaining value H_{i+1} = Compress(H_i, M_i)):
// the x86 SHA extensions where the CPU has them, portable code otherwise,
// chosen at run time with identical results.
#pragma once
#include <stdint.h>

namespace bitz {
namespace chain {

// FIPS 180-4 4.2.2
extern const uint32_t kSha256K[64];

// out = Compress(state, block); block as 16 message words (big-endian words
// of the 64-byte block, as integers). out may alias state.
void compress(const uint32_t state[8], const uint32_t block[16], uint32_t out[8]);
void compress_portable(const uint32_t state[8], const uint32_t block[16], uint32_t out[8]);
void compress_shani(const uint32_t state[8], const uint32_t block[16], uint32_t out[8]);
bool using_shani();

}  // namespace chain
}  // namespace bitz
