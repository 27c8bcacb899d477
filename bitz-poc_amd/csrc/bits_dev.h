// Wavefront bit transpose shared by the bit-layout kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bitz {

// 64 x 64 bit transpose across a wavefront (64 lanes, all active): lane g
// holds row g on entry, lane k holds sum_g bit_k(row g) << g on exit. Column
// k of the block is the wave's ballot of bit k (a compare into a scalar
// mask), selected into lane k: scalar and VALU work only. (A shuffle-based
// transpose through ds_bpermute gave nondeterministic words at full
// occupancy on gfx950, where the compiler reused a shuffle's source
// register before the shuffle had read it.)
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t v, int lane) {
    uint64_t o = 0;
#pragma unroll
    for (int k = 0; k < 64; k++) {
        const uint64_t m = __ballot((v >> k) & 1ull);
        o = lane == k ? m : o;
    }
    return o;
}

}  // namespace bitz
