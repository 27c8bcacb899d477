// Block- and grid-level XOR reductions of GF(2^128) sums. Addition in
// characteristic two is XOR, so partial sums combine in any order and the
// result does not depend on the launch shape. Wide (unreduced 256-bit)
// accumulators are reduced once per thread before the block reduction:
// reduction mod the field polynomial is F2-linear.
#pragma once
#include "gf128_dev.h"

namespace bitz {

constexpr int kRedThreads = 256;

// Reduce K field values across the block; thread 0 writes them to out[0..K).
template <int K>
__device__ __forceinline__ void block_xor_store(gf128 (&v)[K], uint4* __restrict__ out) {
    __shared__ uint4 sh[K][kRedThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
#pragma unroll
        for (int w = 0; w < 4; w++) {
            uint32_t x = v[k].w[w];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) x ^= __shfl_xor(x, off, 64);
            v[k].w[w] = x;
        }
        if (lane == 0) sh[k][wid] = make_uint4(v[k].w[0], v[k].w[1], v[k].w[2], v[k].w[3]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            uint4 a = sh[k][0];
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
                uint4 b = sh[k][w];
                a.x ^= b.x;
                a.y ^= b.y;
                a.z ^= b.z;
                a.w ^= b.w;
            }
            out[k] = a;
        }
    }
}

// partials[b*K + k] for b < nblocks -> out[k]; one block.
template <int K>
__global__ void k_xor_partials(const uint4* __restrict__ partials, uint32_t nblocks,
                               uint4* __restrict__ out) {
    gf128 acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = gf_zero();
    for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = gf_add(acc[k], gf_load(&partials[(size_t)b * K + k]));
    }
    block_xor_store<K>(acc, out);
}

// The same reduction, also written to a host-mapped mailbox: the K values,
// then a system-scope fence, then the sequence flag the host polls on.
template <int K>
__global__ void k_xor_partials_mbox(const uint4* __restrict__ partials, uint32_t nblocks, uint4* __restrict__ out,
                                    uint4* mbox_vals, uint32_t* mbox_flag, uint32_t seq) {
    __shared__ uint4 res[K];
    gf128 acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = gf_zero();
    for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = gf_add(acc[k], gf_load(&partials[(size_t)b * K + k]));
    }
    block_xor_store<K>(acc, res);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            out[k] = res[k];
            mbox_vals[k] = res[k];
        }
        __threadfence_system();
        *(volatile uint32_t*)mbox_flag = seq;
    }
}

}  // namespace bitz
