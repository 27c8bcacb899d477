// Reed-Solomon encode (interleaved additive NTT) and BLAKE3 Merkle commit.
//
// Codeword layout in HBM (the reference's SoA, pcs/commit.rs:142-176):
// cw[pos * L + lane], L = 2^log_lanes lanes of 16 B; one Merkle leaf is one
// position (16L contiguous bytes).
//
// Pipeline for a commit of 2^log_d positions from layer r:
//  1. global radix-2^F passes for layers [r, log_d - T): each thread owns one
//     lane of 2^F positions spaced 2^(log_d-l0-F) apart and runs F butterfly
//     layers in registers; the first pass reads the packed message directly
//     (position p takes message position p mod 2^(log_d-r)), which is the
//     replicate-fill of commit.rs:239-258 without a separate copy;
//  2. one LDS-tiled pass for the deepest layers [log_d - T, log_d): a block
//     owns 2^T consecutive positions, runs the layers in LDS, writes the
//     codeword tile, hashes its 2^T leaves from LDS and builds its T-level
//     subtree;
//  3. the Merkle levels above the tiles.
#include "common.h"
#include "gf128_dev.h"
#include "blake3_dev.h"

namespace {

constexpr int kThreads = 256;

// ---- twiddles ----
// twiddle(l, b) = sum_j bit_j(b) * evals[ell-l-1][1+j] (additive_ntt_f128.rs:153-156)
// is F2-linear in b, so it is stored per layer as two span tables over the
// low kTwLoBits bits of b and the bits above: twiddle(l, b) = lo[b & mask] ^
// hi[b >> kTwLoBits]. Buffer: 64 u64 layer offsets (in 16-B units from the
// buffer start), then the tables; O(ell * 2^max(12, ell-12)) entries in all.
constexpr int kTwLoBits = 12;

__device__ __host__ __forceinline__ int tw_lo_bits(int layer) {
    return layer < kTwLoBits ? layer : kTwLoBits;
}

__global__ void k_twiddle_layer(const uint4* __restrict__ row /* evals[ell-l-1] */, int layer,
                                uint4* __restrict__ out /* lo table then hi table */) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int lo_bits = tw_lo_bits(layer), hi_bits = layer - lo_bits;
    uint64_t n_lo = 1ull << lo_bits, n_hi = 1ull << hi_bits;
    if (idx >= n_lo + n_hi) return;
    int first = idx < n_lo ? 0 : lo_bits;
    uint64_t bits = idx < n_lo ? idx : idx - n_lo;
    int count = idx < n_lo ? lo_bits : hi_bits;
    uint32_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;
    for (int j = 0; j < count; j++) {
        if ((bits >> j) & 1) {
            uint4 e = row[1 + first + j];
            x0 ^= e.x;
            x1 ^= e.y;
            x2 ^= e.z;
            x3 ^= e.w;
        }
    }
    out[idx] = make_uint4(x0, x1, x2, x3);
}

__device__ __forceinline__ gf128 twiddle(const uint4* tw, int layer, uint64_t block) {
    uint64_t off = ((const uint64_t*)tw)[layer];
    int lo_bits = tw_lo_bits(layer);
    gf128 a = gf_load(&tw[off + (block & ((1ull << lo_bits) - 1))]);
    gf128 b = gf_load(&tw[off + (1ull << lo_bits) + (block >> lo_bits)]);
    return gf_add(a, b);
}

// ---- global radix-2^F pass ----
template <int F>
__global__ __launch_bounds__(kThreads) void k_ntt_pass(uint4* __restrict__ data,
                                                       const uint4* __restrict__ src,
                                                       uint64_t src_pos_mask, int log_d,
                                                       int log_lanes, int l0,
                                                       const uint4* __restrict__ tw) {
    uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t total = 1ull << (log_d - F + log_lanes);
    if (tid >= total) return;
    uint64_t lane = tid & ((1ull << log_lanes) - 1);
    uint64_t g = tid >> log_lanes;
    int sh = log_d - l0 - F;
    uint64_t offset = g & ((1ull << sh) - 1);
    uint64_t blk0 = g >> sh;
    uint64_t base = (blk0 << (log_d - l0)) + offset;
    gf128 v[1 << F];
#pragma unroll
    for (int j = 0; j < (1 << F); j++) {
        uint64_t p = base + ((uint64_t)j << sh);
        if (src)
            v[j] = gf_load(&src[((p & src_pos_mask) << log_lanes) + lane]);
        else
            v[j] = gf_load(&data[(p << log_lanes) + lane]);
    }
#pragma unroll
    for (int f = 0; f < F; f++) {
        int layer = l0 + f;
        int half = 1 << (F - f - 1);
#pragma unroll
        for (int j = 0; j < (1 << F); j++) {
            if (j & half) continue;
            uint64_t p = base + ((uint64_t)j << sh);
            gf128 t = twiddle(tw, layer, p >> (log_d - layer));
            gf128 u = gf_add(v[j], gf_mul(v[j | half], t));
            v[j] = u;
            v[j | half] = gf_add(v[j | half], u);
        }
    }
#pragma unroll
    for (int j = 0; j < (1 << F); j++) {
        uint64_t p = base + ((uint64_t)j << sh);
        gf_store(&data[(p << log_lanes) + lane], v[j]);
    }
}

// The same pass for blocks whose threads all share blk0 (2^(sh + log_lanes)
// a multiple of the block size): then every twiddle of the block is one of
// 2^f constants per layer f, the twiddles of layers l0..l0+F-1 at blocks
// (blk0 << f) | k, k < 2^f. They become 4-bit window tables in LDS (built by
// 32 threads each, eight at a time): a butterfly's product is a table
// product (~125 VALU) instead of a variable one (~430). Layers 0-2 use
// tables 0..6; for F = 4, layer 3's eight tables replace them afterwards.
template <int F>
__global__ __launch_bounds__(kThreads) void k_ntt_pass_tab(uint4* __restrict__ data,
                                                           const uint4* __restrict__ src,
                                                           uint64_t src_pos_mask, int log_d,
                                                           int log_lanes, int l0,
                                                           const uint4* __restrict__ tw) {
    __shared__ uint4 tabs[8][512];
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t lane = tid & ((1ull << log_lanes) - 1);
    const uint64_t g = tid >> log_lanes;
    const int sh = log_d - l0 - F;
    const uint64_t offset = g & ((1ull << sh) - 1);
    const uint64_t blk0 = g >> sh;  // the same for the whole block
    const uint64_t base = (blk0 << (log_d - l0)) + offset;
    const int grp = threadIdx.x >> 5, gt = threadIdx.x & 31;
    // tables of layers 0..min(F,3)-1: table (1 << f) - 1 + k
    if (grp < 7) {
        const int f = grp == 0 ? 0 : (grp < 3 ? 1 : 2);
        const int k = grp - ((1 << f) - 1);
        if (f < F) gf_table_build(tabs[grp], twiddle(tw, l0 + f, (blk0 << f) | (uint64_t)k), gt, 32);
    }
    gf128 v[1 << F];
#pragma unroll
    for (int j = 0; j < (1 << F); j++) {
        uint64_t p = base + ((uint64_t)j << sh);
        if (src)
            v[j] = gf_load(&src[((p & src_pos_mask) << log_lanes) + lane]);
        else
            v[j] = gf_load(&data[(p << log_lanes) + lane]);
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < F && f < 3; f++) {
        const int half = 1 << (F - f - 1);
#pragma unroll
        for (int j = 0; j < (1 << F); j++) {
            if (j & half) continue;
            const uint4* tab = tabs[(1 << f) - 1 + (j >> (F - f))];
            gf128 u = gf_add(v[j], gf_mul_tab_lean(tab, v[j | half]));
            v[j] = u;
            v[j | half] = gf_add(v[j | half], u);
        }
    }
    if constexpr (F == 4) {
        __syncthreads();  // every table of layers 0-2 is read
        gf_table_build(tabs[grp], twiddle(tw, l0 + 3, (blk0 << 3) | (uint64_t)grp), gt, 32);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; j += 2) {
            gf128 u = gf_add(v[j], gf_mul_tab_lean(tabs[j >> 1], v[j | 1]));
            v[j] = u;
            v[j | 1] = gf_add(v[j | 1], u);
        }
    }
#pragma unroll
    for (int j = 0; j < (1 << F); j++) {
        uint64_t p = base + ((uint64_t)j << sh);
        gf_store_nt(&data[(p << log_lanes) + lane], v[j]);
    }
}

// ---- LDS-tiled deepest layers, optional fused leaf hashing + subtree ----
// Tile: 2^T positions x L lanes, position stride L+1 uint4 in LDS (one pad
// slot per position spreads the per-leaf reads of the hashing phase across
// banks). Dynamic LDS: 2^T*(L+1)*16 + 2^T*9*4 bytes.
__global__ __launch_bounds__(kThreads) void k_ntt_local(
    uint4* __restrict__ data, const uint4* __restrict__ src, uint64_t src_pos_mask, int log_d,
    int log_lanes, int T, int l_first, const uint4* __restrict__ tw,
    uint32_t* __restrict__ tree /* nullable: (2N-1)*8 words */) {
    extern __shared__ uint4 smem[];
    const int L = 1 << log_lanes;
    const int stride = L + 1;
    const uint32_t tile_pos = 1u << T;
    uint4* tile = smem;
    uint32_t* cvs = (uint32_t*)(smem + (size_t)tile_pos * stride);
    const uint64_t tile_base = (uint64_t)blockIdx.x << T;

    const uint32_t elems = tile_pos << log_lanes;
    for (uint32_t i = threadIdx.x; i < elems; i += blockDim.x) {
        uint32_t pl = i >> log_lanes, lane = i & (L - 1);
        uint64_t p = tile_base + pl;
        uint4 x = src ? src[((p & src_pos_mask) << log_lanes) + lane]
                      : data[(p << log_lanes) + lane];
        tile[pl * stride + lane] = x;
    }
    __syncthreads();

    const uint32_t pairs = elems >> 1;
    for (int layer = l_first; layer < log_d; layer++) {
        uint32_t half = 1u << (log_d - layer - 1);
        uint32_t lowmask = half - 1;
        for (uint32_t q = threadIdx.x; q < pairs; q += blockDim.x) {
            uint32_t lane = q & (L - 1);
            uint32_t pr = q >> log_lanes;
            uint32_t top = ((pr & ~lowmask) << 1) | (pr & lowmask);
            uint64_t p = tile_base + top;
            gf128 t = twiddle(tw, layer, p >> (log_d - layer));
            uint4* pu = &tile[top * stride + lane];
            uint4* pv = &tile[(top + half) * stride + lane];
            gf128 u = gf_load(pu), v = gf_load(pv);
            u = gf_add(u, gf_mul(v, t));
            v = gf_add(v, u);
            gf_store(pu, u);
            gf_store(pv, v);
        }
        __syncthreads();
    }

    for (uint32_t i = threadIdx.x; i < elems; i += blockDim.x) {
        uint32_t pl = i >> log_lanes, lane = i & (L - 1);
        data[((tile_base + pl) << log_lanes) + lane] = tile[pl * stride + lane];
    }
    if (!tree) return;

    // Leaf chaining values: leaf = 16L bytes = L/4 BLAKE3 blocks (L >= 4),
    // or one partial block for L < 4.
    const uint64_t n_leaves = 1ull << log_d;
    const uint32_t leaf_words = 4u * L;
    for (uint32_t pl = threadIdx.x; pl < tile_pos; pl += blockDim.x) {
        const uint32_t* leaf = (const uint32_t*)&tile[pl * stride];
        uint32_t cv[8];
        b3::iv_init(cv);
        uint32_t nblocks = (leaf_words + 15) / 16;
        for (uint32_t b = 0; b < nblocks; b++) {
            uint32_t m[16];
            uint32_t avail = leaf_words - b * 16;
#pragma unroll
            for (int k = 0; k < 16; k++) m[k] = (uint32_t)k < avail ? leaf[b * 16 + k] : 0u;
            uint32_t flags =
                (b == 0 ? b3::CHUNK_START : 0u) | (b + 1 == nblocks ? b3::CHUNK_END : 0u);
            uint32_t blen = b + 1 == nblocks ? (avail >= 16 ? 64u : avail * 4u) : 64u;
            b3::compress(cv, m, 0, blen, flags);
        }
        uint64_t node = tile_base + pl;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            tree[node * 8 + k] = cv[k];
            cvs[pl * 9 + k] = cv[k];
        }
    }
    __syncthreads();
    // Subtree levels 1..T inside the tile.
    uint64_t level_off = 0, level_len = n_leaves;
    for (int j = 1; j <= T; j++) {
        level_off += level_len;
        level_len >>= 1;
        uint32_t nodes = tile_pos >> j;
        uint32_t cv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t i = threadIdx.x;
        if (i < nodes) {
            uint32_t m[16];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                m[k] = cvs[(2 * i) * 9 + k];
                m[8 + k] = cvs[(2 * i + 1) * 9 + k];
            }
            b3::iv_init(cv);
            b3::compress(cv, m, 0, 64, b3::PARENT);
        }
        __syncthreads();
        if (i < nodes) {
            uint64_t node = level_off + ((uint64_t)blockIdx.x << (T - j)) + i;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                cvs[i * 9 + k] = cv[k];
                tree[node * 8 + k] = cv[k];
            }
        }
        __syncthreads();
    }
}

// ---- Merkle levels ----
__global__ void k_merkle_level(uint32_t* __restrict__ tree, uint64_t read_off, uint64_t write_off,
                               uint64_t nodes) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nodes) return;
    uint32_t m[16];
    const uint4* c = (const uint4*)(tree + (read_off + 2 * i) * 8);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint4 x = c[k];
        m[4 * k] = x.x;
        m[4 * k + 1] = x.y;
        m[4 * k + 2] = x.z;
        m[4 * k + 3] = x.w;
    }
    uint32_t cv[8];
    b3::iv_init(cv);
    b3::compress(cv, m, 0, 64, b3::PARENT);
    uint4* o = (uint4*)(tree + (write_off + i) * 8);
    o[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
    o[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

// Single block finishes all levels from a level of `len` (<= 2*kThreads) nodes.
__global__ __launch_bounds__(kThreads) void k_merkle_top(uint32_t* __restrict__ tree,
                                                         uint64_t read_off, uint64_t len) {
    __shared__ uint32_t cvs[2 * kThreads * 9];
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x)
        for (int k = 0; k < 8; k++) cvs[i * 9 + k] = tree[(read_off + i) * 8 + k];
    __syncthreads();
    uint64_t off = read_off;
    while (len > 1) {
        uint64_t nodes = len >> 1;
        uint64_t write_off = off + len;
        uint32_t cv[8];
        uint32_t i = threadIdx.x;
        if (i < nodes) {
            uint32_t m[16];
            for (int k = 0; k < 8; k++) {
                m[k] = cvs[(2 * i) * 9 + k];
                m[8 + k] = cvs[(2 * i + 1) * 9 + k];
            }
            b3::iv_init(cv);
            b3::compress(cv, m, 0, 64, b3::PARENT);
        }
        __syncthreads();
        if (i < nodes) {
            for (int k = 0; k < 8; k++) {
                cvs[i * 9 + k] = cv[k];
                tree[(write_off + i) * 8 + k] = cv[k];
            }
        }
        __syncthreads();
        off = write_off;
        len = nodes;
    }
}

__global__ void k_leaf_hash_global(const uint32_t* __restrict__ data, uint64_t n,
                                   uint32_t leaf_words, uint32_t* __restrict__ tree) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* leaf = data + i * leaf_words;
    uint32_t cv[8];
    b3::iv_init(cv);
    uint32_t nblocks = leaf_words == 0 ? 1 : (leaf_words + 15) / 16;
    for (uint32_t b = 0; b < nblocks; b++) {
        uint32_t m[16];
        uint32_t avail = leaf_words - b * 16;
        for (int k = 0; k < 16; k++) m[k] = (uint32_t)k < avail ? leaf[b * 16 + k] : 0u;
        uint32_t flags = (b == 0 ? b3::CHUNK_START : 0u) | (b + 1 == nblocks ? b3::CHUNK_END : 0u);
        uint32_t blen = b + 1 == nblocks ? (avail >= 16 ? 64u : avail * 4u) : 64u;
        b3::compress(cv, m, 0, blen, flags);
    }
    for (int k = 0; k < 8; k++) tree[i * 8 + k] = cv[k];
}

}  // namespace

namespace bitz {

static size_t row_off(int ell, int i) { return (size_t)i * ell - (size_t)i * (i - 1) / 2; }

// Device twiddle tables for domain dimension ell, cached on the context.
int ensure_twiddles(bitz_ctx* ctx, int ell, const uint4** out) {
    auto it = ctx->twiddles.find(ell);
    if (it != ctx->twiddles.end()) {
        *out = (const uint4*)it->second->ptr;
        return BITZ_OK;
    }
    BITZ_REQUIRE(ell >= 1 && ell <= 64, BITZ_ERR_SHAPE, "twiddles: ell out of range");
    // evals rows (generate_evals_from_subspace, additive_ntt_f128.rs:67-103)
    std::vector<hgf> ev(row_off(ell, ell) + 1);
    for (int k = 0; k < ell; k++) ev[k] = hgf{k < 64 ? 1ull << k : 0, k >= 64 ? 1ull << (k - 64) : 0};
    for (int i = 1; i < ell; i++) {
        hgf* prev = &ev[row_off(ell, i - 1)];
        hgf* cur = &ev[row_off(ell, i)];
        for (int k = 1; k < ell - (i - 1); k++) cur[k - 1] = hgf_mul(prev[k], hgf_add(prev[k], prev[0]));
    }
    for (int i = 0; i < ell; i++) {
        hgf* row = &ev[row_off(ell, i)];
        hgf inv = hgf_inv(row[0]);
        for (int k = 0; k < ell - i; k++) row[k] = hgf_mul(row[k], inv);
    }
    std::vector<uint64_t> offs(64, 0);
    uint64_t cursor = 64 / 2;  // the offset header occupies 32 uint4 slots
    for (int l = 0; l < ell; l++) {
        offs[l] = cursor;
        int lo = tw_lo_bits(l);
        cursor += (1ull << lo) + (1ull << (l - lo));
    }
    BITZ_REQUIRE(cursor < (1ull << 31), BITZ_ERR_SHAPE, "twiddles: domain too large");
    DevBuf d_ev;
    BITZ_REQUIRE(d_ev.alloc(ev.size() * 16), BITZ_ERR_ALLOC, "twiddle setup: allocation failed");
    auto tab = std::make_unique<DevBuf>();
    BITZ_REQUIRE(tab->alloc(cursor * 16), BITZ_ERR_ALLOC, "twiddle table: allocation failed");
    BITZ_HIP(hipMemcpyAsync(d_ev.ptr, ev.data(), ev.size() * 16, hipMemcpyHostToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(tab->ptr, offs.data(), 64 * 8, hipMemcpyHostToDevice, ctx->stream));
    for (int l = 0; l < ell; l++) {
        int lo = tw_lo_bits(l);
        uint64_t n = (1ull << lo) + (1ull << (l - lo));
        hipLaunchKernelGGL(k_twiddle_layer, dim3((unsigned)((n + kThreads - 1) / kThreads)),
                           dim3(kThreads), 0, ctx->stream,
                           (const uint4*)d_ev.ptr + row_off(ell, ell - l - 1), l,
                           (uint4*)tab->ptr + offs[l]);
        BITZ_LAUNCH_CHECK();
    }
    BITZ_HIP(hipStreamSynchronize(ctx->stream));  // d_ev is freed on return
    *out = (const uint4*)tab->ptr;
    ctx->twiddles[ell] = std::move(tab);
    return BITZ_OK;
}

// cost model of the algorithmic work: one GF(2^128) multiplication = 3
// 64-bit carry-less products + reduction, one BLAKE3 compression = 7 rounds
// x 8 G x 14 ops + 16, both ~800 32-bit integer operations

static int tile_log(int log_d, int log_lanes) { return std::min(std::min(log_d, 11 - log_lanes), 9); }

static size_t local_lds_bytes(int T, int log_lanes) {
    size_t tp = (size_t)1 << T;
    return tp * ((1u << log_lanes) + 1) * 16 + tp * 9 * 4;
}

// Runs layers [start, log_d) of the interleaved NTT, reading the initial
// state from `src` (position p <- src position p & src_mask) when given.
// With `tree`, the LDS pass also hashes the leaves and the tile subtrees.
static int run_ntt(bitz_ctx* ctx, uint4* data, const uint4* src, uint64_t src_mask, int log_d,
                   int log_lanes, int start, int ell, uint32_t* tree) {
    const uint4* tw = nullptr;
    int rc = ensure_twiddles(ctx, ell, &tw);
    if (rc) return rc;
    int T = tile_log(log_d, log_lanes);
    int local_first = std::max(start, log_d - T);
    int l = start;
    while (l < log_d - T) {
        int F = std::min(std::max(1, std::min(4, ctx->tune.ntt_radix_log)), log_d - T - l);
        uint64_t threads = 1ull << (log_d - F + log_lanes);
        dim3 grid((unsigned)((threads + kThreads - 1) / kThreads));
        double elems = (double)(1ull << (log_d + log_lanes));
        TimedLaunch tl(ctx, "ntt_pass", 2.0 * 16.0 * elems, F * (elems / 2) * kGfMulOps);
        // blocks whose threads share their twiddles take them from LDS tables
        const bool tab = (log_d - l - F) + log_lanes >= 8;
        if (tab) {
            switch (F) {
                case 4:
                    hipLaunchKernelGGL(k_ntt_pass_tab<4>, grid, dim3(kThreads), 0, ctx->stream, data, src, src_mask,
                                       log_d, log_lanes, l, tw);
                    break;
                case 3:
                    hipLaunchKernelGGL(k_ntt_pass_tab<3>, grid, dim3(kThreads), 0, ctx->stream, data, src, src_mask,
                                       log_d, log_lanes, l, tw);
                    break;
                case 2:
                    hipLaunchKernelGGL(k_ntt_pass_tab<2>, grid, dim3(kThreads), 0, ctx->stream, data, src, src_mask,
                                       log_d, log_lanes, l, tw);
                    break;
                default:
                    hipLaunchKernelGGL(k_ntt_pass_tab<1>, grid, dim3(kThreads), 0, ctx->stream, data, src, src_mask,
                                       log_d, log_lanes, l, tw);
                    break;
            }
        } else switch (F) {
            case 4:
                hipLaunchKernelGGL(k_ntt_pass<4>, grid, dim3(kThreads), 0, ctx->stream, data, src,
                                   src_mask, log_d, log_lanes, l, tw);
                break;
            case 3:
                hipLaunchKernelGGL(k_ntt_pass<3>, grid, dim3(kThreads), 0, ctx->stream, data, src,
                                   src_mask, log_d, log_lanes, l, tw);
                break;
            case 2:
                hipLaunchKernelGGL(k_ntt_pass<2>, grid, dim3(kThreads), 0, ctx->stream, data, src,
                                   src_mask, log_d, log_lanes, l, tw);
                break;
            default:
                hipLaunchKernelGGL(k_ntt_pass<1>, grid, dim3(kThreads), 0, ctx->stream, data, src,
                                   src_mask, log_d, log_lanes, l, tw);
                break;
        }
        BITZ_LAUNCH_CHECK();
        src = nullptr;
        l += F;
    }
    size_t lds = local_lds_bytes(T, log_lanes);
    unsigned tiles = (unsigned)(1ull << (log_d - T));
    {
        double elems = (double)(1ull << (log_d + log_lanes));
        double leaves = (double)(1ull << log_d);
        double hashes = tree ? leaves * (((16u << log_lanes) + 63) / 64) + leaves : 0;
        TimedLaunch tl(ctx, tree ? "ntt_local_merkle" : "ntt_local",
                       2.0 * 16.0 * elems + (tree ? 64.0 * leaves : 0.0),
                       (log_d - local_first) * (elems / 2) * kGfMulOps + hashes * kBlake3Ops);
        hipLaunchKernelGGL(k_ntt_local, dim3(tiles), dim3(kThreads), lds, ctx->stream, data, src,
                           src_mask, log_d, log_lanes, T, local_first, tw, tree);
        BITZ_LAUNCH_CHECK();
    }
    if (tree) {
        // levels above the tile subtrees
        uint64_t n = 1ull << log_d;
        uint64_t off = 0, len = n;
        for (int j = 0; j < T; j++) {
            off += len;
            len >>= 1;
        }
        while (len > 2 * (uint64_t)kThreads) {
            uint64_t nodes = len >> 1;
            hipLaunchKernelGGL(k_merkle_level, dim3((unsigned)((nodes + kThreads - 1) / kThreads)),
                               dim3(kThreads), 0, ctx->stream, tree, off, off + len, nodes);
            BITZ_LAUNCH_CHECK();
            off += len;
            len = nodes;
        }
        if (len > 1) {
            hipLaunchKernelGGL(k_merkle_top, dim3(1), dim3(kThreads), 0, ctx->stream, tree, off,
                               len);
            BITZ_LAUNCH_CHECK();
        }
    }
    return BITZ_OK;
}

int commit_device(bitz_ctx* ctx, const void* packed, int log_msg_len, int log_batch,
                  int log_inv_rate, void* codeword, void* tree) {
    BITZ_REQUIRE(ctx && packed && codeword && tree, BITZ_ERR_ARG, "commit: null argument");
    // PcsParams::validate (commit.rs:125-136) plus the library's own size limits
    BITZ_REQUIRE(log_batch >= 0 && log_batch <= 6, BITZ_ERR_SHAPE,
                 "commit: log_batch must be in [0, 6] (leaf <= 1024 B)");
    BITZ_REQUIRE(log_msg_len > log_batch, BITZ_ERR_SHAPE,
                 "commit: m too small (need m >= LOG_PACKING + log_batch_size + 1)");
    BITZ_REQUIRE(log_inv_rate >= 1, BITZ_ERR_SHAPE,
                 "commit: log_inv_rate must be >= 1 for a non-trivial RS code");
    int log_dim = log_msg_len - log_batch;
    int k_code = log_dim + log_inv_rate;
    BITZ_REQUIRE(k_code <= 40, BITZ_ERR_SHAPE, "commit: codeword too large");
    BITZ_HIP(hipSetDevice(ctx->device));
    return run_ntt(ctx, (uint4*)codeword, (const uint4*)packed, (1ull << log_dim) - 1, k_code,
                   log_batch, log_inv_rate, k_code, (uint32_t*)tree);
}

}  // namespace bitz

extern "C" int bitz_ntt_forward_dev(bitz_ctx* ctx, void* data, int log_d, int log_lanes,
                                    int start_layer, int ell) {
    BITZ_REQUIRE(ctx && data, BITZ_ERR_ARG, "bitz_ntt_forward_dev: null argument");
    BITZ_REQUIRE(log_d >= 1 && log_d <= ell && ell <= 64 && log_d <= 40, BITZ_ERR_SHAPE,
                 "bitz_ntt_forward_dev: need 1 <= log_d <= ell <= 64");
    BITZ_REQUIRE(start_layer >= 0 && start_layer <= log_d, BITZ_ERR_SHAPE,
                 "bitz_ntt_forward_dev: start_layer out of range");
    BITZ_REQUIRE(log_lanes >= 0 && log_lanes <= 6, BITZ_ERR_SHAPE,
                 "bitz_ntt_forward_dev: log_lanes must be in [0, 6]");
    if (start_layer == log_d) return BITZ_OK;
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    return bitz::run_ntt(ctx, (uint4*)data, nullptr, ~0ull, log_d, log_lanes, start_layer, ell,
                         nullptr);
}

extern "C" int bitz_merkle_tree_dev(bitz_ctx* ctx, const void* data, uint64_t num_leaves,
                                    uint64_t leaf_size, void* tree) {
    BITZ_REQUIRE(ctx && data && tree, BITZ_ERR_ARG, "bitz_merkle_tree_dev: null argument");
    BITZ_REQUIRE(num_leaves > 0 && (num_leaves & (num_leaves - 1)) == 0, BITZ_ERR_SHAPE,
                 "bitz_merkle_tree_dev: num_leaves must be a power of 2");
    BITZ_REQUIRE(leaf_size <= 1024 && leaf_size % 4 == 0, BITZ_ERR_SHAPE,
                 "bitz_merkle_tree_dev: leaf_size must be a multiple of 4 and <= 1024");
    BITZ_HIP(hipSetDevice(ctx->device));
    uint32_t* t = (uint32_t*)tree;
    hipLaunchKernelGGL(k_leaf_hash_global, dim3((unsigned)((num_leaves + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, ctx->stream, (const uint32_t*)data, num_leaves,
                       (uint32_t)(leaf_size / 4), t);
    BITZ_LAUNCH_CHECK();
    uint64_t off = 0, len = num_leaves;
    while (len > 2 * (uint64_t)kThreads) {
        uint64_t nodes = len >> 1;
        hipLaunchKernelGGL(k_merkle_level, dim3((unsigned)((nodes + kThreads - 1) / kThreads)),
                           dim3(kThreads), 0, ctx->stream, t, off, off + len, nodes);
        BITZ_LAUNCH_CHECK();
        off += len;
        len = nodes;
    }
    if (len > 1) {
        hipLaunchKernelGGL(k_merkle_top, dim3(1), dim3(kThreads), 0, ctx->stream, t, off, len);
        BITZ_LAUNCH_CHECK();
    }
    return BITZ_OK;
}

extern "C" int bitz_commit_dev(bitz_ctx* ctx, const void* packed, int log_msg_len, int log_batch,
                               int log_inv_rate, void* codeword, void* tree, uint8_t* root) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_commit_dev: null context");
    bitz::PoolScope scope(&ctx->pool);
    int rc = bitz::commit_device(ctx, packed, log_msg_len, log_batch, log_inv_rate, codeword, tree);
    if (rc) return rc;
    if (root) {
        uint64_t n = 1ull << (log_msg_len - log_batch + log_inv_rate);
        BITZ_HIP(hipMemcpyAsync(root, (const uint8_t*)tree + (2 * n - 2) * 32, 32,
                                hipMemcpyDeviceToHost, ctx->stream));
        BITZ_HIP(hipStreamSynchronize(ctx->stream));
    }
    return BITZ_OK;
}

// ---- host-buffer commit with a device-resident handle ----

static int commit_impl(bitz_ctx* ctx, const uint64_t* rows, bool rows_on_device, uint32_t t, uint32_t s,
                       uint32_t log_batch, uint32_t log_inv_rate, uint8_t root[32], bitz_commit_handle** out);

extern "C" int bitz_commit(bitz_ctx* ctx, const uint64_t* rows, uint32_t t, uint32_t s,
                           uint32_t log_batch, uint32_t log_inv_rate, uint8_t root[32],
                           bitz_commit_handle** out) {
    return commit_impl(ctx, rows, false, t, s, log_batch, log_inv_rate, root, out);
}

extern "C" int bitz_commit_device_rows(bitz_ctx* ctx, const void* rows, uint32_t t, uint32_t s,
                                       uint32_t log_batch, uint32_t log_inv_rate, uint8_t root[32],
                                       bitz_commit_handle** out) {
    return commit_impl(ctx, (const uint64_t*)rows, true, t, s, log_batch, log_inv_rate, root, out);
}

static int commit_impl(bitz_ctx* ctx, const uint64_t* rows, bool rows_on_device, uint32_t t, uint32_t s,
                       uint32_t log_batch, uint32_t log_inv_rate, uint8_t root[32], bitz_commit_handle** out) {
    BITZ_REQUIRE(ctx && rows && root && out, BITZ_ERR_ARG, "bitz_commit: null argument");
    // Shape::new (src/bitz/params.rs:53-66) and the packing width
    BITZ_REQUIRE(t >= 7 && t <= 35 && s <= 35 - t, BITZ_ERR_SHAPE,
                 "bitz_commit: shape needs 7 <= t and t + s <= 35");
    int log_msg = (int)(t + s) - 7;
    bitz::PoolScope scope(&ctx->pool);
    bitz::PhaseClock phase(ctx);
    phase.mark("commit");
    auto h = std::make_unique<bitz_commit_handle>();
    h->ctx = ctx;
    h->t = t;
    h->s = s;
    h->log_batch = log_batch;
    h->log_inv_rate = log_inv_rate;
    h->n_leaves = 1ull << (log_msg - (int)log_batch + (int)log_inv_rate);
    BITZ_REQUIRE(log_batch <= 6 && (int)log_batch < log_msg && log_inv_rate >= 1 && log_inv_rate <= 8,
                 BITZ_ERR_SHAPE, "bitz_commit: need log_batch <= 6, log_batch < t+s-7, 1 <= log_inv_rate <= 8");
    size_t msg_bytes = (size_t)16 << log_msg;
    BITZ_REQUIRE(h->packed.alloc(msg_bytes) && h->codeword.alloc(msg_bytes << log_inv_rate) &&
                     h->tree.alloc((2 * h->n_leaves - 1) * 32),
                 BITZ_ERR_ALLOC, "bitz_commit: device allocation failed");
    BITZ_HIP(hipSetDevice(ctx->device));
    BITZ_HIP(hipMemcpyAsync(h->packed.ptr, rows, msg_bytes,
                            rows_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    int rc = bitz_commit_dev(ctx, h->packed.ptr, log_msg, (int)log_batch, (int)log_inv_rate,
                             h->codeword.ptr, h->tree.ptr, root);
    if (rc) return rc;
    memcpy(h->root, root, 32);
    *out = h.release();
    return BITZ_OK;
}

extern "C" void bitz_commit_free(bitz_commit_handle* h) { delete h; }
extern "C" const void* bitz_commit_packed(const bitz_commit_handle* h) { return h ? h->packed.ptr : nullptr; }
extern "C" const void* bitz_commit_codeword(const bitz_commit_handle* h) { return h ? h->codeword.ptr : nullptr; }
extern "C" const void* bitz_commit_tree(const bitz_commit_handle* h) { return h ? h->tree.ptr : nullptr; }
extern "C" int bitz_commit_read_tree(bitz_commit_handle* h, uint8_t* out, uint64_t bytes) {
    BITZ_REQUIRE(h && out, BITZ_ERR_ARG, "bitz_commit_read_tree: null argument");
    BITZ_REQUIRE(bytes <= (2 * h->n_leaves - 1) * 32, BITZ_ERR_ARG, "bitz_commit_read_tree: too many bytes");
    BITZ_HIP(hipMemcpyAsync(out, h->tree.ptr, bytes, hipMemcpyDeviceToHost, h->ctx->stream));
    BITZ_HIP(hipStreamSynchronize(h->ctx->stream));
    return BITZ_OK;
}
