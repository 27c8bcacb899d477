// The low levels of the batched grand-product GKR, run off 16-bit row words
// (src/bitz/forest.rs: table-driven levels; their values are functions of a
// few committed bits, so they are never materialised).
//
// Row words. With Y4 = 2^(t-4), the word of slot (y, c), y < Y4, holds the
// bits of column c at the sixteen rows y + w Y4 (w < 16): its low byte the
// rows of even w (bit i <-> w = 2i), its high byte those of odd w (bit i <->
// w = 2i + 1). Stored as u16, nr[(y << s) | c].
//
// A level-l layer (l = 0: the leaves, 1, 2) has halves E (product bit p = 0)
// and O (p = 1) of 2^(t-l-1) row positions each; the GKR binds their row
// bits top down. After k = 2 - l of its rounds (challenges r_0..r_{k-1}),
// round k binds the next bit h, and each of E_h, O_h at (y, c) is a function
// of four bits of the slot's word: the rows y + Y4 (h + 2v + 2^(k+1) p +
// 2^(k+2) u) over the bound bits v < 2^k and the level's leaves u < 2^l,
//   T_{h,p}[j] = sum_v eq(r, v) prod_u (bit j_{v + 2^k u} ? A[row] : 1),
// with v's top bit bound by r_0. In a row word's bytes those four bits are
// the bits Em of one byte (h = 0: the low byte, h = 1: the high byte) for
// E, and the bits Om = Em << sh for O:
//   l = 0 (k = 2): Em = 0x0f, sh = 4;  l = 1 (k = 1): Em = 0x33, sh = 2;
//   l = 2 (k = 0): Em = 0x55, sh = 1.
// So round k's products E_a O_b are functions of one byte per slot,
// P_ab[i] = T_{a,0}[ext(i & Em)] T_{b,1}[ext((i & Om) >> sh)], at i = the
// low byte (ab = 00), the high byte (11), or the mix of the two (01, 10),
// and a row's sums are those tables against the eq_c buckets of its columns.
// The fold by r_k is a lookup in two 256-entry tables per y,
// F_p[i] = (1 + r_k) T_{0,p}[ext(i & Em)] + r_k T_{1,p}[ext((i & Om) >> sh)],
// into halves of 2^(t+s-4) entries, where the layer's dense rounds start.
#include "common.h"
#include "gf128_dev.h"
#include "open_impl.h"
#include "reduce_dev.h"

namespace bitz {
namespace {

constexpr int kT = 256;

__device__ __forceinline__ uint64_t grid_tid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t grid_threads() { return (uint64_t)gridDim.x * blockDim.x; }

unsigned blocks_for(uint64_t n, uint64_t per_thread = 1, unsigned cap = 4096) {
    uint64_t threads = (n + per_thread - 1) / per_thread;
    uint64_t b = (threads + kT - 1) / kT;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// The bit layout of level L's patterns in a row-word byte.
template <int L>
struct Nib {
    static constexpr int k = 2 - L;                                 // rounds bound before
    static constexpr uint32_t em = L == 0 ? 0x0fu : L == 1 ? 0x33u : 0x55u;  // E bits of a byte
    static constexpr int sh = L == 0 ? 4 : L == 1 ? 2 : 1;          // O bits = E bits << sh
    static constexpr uint32_t om = em << sh;
    // E-position bits of a byte (x & em) -> the pattern index j = v + 2^k u
    __device__ __forceinline__ static uint32_t ext(uint32_t x) {
        if (L == 0) return x & 15u;
        if (L == 1) return (x & 3u) | ((x >> 2) & 12u);
        return (x & 1u) | ((x >> 1) & 2u) | ((x >> 2) & 4u) | ((x >> 3) & 8u);
    }
};

// Challenges of the layer's rounds: host values as scalars, or (with the
// transcript on the device) a device array.
struct Chal {
    uint64_t lo[3], hi[3];
    const uint4* dev;
};
__device__ __forceinline__ gf128 chal(const Chal& c, int i) {
    return c.dev ? gf_load(&c.dev[i]) : gf_from_u64(c.lo[i], c.hi[i]);
}

// rows y + w Y4 of column group g -> the word's two bytes
__global__ __launch_bounds__(kT) void k_nibble_rows(const uint64_t* __restrict__ packed, int t, int s,
                                                    uint16_t* __restrict__ nr) {
    const uint64_t y4 = 1ull << (t - 4), nrows = 1ull << t, n = y4 << s;
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) {
        const uint64_t y = x >> s, c = x & ((1ull << s) - 1);
        const uint64_t* col = packed + (c >> 6) * nrows + y;
        const int bit = (int)(c & 63);
        uint32_t w = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            w |= (uint32_t)((col[(2 * i) * y4] >> bit) & 1ull) << i;
            w |= (uint32_t)((col[(2 * i + 1) * y4] >> bit) & 1ull) << (8 + i);
        }
        nr[x] = (uint16_t)w;
    }
}

// T_{h,p}[j] of row y (e = (h*2 + p)*16 + j)
template <int L>
__device__ __forceinline__ gf128 nibble_tval(const uint4* __restrict__ A, int t, const Chal& ch, uint64_t y, uint32_t e) {
    constexpr int k = Nib<L>::k;
    const uint64_t y4 = 1ull << (t - 4);
    const uint32_t h = (e >> 5) & 1u, p = (e >> 4) & 1u, j = e & 15u;
    gf128 acc = gf_zero();
#pragma unroll
    for (uint32_t v = 0; v < (1u << k); v++) {
        gf128 term = gf_one();
        bool first = true;
#pragma unroll
        for (uint32_t u = 0; u < (1u << L); u++) {
            if ((j >> (v + (u << k))) & 1u) {
                const uint64_t w = h + 2 * v + (p << (k + 1)) + ((uint64_t)u << (k + 2));
                const gf128 a = gf_load(&A[y + w * y4]);
                term = first ? a : gf_mul(term, a);
                first = false;
            }
        }
        if (k == 0) {
            acc = term;
        } else {
            // eq(r, v): v's top bit is bound by r_0
            gf128 wv = gf_one();
            bool wfirst = true;
#pragma unroll
            for (int i = 0; i < k; i++) {
                const gf128 r = chal(ch, i);
                const gf128 f = ((v >> (k - 1 - i)) & 1u) ? r : gf_add(r, gf_one());
                wv = wfirst ? f : gf_mul(wv, f);
                wfirst = false;
            }
            acc = gf_add(acc, first ? wv : gf_mul(wv, term));
        }
    }
    return acc;
}

// T[((y*2 + h)*2 + p)*16 + j], unscaled
template <int L>
__global__ __launch_bounds__(kT) void k_nibble_ttab(const uint4* __restrict__ A, int t, Chal ch,
                                                    uint4* __restrict__ T) {
    const uint64_t n = (1ull << (t - 4)) * 64;
    for (uint64_t idx = grid_tid(); idx < n; idx += grid_threads())
        gf_store(&T[idx], nibble_tval<L>(A, t, ch, idx >> 6, (uint32_t)idx & 63u));
}

// Round sums by buckets. Block y: each column scatters eq_c[c] into the
// buckets of its four indices (B_ab[i] = sum of eq_c over the row's columns
// whose ab-index is i, one LDS array per 32-bit word so that random indices
// spread over the banks); then
//   sum_ab = sum_i P_ab[i] B_ab[i] = sum_e T_{a,0}[e] sum_o T_{b,1}[o] B_ab[i(e, o)]
// takes 1280 products per row instead of one per slot, with T_{a,0}
// pre-scaled by eq_y[y]. end = sum_{ss}, ss = send_one; inf = sum over ab.
template <int L>
__device__ __forceinline__ uint32_t nibble_place(uint32_t j) {
    if (L == 0) return j;
    if (L == 1) return (j & 3u) | ((j & 12u) << 2);
    return (j & 1u) | ((j & 2u) << 1) | ((j & 4u) << 2) | ((j & 8u) << 3);
}

template <int L>
__global__ __launch_bounds__(kT) void k_nibble_bucket_sums(const uint16_t* __restrict__ nr, const uint4* __restrict__ A,
                                                           int t, int s, Chal ch, const uint4* __restrict__ eq_c,
                                                           const uint4* __restrict__ eq_y, int send_one,
                                                           uint4* __restrict__ partials) {
    using N = Nib<L>;
    __shared__ uint32_t B[4][4][256];  // [word][ab][index]
    __shared__ uint4 T[64];
    const uint64_t y = blockIdx.x, cols = 1ull << s;
    for (uint32_t i = threadIdx.x; i < 4 * 4 * 256; i += kT) (&B[0][0][0])[i] = 0u;
    if (threadIdx.x < 64) {
        gf128 v = nibble_tval<L>(A, t, ch, y, threadIdx.x);
        if (((threadIdx.x >> 4) & 1u) == 0) v = gf_mul(v, gf_load(&eq_y[y]));  // p = 0: T_{a,0} eq_y[y]
        gf_store(&T[threadIdx.x], v);
    }
    __syncthreads();
    for (uint64_t c = threadIdx.x; c < cols; c += kT) {
        const uint32_t w = nr[(y << s) | c];
        const uint32_t lo = w & 255u, hi = w >> 8;
        const uint32_t ix[4] = {lo, (lo & N::em) | (hi & N::om), (hi & N::em) | (lo & N::om), hi};
        const uint4 e = eq_c[c];
#pragma unroll
        for (int ab = 0; ab < 4; ab++) {
            atomicXor(&B[0][ab][ix[ab]], e.x);
            atomicXor(&B[1][ab][ix[ab]], e.y);
            atomicXor(&B[2][ab][ix[ab]], e.z);
            atomicXor(&B[3][ab][ix[ab]], e.w);
        }
    }
    __syncthreads();
    // thread (ab, e, oq): T_{a,0}[e] * sum over its four o of T_{b,1}[o] B_ab[i(e, o)]
    const uint32_t ab = threadIdx.x >> 6, a = ab >> 1, b = ab & 1u, ej = (threadIdx.x >> 2) & 15u,
                   oq = threadIdx.x & 3u;
    gf256 acc = gf256_zero();
#pragma unroll
    for (uint32_t o = oq * 4; o < oq * 4 + 4; o++) {
        const uint32_t i = nibble_place<L>(ej) | (nibble_place<L>(o) << N::sh);
        const gf128 bv = gf128{{B[0][ab][i], B[1][ab][i], B[2][ab][i], B[3][ab][i]}};
        gf256_acc(&acc, gf_mul_wide(gf_load(&T[(b * 2 + 1) * 16 + o]), bv));
    }
    const gf128 v = gf_mul(gf_reduce(acc), gf_load(&T[(a * 2) * 16 + ej]));
    gf128 sums[2] = {(a == (uint32_t)send_one && b == (uint32_t)send_one) ? v : gf_zero(), v};
    block_xor_store<2>(sums, partials + (size_t)blockIdx.x * 2);
}

// F[(y*2 + p)*256 + i] = (1 + r) T_{0,p}[ext(i & em)] + r T_{1,p}[ext((i & om) >> sh)], r = challenge k
template <int L>
__global__ __launch_bounds__(kT) void k_nibble_ftab(const uint4* __restrict__ T, int t, Chal ch,
                                                    uint4* __restrict__ F) {
    using N = Nib<L>;
    const uint64_t n = (1ull << (t - 4)) * 512;
    const gf128 r = chal(ch, N::k), r1 = gf_add(r, gf_one());
    for (uint64_t idx = grid_tid(); idx < n; idx += grid_threads()) {
        const uint64_t y = idx >> 9;
        const uint32_t p = (uint32_t)(idx >> 8) & 1u, i = (uint32_t)idx & 255u;
        const gf128 lo = gf_load(&T[((y * 2 + 0) * 2 + p) * 16 + N::ext(i & N::em)]);
        const gf128 hi = gf_load(&T[((y * 2 + 1) * 2 + p) * 16 + N::ext((i & N::om) >> N::sh)]);
        gf_store(&F[idx], gf_add(gf_mul(r1, lo), gf_mul(r, hi)));
    }
}

template <int L>
__device__ __forceinline__ void nibble_fold_idx(uint32_t w, uint32_t* ie, uint32_t* io) {
    using N = Nib<L>;
    const uint32_t lo = w & 255u, hi = w >> 8;
    *ie = (lo & N::em) | ((hi & N::em) << N::sh);
    *io = ((lo & N::om) >> N::sh) | (hi & N::om);
}

// The folded halves: l[x] = F_0[y][ie], r[x] = F_1[y][io], x = (y << s) | c.
template <int L>
__global__ __launch_bounds__(kT) void k_nibble_fold(const uint16_t* __restrict__ nr, const uint4* __restrict__ F,
                                                    int t, int s, uint4* __restrict__ l_out,
                                                    uint4* __restrict__ r_out) {
    const uint64_t n = (1ull << (t - 4)) << s;
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) {
        uint32_t ie, io;
        nibble_fold_idx<L>(nr[x], &ie, &io);
        const uint64_t y = x >> s;
        gf_store_nt(&l_out[x], gf_load(&F[y * 512 + ie]));
        gf_store_nt(&r_out[x], gf_load(&F[y * 512 + 256 + io]));
    }
}

// The same with the right half times eq_hi[c >> 8] (s > 8). Workgroups are
// dispatched to the eight XCDs in turn, so block b runs on XCD b mod 8: it
// takes the rows y = b mod 8 (mod 8) only, so that each row's fold tables
// are read into one XCD's L2. Block b = xcd + 8 (g + runs l8) takes the
// 256-column run g (one window table of its factor) and the rows
// y = xcd + 8 (l8 + lanes8 k).
template <int L>
__global__ __launch_bounds__(kT) void k_nibble_fold_hi(const uint16_t* __restrict__ nr, const uint4* __restrict__ F,
                                                       int t, int s, const uint4* __restrict__ eq_hi,
                                                       uint32_t lanes8, uint4* __restrict__ l_out,
                                                       uint4* __restrict__ r_out) {
    __shared__ uint4 tab[512];
    const uint32_t xcd = blockIdx.x & 7u, li = blockIdx.x >> 3;
    const uint32_t g = li & ((1u << (s - 8)) - 1), l8 = li >> (s - 8);
    gf_table_build(tab, gf_load(&eq_hi[g]), threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t y4 = 1ull << (t - 4), c = ((uint64_t)g << 8) + threadIdx.x;
    for (uint64_t y = xcd + 8ull * l8; y < y4; y += 8ull * lanes8) {
        const uint64_t x = (y << s) + c;
        uint32_t ie, io;
        nibble_fold_idx<L>(nr[x], &ie, &io);
        gf_store_nt(&l_out[x], gf_load(&F[y * 512 + ie]));
        gf_store_nt(&r_out[x], gf_mul_tab(tab, gf_load(&F[y * 512 + 256 + io])));
    }
}

Chal chal_args(const hgf* host, int n, const void* dev) {
    Chal c{};
    for (int i = 0; i < n && host; i++) {
        c.lo[i] = host[i].lo;
        c.hi[i] = host[i].hi;
    }
    c.dev = (const uint4*)dev;
    return c;
}

int nibble_ttab(bitz_ctx* ctx, int level, const void* A, int t, const Chal& ch, uint4** T) {
    const uint64_t y4 = 1ull << (t - 4);
    if (!ctx->scratch.nib_t.alloc(y4 * 64 * 16)) return BITZ_ERR_ALLOC;
    *T = (uint4*)ctx->scratch.nib_t.ptr;
    const dim3 grid(blocks_for(y4 * 64)), block(kT);
    if (level == 0)
        hipLaunchKernelGGL(k_nibble_ttab<0>, grid, block, 0, ctx->stream, (const uint4*)A, t, ch, *T);
    else if (level == 1)
        hipLaunchKernelGGL(k_nibble_ttab<1>, grid, block, 0, ctx->stream, (const uint4*)A, t, ch, *T);
    else
        hipLaunchKernelGGL(k_nibble_ttab<2>, grid, block, 0, ctx->stream, (const uint4*)A, t, ch, *T);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

}  // namespace

int gkr_nibble_rows(bitz_ctx* ctx, const uint64_t* packed, int t, int s, void* nr) {
    BITZ_REQUIRE(t >= 4, BITZ_ERR_SHAPE, "gkr_nibble_rows: t < 4");
    const uint64_t n = (1ull << (t - 4)) << s;
    TimedLaunch tl(ctx, "nibble_rows", (double)(n * 2) + (double)((1ull << (t + s)) / 8), 0.0);
    hipLaunchKernelGGL(k_nibble_rows, dim3(blocks_for(n, 1, 8192)), dim3(kT), 0, ctx->stream, packed, t, s,
                       (uint16_t*)nr);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_nibble_round_sums(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s, const hgf* chal_host,
                          const void* chal_dev, const void* eq_c, const void* eq_y, int send_one, void* out_dev) {
    BITZ_REQUIRE(level >= 0 && level <= 2 && t >= 4, BITZ_ERR_SHAPE, "gkr_nibble_round_sums: bad level or shape");
    const Chal ch = chal_args(chal_host, 2 - level, chal_dev);
    const uint64_t y4 = 1ull << (t - 4), cols = 1ull << s;
    BITZ_REQUIRE(y4 <= 0x7fffffffull, BITZ_ERR_SHAPE, "gkr_nibble_round_sums: too many rows");
    const unsigned nb = (unsigned)y4;  // one block per row
    int rc = ctx->scratch.ensure((size_t)nb * 2 * 16);
    if (rc) return rc;
    {
        TimedLaunch tl(ctx, "nibble_round_sums", (double)(y4 * cols) * 2.0 + 16.0 * (double)cols, 0.0);
        const dim3 grid(nb), block(kT);
        const uint16_t* w = (const uint16_t*)nr;
        uint4* parts = (uint4*)ctx->scratch.partials.ptr;
        if (level == 0)
            hipLaunchKernelGGL(k_nibble_bucket_sums<0>, grid, block, 0, ctx->stream, w, (const uint4*)A, t, s, ch,
                               (const uint4*)eq_c, (const uint4*)eq_y, send_one, parts);
        else if (level == 1)
            hipLaunchKernelGGL(k_nibble_bucket_sums<1>, grid, block, 0, ctx->stream, w, (const uint4*)A, t, s, ch,
                               (const uint4*)eq_c, (const uint4*)eq_y, send_one, parts);
        else
            hipLaunchKernelGGL(k_nibble_bucket_sums<2>, grid, block, 0, ctx->stream, w, (const uint4*)A, t, s, ch,
                               (const uint4*)eq_c, (const uint4*)eq_y, send_one, parts);
        BITZ_LAUNCH_CHECK();
    }
    return sums_finish(ctx, nb, 2, out_dev);
}

int gkr_nibble_fold(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s, const hgf* chal_host,
                    const void* chal_dev, const void* eq_hi, void* l_out, void* r_out) {
    BITZ_REQUIRE(level >= 0 && level <= 2 && t >= 4, BITZ_ERR_SHAPE, "gkr_nibble_fold: bad level or shape");
    BITZ_REQUIRE(!eq_hi || s > 8, BITZ_ERR_SHAPE, "gkr_nibble_fold: eq_hi needs s > 8");
    const Chal ch = chal_args(chal_host, 3 - level, chal_dev);
    uint4* T;
    int rc = nibble_ttab(ctx, level, A, t, ch, &T);
    if (rc) return rc;
    const uint64_t y4 = 1ull << (t - 4), n = y4 << s;
    if (!ctx->scratch.nib_tab.alloc(y4 * 512 * 16)) return BITZ_ERR_ALLOC;
    uint4* F = (uint4*)ctx->scratch.nib_tab.ptr;
    {
        const dim3 grid(blocks_for(y4 * 512)), block(kT);
        if (level == 0)
            hipLaunchKernelGGL(k_nibble_ftab<0>, grid, block, 0, ctx->stream, T, t, ch, F);
        else if (level == 1)
            hipLaunchKernelGGL(k_nibble_ftab<1>, grid, block, 0, ctx->stream, T, t, ch, F);
        else
            hipLaunchKernelGGL(k_nibble_ftab<2>, grid, block, 0, ctx->stream, T, t, ch, F);
        BITZ_LAUNCH_CHECK();
    }
    TimedLaunch tl(ctx, "nibble_fold", 2.0 * (double)n + 32.0 * (double)n, 0.0);
    const uint16_t* w = (const uint16_t*)nr;
    if (eq_hi) {
        const uint32_t runs = 1u << (s - 8);
        const uint32_t lanes8 =
            (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(std::max<uint64_t>(y4 / 8, 1), std::max(1u, 1024u / runs)));
        const dim3 grid(8u * runs * lanes8), block(kT);
        if (level == 0)
            hipLaunchKernelGGL(k_nibble_fold_hi<0>, grid, block, 0, ctx->stream, w, F, t, s, (const uint4*)eq_hi, lanes8,
                               (uint4*)l_out, (uint4*)r_out);
        else if (level == 1)
            hipLaunchKernelGGL(k_nibble_fold_hi<1>, grid, block, 0, ctx->stream, w, F, t, s, (const uint4*)eq_hi, lanes8,
                               (uint4*)l_out, (uint4*)r_out);
        else
            hipLaunchKernelGGL(k_nibble_fold_hi<2>, grid, block, 0, ctx->stream, w, F, t, s, (const uint4*)eq_hi, lanes8,
                               (uint4*)l_out, (uint4*)r_out);
    } else {
        const dim3 grid(blocks_for(n, 1, 8192)), block(kT);
        if (level == 0)
            hipLaunchKernelGGL(k_nibble_fold<0>, grid, block, 0, ctx->stream, w, F, t, s, (uint4*)l_out, (uint4*)r_out);
        else if (level == 1)
            hipLaunchKernelGGL(k_nibble_fold<1>, grid, block, 0, ctx->stream, w, F, t, s, (uint4*)l_out, (uint4*)r_out);
        else
            hipLaunchKernelGGL(k_nibble_fold<2>, grid, block, 0, ctx->stream, w, F, t, s, (uint4*)l_out, (uint4*)r_out);
    }
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

}  // namespace bitz
