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
// F_p[i] = (1 + r_k) T_{0,p}[ext(i & Em)] + r_k T_{1,p}[ext((i & Om) >> sh)]:
// k_nibble_fold writes it out as halves of 2^(t+s-4) entries (the C ABI's
// per-stage entry point); the prover keeps the tables F instead and runs one
// more round off them (the table round below, which level 3 starts with),
// whose fold writes halves of 2^(t+s-5) entries where the dense rounds
// start.
#include "common.h"
#include "gf128_dev.h"
#include "open_impl.h"
#include "reduce_dev.h"
#include "bits_dev.h"

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
    uint64_t lo[4], hi[4];
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

// The same words by wavefront transposes (s >= 6, t >= 6): a wave takes
// column group g (64 columns) and rows y0..y0+3; lane L = 4w + yo loads the
// packed word of row y0 + yo + w Y4, and after the 64 x 64 bit transpose
// lane j holds bit(c, y0 + yo + w Y4) at bit 4w + yo for its column c =
// 64g + j. Each of its four words is then two 8-bit gathers (bits 8i and
// 8i + 4 of m >> yo: the even and the odd w) by one multiply each.
__device__ __forceinline__ uint32_t gather8(uint64_t x) {
    return (uint32_t)(((x & 0x0101010101010101ull) * 0x0102040810204080ull) >> 56);
}

__global__ __launch_bounds__(kT) void k_nibble_rows_wave(const uint64_t* __restrict__ packed, int t, int s,
                                                         uint16_t* __restrict__ nr) {
    const uint64_t y4 = 1ull << (t - 4), nrows = 1ull << t;
    const int ylog = t - 6;  // log2 of the row quads
    const uint64_t steps = 1ull << (ylog + s - 6);
    const int lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint32_t w = (uint32_t)lane >> 2, yo = (uint32_t)lane & 3u;
    for (uint64_t st = grid_tid() >> 6; st < steps; st += nwaves) {
        const uint64_t y0 = (st & ((1ull << ylog) - 1)) << 2, g = st >> ylog;
        const uint64_t m = wave_transpose64(packed[g * nrows + y0 + yo + w * y4], lane);
        uint16_t* out = nr + (y0 << s) + (g << 6) + lane;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint64_t x = m >> q;
            out[(uint64_t)q << s] = (uint16_t)(gather8(x) | (gather8(x >> 4) << 8));
        }
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

// ---- the table round and fold (Forest::jit_round_sums, jit_fold_round_with,
// src/bitz/forest.rs:990-1176) ----
// After its bit and row-word rounds (level 3: none), a low layer's halves at
// (y, c), y < Y4, are F_p[y][sel_p(w)], w the slot's row word, F the per-row
// 256-entry tables F[(y*2 + p)*256 + i]: for levels 0-2 the fold tables of
// the row-word round (k_nibble_ftab, indices ie / io), for level 3 the
// products of the eight leaves of the word's low (E) or high (O) byte. The
// next round binds y's top bit, so E_h, O_h (h = 0, 1) at (y', c),
// y' < Y5 = 2^(t-5), read the word of row y' + h Y5. Its sums bucket
// eq_c[c] O_b(c) by E's index per row (one LDS XOR per bucket, then 768
// products per row); its fold is two lookups per half in the tables
// pre-scaled by 1 + r (h = 0) and r (h = 1), into halves of 2^(t+s-5)
// entries where the layer's dense rounds start.
template <int L>
struct Sel {
    __device__ __forceinline__ static uint32_t e(uint32_t w) {
        if constexpr (L == 3) {
            return w & 255u;
        } else {
            uint32_t ie, io;
            nibble_fold_idx<L>(w, &ie, &io);
            return ie;
        }
    }
    __device__ __forceinline__ static uint32_t o(uint32_t w) {
        if constexpr (L == 3) {
            return w >> 8;
        } else {
            uint32_t ie, io;
            nibble_fold_idx<L>(w, &ie, &io);
            return io;
        }
    }
};

// Level 3's tables: F[(y*2 + p)*256 + i] = prod over the set bits u of i of
// A[y + Y4 (p + 2u)] (the rows of word bit w = p + 2u), as the product of
// the tables of its low and high nibble.
__global__ __launch_bounds__(kT) void k_level3_ftab(const uint4* __restrict__ A, int t, uint4* __restrict__ F) {
    __shared__ uint4 part[2][16];
    const uint64_t y4 = 1ull << (t - 4), y = blockIdx.x >> 1;
    const uint32_t p = blockIdx.x & 1u, i = threadIdx.x;
    if (i < 32) {
        const uint32_t half = i >> 4, j = i & 15u;
        gf128 acc = gf_one();
        bool first = true;
#pragma unroll
        for (uint32_t u4 = 0; u4 < 4; u4++) {
            if ((j >> u4) & 1u) {
                const uint32_t u = u4 + 4 * half;
                const gf128 a = gf_load(&A[y + y4 * (p + 2 * u)]);
                acc = first ? a : gf_mul(acc, a);
                first = false;
            }
        }
        gf_store(&part[half][j], acc);
    }
    __syncthreads();
    gf_store(&F[(uint64_t)blockIdx.x * 256 + i], gf_mul(gf_load(&part[0][i & 15u]), gf_load(&part[1][i >> 4])));
}

// Level 4 from level 3's tables: out[(y << s) | c] = F_0[y][lo(w)] F_1[y][hi(w)].
__global__ __launch_bounds__(kT) void k_level4_from_tab(const uint16_t* __restrict__ nr, const uint4* __restrict__ F,
                                                        int s, uint4* __restrict__ out) {
    __shared__ uint4 tab[512];
    const uint64_t y = blockIdx.x, cols = 1ull << s;
    for (uint32_t i = threadIdx.x; i < 512; i += kT) tab[i] = F[y * 512 + i];
    __syncthreads();
    for (uint64_t c = threadIdx.x; c < cols; c += kT) {
        const uint32_t w = nr[(y << s) | c];
        gf_store(&out[(y << s) | c], gf_mul(gf_load(&tab[w & 255u]), gf_load(&tab[256 + (w >> 8)])));
    }
}

// The table round's sums, block y' < Y5: buckets over E's index of
//   END[i]  = sum of eq_c[c] O_ss(c) over the columns with E_ss index i,
//   INFh[i] = sum of eq_c[c] (O_0(c) + O_1(c)) over those with E_h index i,
// then end = sum_i E_ss[i] END[i], inf = sum_h sum_i E_h[i] INFh[i]
// (= sum eq (E_0 + E_1)(O_0 + O_1)), the E tables pre-scaled by eq_y[y'].
template <int L>
__global__ __launch_bounds__(kT) void k_jit_round_sums(const uint16_t* __restrict__ nr, const uint4* __restrict__ F,
                                                       int t, int s, const uint4* __restrict__ eq_c,
                                                       const uint4* __restrict__ eq_y, int send_one,
                                                       uint4* __restrict__ partials) {
    __shared__ uint4 TE[2][256], TO[2][256];
    __shared__ uint32_t B[4][3][256];  // [word][END, INF0, INF1][index]
    const uint64_t y = blockIdx.x, y5 = 1ull << (t - 5), cols = 1ull << s;
    for (uint32_t i = threadIdx.x; i < 4 * 3 * 256; i += kT) (&B[0][0][0])[i] = 0u;
    {
        const gf128 wy = gf_load(&eq_y[y]);
        const uint32_t i = threadIdx.x;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint64_t yh = y + (uint64_t)h * y5;
            gf_store(&TE[h][i], gf_mul(gf_load(&F[(yh * 2 + 0) * 256 + i]), wy));
            TO[h][i] = F[(yh * 2 + 1) * 256 + i];
        }
    }
    __syncthreads();
    const uint16_t* w0p = nr + (y << s);
    const uint16_t* w1p = nr + ((y + y5) << s);
    for (uint64_t c = threadIdx.x; c < cols; c += kT) {
        const uint32_t w0 = w0p[c], w1 = w1p[c];
        const gf128 e = gf_load(&eq_c[c]);
        const gf128 u0 = gf_mul(e, gf_load(&TO[0][Sel<L>::o(w0)]));
        const gf128 u1 = gf_mul(e, gf_load(&TO[1][Sel<L>::o(w1)]));
        const uint32_t a0 = Sel<L>::e(w0), a1 = Sel<L>::e(w1);
        const uint32_t as = send_one ? a1 : a0;
        const gf128 us = send_one ? u1 : u0, uu = gf_add(u0, u1);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            atomicXor(&B[k][0][as], us.w[k]);
            atomicXor(&B[k][1][a0], uu.w[k]);
            atomicXor(&B[k][2][a1], uu.w[k]);
        }
    }
    __syncthreads();
    const uint32_t i = threadIdx.x;
    const gf128 bend = gf128{{B[0][0][i], B[1][0][i], B[2][0][i], B[3][0][i]}};
    const gf128 b0 = gf128{{B[0][1][i], B[1][1][i], B[2][1][i], B[3][1][i]}};
    const gf128 b1 = gf128{{B[0][2][i], B[1][2][i], B[2][2][i], B[3][2][i]}};
    gf256 inf = gf_mul_wide(gf_load(&TE[0][i]), b0);
    gf256_acc(&inf, gf_mul_wide(gf_load(&TE[1][i]), b1));
    gf128 sums[2] = {gf_mul(gf_load(&TE[send_one ? 1 : 0][i]), bend), gf_reduce(inf)};
    block_xor_store<2>(sums, partials + (size_t)blockIdx.x * 2);
}

// F rows y < Y5 times (1 + r), rows Y5 <= y < Y4 times r (r = challenge idx)
__global__ __launch_bounds__(kT) void k_jit_prescale(uint4* __restrict__ F, int t, Chal ch, int idx) {
    __shared__ uint4 tab0[512], tab1[512];
    const gf128 r = chal(ch, idx);
    gf_table_build(tab0, gf_add(r, gf_one()), threadIdx.x, blockDim.x);
    gf_table_build(tab1, r, threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t y5 = 1ull << (t - 5), n = (1ull << (t - 4)) * 512;
    for (uint64_t e = grid_tid(); e < n; e += grid_threads())
        gf_store(&F[e], gf_mul_tab((e >> 9) < y5 ? tab0 : tab1, gf_load(&F[e])));
}

// The folded halves from the pre-scaled tables, x = (y' << s) | c:
//   l[x] = F_0[y'][e(w0)] + F_0[y' + Y5][e(w1)], r[x] likewise with F_1, o().
template <int L>
__device__ __forceinline__ void jit_fold_pair(const uint16_t* __restrict__ nr, const uint4* __restrict__ F, int s,
                                              uint64_t y5, uint64_t y, uint64_t c, gf128* lv, gf128* rv) {
    const uint32_t w0 = nr[(y << s) | c], w1 = nr[((y + y5) << s) | c];
    const uint64_t f0 = y * 512, f1 = (y + y5) * 512;
    *lv = gf_add(gf_load(&F[f0 + Sel<L>::e(w0)]), gf_load(&F[f1 + Sel<L>::e(w1)]));
    *rv = gf_add(gf_load(&F[f0 + 256 + Sel<L>::o(w0)]), gf_load(&F[f1 + 256 + Sel<L>::o(w1)]));
}

template <int L>
__global__ __launch_bounds__(kT) void k_jit_fold(const uint16_t* __restrict__ nr, const uint4* __restrict__ F, int t,
                                                 int s, uint4* __restrict__ l_out, uint4* __restrict__ r_out) {
    const uint64_t y5 = 1ull << (t - 5), n = y5 << s;
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) {
        gf128 lv, rv;
        jit_fold_pair<L>(nr, F, s, y5, x >> s, x & ((1ull << s) - 1), &lv, &rv);
        gf_store_nt(&l_out[x], lv);
        gf_store_nt(&r_out[x], rv);
    }
}

// The same with the right half times eq_hi[c >> 8] (s > 8), blocks mapped
// to rows as in k_nibble_fold_hi (XCD-aware: block b takes rows y' = b mod 8).
template <int L>
__global__ __launch_bounds__(kT) void k_jit_fold_hi(const uint16_t* __restrict__ nr, const uint4* __restrict__ F,
                                                    int t, int s, const uint4* __restrict__ eq_hi, uint32_t lanes8,
                                                    uint4* __restrict__ l_out, uint4* __restrict__ r_out) {
    __shared__ uint4 tab[512];
    const uint32_t xcd = blockIdx.x & 7u, li = blockIdx.x >> 3;
    const uint32_t g = li & ((1u << (s - 8)) - 1), l8 = li >> (s - 8);
    gf_table_build(tab, gf_load(&eq_hi[g]), threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t y5 = 1ull << (t - 5), c = ((uint64_t)g << 8) + threadIdx.x;
    for (uint64_t y = xcd + 8ull * l8; y < y5; y += 8ull * lanes8) {
        gf128 lv, rv;
        jit_fold_pair<L>(nr, F, s, y5, y, c, &lv, &rv);
        const uint64_t x = (y << s) + c;
        gf_store_nt(&l_out[x], lv);
        gf_store_nt(&r_out[x], gf_mul_tab(tab, rv));
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

// F[(y*2 + p)*256 + i] (ctx->scratch.nib_tab): the row-word round's fold tables
// of level 0-2 by challenge 2 - level (k_nibble_ftab), or level 3's leaf
// product tables (k_level3_ftab).
int nibble_ftab(bitz_ctx* ctx, int level, const void* A, int t, const Chal& ch, uint4** F) {
    const uint64_t y4 = 1ull << (t - 4);
    if (!ctx->scratch.nib_tab.alloc(y4 * 512 * 16)) return BITZ_ERR_ALLOC;
    *F = (uint4*)ctx->scratch.nib_tab.ptr;
    if (level == 3) {
        BITZ_REQUIRE(2 * y4 <= 0x7fffffffull, BITZ_ERR_SHAPE, "level-3 tables: too many rows");
        hipLaunchKernelGGL(k_level3_ftab, dim3((unsigned)(2 * y4)), dim3(kT), 0, ctx->stream, (const uint4*)A, t, *F);
        BITZ_LAUNCH_CHECK();
        return BITZ_OK;
    }
    uint4* T;
    int rc = nibble_ttab(ctx, level, A, t, ch, &T);
    if (rc) return rc;
    const dim3 grid(blocks_for(y4 * 512)), block(kT);
    if (level == 0)
        hipLaunchKernelGGL(k_nibble_ftab<0>, grid, block, 0, ctx->stream, T, t, ch, *F);
    else if (level == 1)
        hipLaunchKernelGGL(k_nibble_ftab<1>, grid, block, 0, ctx->stream, T, t, ch, *F);
    else
        hipLaunchKernelGGL(k_nibble_ftab<2>, grid, block, 0, ctx->stream, T, t, ch, *F);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

}  // namespace

int gkr_nibble_rows(bitz_ctx* ctx, const uint64_t* packed, int t, int s, void* nr) {
    BITZ_REQUIRE(t >= 4, BITZ_ERR_SHAPE, "gkr_nibble_rows: t < 4");
    const uint64_t n = (1ull << (t - 4)) << s;
    TimedLaunch tl(ctx, "nibble_rows", (double)(n * 2) + (double)((1ull << (t + s)) / 8), 0.0);
    if (s >= 6 && t >= 6)
        hipLaunchKernelGGL(k_nibble_rows_wave, dim3(blocks_for(n / 64, 1, 8192)), dim3(kT), 0, ctx->stream, packed,
                           t, s, (uint16_t*)nr);
    else
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
    uint4* F;
    int rc = nibble_ftab(ctx, level, A, t, ch, &F);
    if (rc) return rc;
    const uint64_t y4 = 1ull << (t - 4), n = y4 << s;
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

int gkr_jit_tables(bitz_ctx* ctx, int level, const void* A, int t, const hgf* chal_host, const void* chal_dev) {
    BITZ_REQUIRE(level >= 0 && level <= 3 && t >= 5, BITZ_ERR_SHAPE, "gkr_jit_tables: bad level or shape");
    const Chal ch = chal_args(chal_host, 3 - level, chal_dev);
    uint4* F;
    return nibble_ftab(ctx, level, A, t, ch, &F);
}

int gkr_jit_round_sums(bitz_ctx* ctx, int level, const void* nr, int t, int s, const void* eq_c, const void* eq_y,
                       int send_one, void* out_dev) {
    BITZ_REQUIRE(level >= 0 && level <= 3 && t >= 5, BITZ_ERR_SHAPE, "gkr_jit_round_sums: bad level or shape");
    const uint64_t y4 = 1ull << (t - 4), y5 = y4 >> 1, cols = 1ull << s;
    BITZ_REQUIRE(ctx->scratch.nib_tab.ptr && ctx->scratch.nib_tab.bytes >= y4 * 512 * 16, BITZ_ERR_ARG,
                 "gkr_jit_round_sums: tables not built");
    BITZ_REQUIRE(y5 <= 0x7fffffffull, BITZ_ERR_SHAPE, "gkr_jit_round_sums: too many rows");
    const unsigned nb = (unsigned)y5;  // one block per row
    int rc = ctx->scratch.ensure((size_t)nb * 2 * 16);
    if (rc) return rc;
    {
        TimedLaunch tl(ctx, "jit_round_sums", (double)(y5 * cols) * (4.0 + 16.0) + 16.0 * 1024.0 * (double)y5,
                       (double)(y5 * cols) * 2.0 * kGfMulOps);
        const dim3 grid(nb), block(kT);
        const uint16_t* w = (const uint16_t*)nr;
        const uint4* F = (const uint4*)ctx->scratch.nib_tab.ptr;
        uint4* parts = (uint4*)ctx->scratch.partials.ptr;
        const uint4* ec = (const uint4*)eq_c;
        const uint4* ey = (const uint4*)eq_y;
        if (level == 0)
            hipLaunchKernelGGL(k_jit_round_sums<0>, grid, block, 0, ctx->stream, w, F, t, s, ec, ey, send_one, parts);
        else if (level == 1)
            hipLaunchKernelGGL(k_jit_round_sums<1>, grid, block, 0, ctx->stream, w, F, t, s, ec, ey, send_one, parts);
        else if (level == 2)
            hipLaunchKernelGGL(k_jit_round_sums<2>, grid, block, 0, ctx->stream, w, F, t, s, ec, ey, send_one, parts);
        else
            hipLaunchKernelGGL(k_jit_round_sums<3>, grid, block, 0, ctx->stream, w, F, t, s, ec, ey, send_one, parts);
        BITZ_LAUNCH_CHECK();
    }
    return sums_finish(ctx, nb, 2, out_dev);
}

int gkr_jit_fold(bitz_ctx* ctx, int level, const void* nr, int t, int s, const hgf* chal_host, const void* chal_dev,
                 const void* eq_hi, void* l_out, void* r_out) {
    BITZ_REQUIRE(level >= 0 && level <= 3 && t >= 5, BITZ_ERR_SHAPE, "gkr_jit_fold: bad level or shape");
    BITZ_REQUIRE(!eq_hi || s > 8, BITZ_ERR_SHAPE, "gkr_jit_fold: eq_hi needs s > 8");
    const uint64_t y4 = 1ull << (t - 4), y5 = y4 >> 1, n = y5 << s;
    BITZ_REQUIRE(ctx->scratch.nib_tab.ptr && ctx->scratch.nib_tab.bytes >= y4 * 512 * 16, BITZ_ERR_ARG,
                 "gkr_jit_fold: tables not built");
    const Chal ch = chal_args(chal_host, 4 - level, chal_dev);
    uint4* F = (uint4*)ctx->scratch.nib_tab.ptr;
    hipLaunchKernelGGL(k_jit_prescale, dim3(blocks_for(y4 * 512, 4, 1024)), dim3(kT), 0, ctx->stream, F, t, ch,
                       3 - level);
    BITZ_LAUNCH_CHECK();
    TimedLaunch tl(ctx, "jit_fold", 4.0 * (double)n + 32.0 * (double)n, eq_hi ? (double)n * kGfTabMulOps : 0.0);
    const uint16_t* w = (const uint16_t*)nr;
    if (eq_hi) {
        const uint32_t runs = 1u << (s - 8);
        const uint32_t lanes8 = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(std::max<uint64_t>(y5 / 8, 1), std::max(1u, 1024u / runs)));
        const dim3 grid(8u * runs * lanes8), block(kT);
        const uint4* eh = (const uint4*)eq_hi;
        if (level == 0)
            hipLaunchKernelGGL(k_jit_fold_hi<0>, grid, block, 0, ctx->stream, w, F, t, s, eh, lanes8, (uint4*)l_out,
                               (uint4*)r_out);
        else if (level == 1)
            hipLaunchKernelGGL(k_jit_fold_hi<1>, grid, block, 0, ctx->stream, w, F, t, s, eh, lanes8, (uint4*)l_out,
                               (uint4*)r_out);
        else if (level == 2)
            hipLaunchKernelGGL(k_jit_fold_hi<2>, grid, block, 0, ctx->stream, w, F, t, s, eh, lanes8, (uint4*)l_out,
                               (uint4*)r_out);
        else
            hipLaunchKernelGGL(k_jit_fold_hi<3>, grid, block, 0, ctx->stream, w, F, t, s, eh, lanes8, (uint4*)l_out,
                               (uint4*)r_out);
    } else {
        const dim3 grid(blocks_for(n, 1, 8192)), block(kT);
        if (level == 0)
            hipLaunchKernelGGL(k_jit_fold<0>, grid, block, 0, ctx->stream, w, F, t, s, (uint4*)l_out, (uint4*)r_out);
        else if (level == 1)
            hipLaunchKernelGGL(k_jit_fold<1>, grid, block, 0, ctx->stream, w, F, t, s, (uint4*)l_out, (uint4*)r_out);
        else if (level == 2)
            hipLaunchKernelGGL(k_jit_fold<2>, grid, block, 0, ctx->stream, w, F, t, s, (uint4*)l_out, (uint4*)r_out);
        else
            hipLaunchKernelGGL(k_jit_fold<3>, grid, block, 0, ctx->stream, w, F, t, s, (uint4*)l_out, (uint4*)r_out);
    }
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_level4(bitz_ctx* ctx, const void* nr, const void* A, int t, int s, void* out) {
    BITZ_REQUIRE(t >= 5, BITZ_ERR_SHAPE, "gkr_level4: t < 5");
    const uint64_t y4 = 1ull << (t - 4), n = y4 << s;
    BITZ_REQUIRE(y4 <= 0x7fffffffull, BITZ_ERR_SHAPE, "gkr_level4: too many rows");
    const Chal ch{};
    uint4* F;
    int rc = nibble_ftab(ctx, 3, A, t, ch, &F);
    if (rc) return rc;
    TimedLaunch tl(ctx, "level4", 2.0 * (double)n + 16.0 * (double)n, (double)n * kGfMulOps);
    hipLaunchKernelGGL(k_level4_from_tab, dim3((unsigned)y4), dim3(kT), 0, ctx->stream, (const uint16_t*)nr, F, s,
                       (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

}  // namespace bitz
