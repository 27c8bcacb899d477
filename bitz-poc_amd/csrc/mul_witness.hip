// Committed bit grid of the native multiplication relations: MulLayout
// (src/piop/spartan/mul.rs:62-153) and MulWitness::bitz_bit_rows
// (mul.rs:268-330, word packer src/piop/spartan/slot_rows.rs:223-270).
//
// Gate i of a batch of 2^g (padded) multiplications carries four limbs
// [x | y | lo | hi] of B bits (B = 32, 64, 128; lo, hi = the 2B-bit product
// x*y unless supplied). Bit b of limb l is slot k = l*B + b of 4B slots, and
// lands in column c = i mod 2^s at row bit (k << (g - s)) | (i >> s)
// (bitz_cell, mul.rs:140-150). Columns are 2^t bits, t = g - s + log2(4B),
// stored as 2^(t-6) u64 words, column c at rows + c * 2^(t-6).
//
// Per gate the 4B slots are NW = 4B/64 slot words (word w holds slots
// 64w..64w+63: u64 [x, y, lo, hi]; u32 [x | y<<32, lo | hi<<32]; u128 the
// eight 64-bit halves). With 2^(g-s) >= 64 high gates, output word
// (64w + k) * L + j of column c (L = 2^(g-s) / 64) holds bit k of slot word
// w of the 64 gates ((64 j + lane) << s) | c: one wavefront loads those 64
// gates (lane = gate), transposes each 64 x 64 bit block across the lanes
// (64 ballots) and lane k stores word k. Consecutive waves
// take consecutive columns, so the strided gate loads of one wave share cache
// lines with its neighbours'. With fewer high gates (a shifted split on a
// small batch) each thread gathers the 64 bits of one output word.
#include "../../include/bitz_hip.h"
#include "bits_dev.h"
#include "common.h"

namespace bitz {
namespace {

template <int B>
struct Limbs;

template <>
struct Limbs<32> {
    static constexpr int NW = 2;
    // slot words of gate i (zero for padding gates)
    __device__ __forceinline__ static void words(const void* x, const void* y, const void* lo, const void* hi,
                                                 uint64_t i, uint64_t n, uint64_t* w) {
        if (i >= n) {
            w[0] = w[1] = 0;
            return;
        }
        const uint32_t a = ((const uint32_t*)x)[i], b = ((const uint32_t*)y)[i];
        uint32_t zl, zh;
        if (lo) {
            zl = ((const uint32_t*)lo)[i];
            zh = ((const uint32_t*)hi)[i];
        } else {
            const uint64_t p = (uint64_t)a * b;
            zl = (uint32_t)p;
            zh = (uint32_t)(p >> 32);
        }
        w[0] = (uint64_t)a | ((uint64_t)b << 32);
        w[1] = (uint64_t)zl | ((uint64_t)zh << 32);
    }
};

template <>
struct Limbs<64> {
    static constexpr int NW = 4;
    __device__ __forceinline__ static void words(const void* x, const void* y, const void* lo, const void* hi,
                                                 uint64_t i, uint64_t n, uint64_t* w) {
        if (i >= n) {
            w[0] = w[1] = w[2] = w[3] = 0;
            return;
        }
        const uint64_t a = ((const uint64_t*)x)[i], b = ((const uint64_t*)y)[i];
        w[0] = a;
        w[1] = b;
        if (lo) {
            w[2] = ((const uint64_t*)lo)[i];
            w[3] = ((const uint64_t*)hi)[i];
        } else {
            w[2] = a * b;
            w[3] = __umul64hi(a, b);
        }
    }
};

template <>
struct Limbs<128> {
    static constexpr int NW = 8;
    __device__ __forceinline__ static void words(const void* x, const void* y, const void* lo, const void* hi,
                                                 uint64_t i, uint64_t n, uint64_t* w) {
        if (i >= n) {
            for (int k = 0; k < 8; k++) w[k] = 0;
            return;
        }
        const uint64_t* xp = (const uint64_t*)x + 2 * i;
        const uint64_t* yp = (const uint64_t*)y + 2 * i;
        const uint64_t a0 = xp[0], a1 = xp[1], b0 = yp[0], b1 = yp[1];
        w[0] = a0;
        w[1] = a1;
        w[2] = b0;
        w[3] = b1;
        if (lo) {
            const uint64_t* lp = (const uint64_t*)lo + 2 * i;
            const uint64_t* hp = (const uint64_t*)hi + 2 * i;
            w[4] = lp[0];
            w[5] = lp[1];
            w[6] = hp[0];
            w[7] = hp[1];
        } else {
            // 256-bit schoolbook product of two 128-bit values, 64-bit digits
            const uint64_t p00l = a0 * b0, p00h = __umul64hi(a0, b0);
            const uint64_t p01l = a0 * b1, p01h = __umul64hi(a0, b1);
            const uint64_t p10l = a1 * b0, p10h = __umul64hi(a1, b0);
            const uint64_t p11l = a1 * b1, p11h = __umul64hi(a1, b1);
            uint64_t d1 = p00h + p01l;
            uint64_t c1 = d1 < p00h;
            d1 += p10l;
            c1 += d1 < p10l;
            uint64_t d2 = p01h + p10h;
            uint64_t c2 = d2 < p01h;
            d2 += p11l;
            c2 += d2 < p11l;
            d2 += c1;
            c2 += d2 < c1;
            const uint64_t d3 = p11h + c2;
            w[4] = p00l;
            w[5] = d1;
            w[6] = d2;
            w[7] = d3;
        }
    }
};

template <int B>
__global__ __launch_bounds__(256) void k_mul_rows_transpose(const void* __restrict__ x, const void* __restrict__ y,
                                                            const void* __restrict__ lo,
                                                            const void* __restrict__ hi, uint64_t n, int s,
                                                            int lw_log, int cg_log, uint64_t waves,
                                                            uint64_t* __restrict__ rows, int t) {
    constexpr int NW = Limbs<B>::NW;
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wid >= waves) return;  // whole waves only (blockDim is a multiple of 64)
    // wave -> (column c, block j): 2^cg_log consecutive columns share j
    const uint64_t cg_mask = (1ull << cg_log) - 1, lw_mask = (1ull << lw_log) - 1;
    const uint64_t c = ((wid >> (cg_log + lw_log)) << cg_log) | (wid & cg_mask);
    const uint64_t j = (wid >> cg_log) & lw_mask;
    const uint64_t gate = (((j << 6) | (uint64_t)lane) << s) | c;
    uint64_t w[NW];
    Limbs<B>::words(x, y, lo, hi, gate, n, w);
    uint64_t* col = rows + (c << (t - 6));
#pragma unroll
    for (int k = 0; k < NW; k++) {
        const uint64_t o = wave_transpose64(w[k], lane);
        col[((((uint64_t)k << 6) | (uint64_t)lane) << lw_log) | j] = o;
    }
}

// high = 2^hl < 64 high gates: thread per output word
template <int B>
__global__ __launch_bounds__(256) void k_mul_rows_gather(const void* __restrict__ x, const void* __restrict__ y,
                                                         const void* __restrict__ lo, const void* __restrict__ hi,
                                                         uint64_t n, int s, int hl, uint64_t* __restrict__ rows,
                                                         int t) {
    constexpr int NW = Limbs<B>::NW;
    const uint64_t col_words = 1ull << (t - 6), total = col_words << s;
    for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
         o += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c = o >> (t - 6), wd = o & (col_words - 1);
        uint64_t out = 0;
        for (int b = 0; b < 64; b++) {
            const uint64_t bit = (wd << 6) | (uint64_t)b;
            const uint64_t slot = bit >> hl, ghi = bit & ((1ull << hl) - 1);
            uint64_t wv[NW];
            Limbs<B>::words(x, y, lo, hi, (ghi << s) | c, n, wv);
            uint64_t sw = 0;
#pragma unroll
            for (int k = 0; k < NW; k++) sw = (slot >> 6) == (uint64_t)k ? wv[k] : sw;
            out |= ((sw >> (slot & 63)) & 1ull) << b;
        }
        rows[o] = out;
    }
}

template <int B>
int launch_rows(bitz_ctx* ctx, const void* x, const void* y, const void* lo, const void* hi, uint64_t n, int g,
                int s, int t, uint64_t* rows) {
    const int hl = g - s;
    const double bytes = (double)n * (B / 8) * (lo ? 4 : 2) + (double)(1ull << (t + s - 3));
    TimedLaunch tl(ctx, "mul_bit_rows", bytes, 0.0);
    if (hl >= 6) {
        const int lw_log = hl - 6;
        const int cg_log = std::min(s, 4);
        const uint64_t waves = 1ull << (s + lw_log);
        const uint64_t blocks = (waves * 64 + 255) / 256;
        BITZ_REQUIRE(blocks <= 0x7FFFFFFFull, BITZ_ERR_SHAPE, "bitz_mul_bit_rows_dev: grid too large");
        hipLaunchKernelGGL(k_mul_rows_transpose<B>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, x, y, lo, hi,
                           n, s, lw_log, cg_log, waves, rows, t);
    } else {
        const uint64_t total = 1ull << (t + s - 6);
        const unsigned blocks = (unsigned)std::min<uint64_t>(65536, (total + 255) / 256);
        hipLaunchKernelGGL(k_mul_rows_gather<B>, dim3(blocks), dim3(256), 0, ctx->stream, x, y, lo, hi, n, s, hl,
                           rows, t);
    }
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int log2_exact(uint64_t v) { return 63 - __builtin_clzll(v); }

}  // namespace
}  // namespace bitz

// MulLayout::new / default_col_vars / with_split_shift / bitz_params
// (src/piop/spartan/mul.rs:71-139): capacity = max(n, 256) rounded up to a
// power of two (2^g), s = min(N - reference_log_rows(N), g - 6) + shift with
// N = g + log2(4B) (src/bitz/params.rs:30-35), t = g - s + log2(4B).
extern "C" int bitz_mul_layout(uint32_t word_bits, uint64_t multiplications, int32_t split_shift, uint32_t out[4]) {
    BITZ_REQUIRE(out != nullptr, BITZ_ERR_ARG, "bitz_mul_layout: null out");
    BITZ_REQUIRE(word_bits == 32 || word_bits == 64 || word_bits == 128, BITZ_ERR_ARG,
                 "bitz_mul_layout: word_bits must be 32, 64 or 128");
    BITZ_REQUIRE(multiplications > 0, BITZ_ERR_SHAPE, "bitz_mul_layout: a multiplication batch must not be empty");
    const uint64_t m = std::max<uint64_t>(multiplications, 256);
    BITZ_REQUIRE(m <= (1ull << 63), BITZ_ERR_SHAPE, "bitz_mul_layout: the multiplication domain is too large");
    int g = bitz::log2_exact(m);
    if ((1ull << g) < m) g++;
    const int slot_log = bitz::log2_exact(4ull * word_bits);
    BITZ_REQUIRE(g + slot_log < 64, BITZ_ERR_SHAPE, "bitz_mul_layout: the multiplication domain is too large");
    const int N = g + slot_log;
    // reference_log_rows: ceil(3N/5) - 1, at least 7 (the packing width), at most N - 1
    int rows = (3 * N + 4) / 5 - 1;
    rows = std::max(rows, 7);
    rows = std::min(rows, N - 1);
    const int def = std::min(std::max(N - rows, 0), std::max(g - 6, 0));
    const int64_t s = (int64_t)def + split_shift;
    BITZ_REQUIRE(s >= 0 && s <= g, BITZ_ERR_SHAPE, "bitz_mul_layout: split shift out of range");
    out[0] = (uint32_t)g;
    out[1] = (uint32_t)s;
    out[2] = (uint32_t)(g - s + slot_log);
    out[3] = (uint32_t)N;
    return BITZ_OK;
}

extern "C" int bitz_mul_bit_rows_dev(bitz_ctx* ctx, uint32_t word_bits, uint64_t multiplications,
                                     int32_t split_shift, const void* x, const void* y, const void* lo,
                                     const void* hi, void* rows, uint64_t rows_words) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "null context");
    BITZ_HIP(hipSetDevice(ctx->device));
    uint32_t lay[4];
    if (int rc = bitz_mul_layout(word_bits, multiplications, split_shift, lay)) return rc;
    BITZ_REQUIRE(x && y && rows, BITZ_ERR_ARG, "bitz_mul_bit_rows_dev: null argument");
    BITZ_REQUIRE((lo == nullptr) == (hi == nullptr), BITZ_ERR_ARG,
                 "bitz_mul_bit_rows_dev: supply both product limbs or neither");
    const int g = (int)lay[0], s = (int)lay[1], t = (int)lay[2];
    BITZ_REQUIRE(t + s <= 40, BITZ_ERR_SHAPE, "bitz_mul_bit_rows_dev: grid above 2^40 bits");
    BITZ_REQUIRE(rows_words == (1ull << (t + s - 6)), BITZ_ERR_SHAPE,
                 "bitz_mul_bit_rows_dev: packed rows do not match the committed layout");
    switch (word_bits) {
        case 32:
            return bitz::launch_rows<32>(ctx, x, y, lo, hi, multiplications, g, s, t, (uint64_t*)rows);
        case 64:
            return bitz::launch_rows<64>(ctx, x, y, lo, hi, multiplications, g, s, t, (uint64_t*)rows);
        default:
            return bitz::launch_rows<128>(ctx, x, y, lo, hi, multiplications, g, s, t, (uint64_t*)rows);
    }
}
