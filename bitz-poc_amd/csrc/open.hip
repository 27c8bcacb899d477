// Opening-side kernels of the BitZ PCS: eq tables, integer column folds and
// their images, the grand-product GKR rounds, the inner-product sumcheck,
// the ring switch, the Ligerito sumcheck, the induced basis and the
// proof-of-work scan.
//
// Data layouts in HBM:
//  * Gf128 arrays: 16-byte elements (uint4), reference wire form.
//  * committed bits, column-major ("rows"): column c is 2^(t-6) u64 words,
//    word w bit j = row 64w+j (the packed message is these words as Gf128).
//  * committed bits, row-major ("packed columns"): packed[g*2^t + b] = bit b
//    of columns 64g..64g+63 (src/ligerito.rs:655-690 pack_columns_from_rows).
#include <atomic>
#include <thread>

#include "common.h"
#include "gf128_dev.h"
#include "blake3_dev.h"
#include "reduce_dev.h"
#include "open_impl.h"
#include "transcript_dev.h"
#include "bits_dev.h"

namespace bitz {
namespace {

constexpr int kT = 256;

__device__ __forceinline__ uint64_t grid_tid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
// A challenge argument: read from device memory when the transcript runs on
// the device (p != nullptr), else passed as two u64 scalars.
__device__ __forceinline__ gf128 challenge_arg(uint64_t lo, uint64_t hi, const uint4* p) {
    return p ? gf_load(p) : gf_from_u64(lo, hi);
}
__device__ __forceinline__ uint64_t grid_threads() { return (uint64_t)gridDim.x * blockDim.x; }

unsigned blocks_for(uint64_t n, uint64_t per_thread = 1, unsigned cap = 4096) {
    uint64_t threads = (n + per_thread - 1) / per_thread;
    uint64_t b = (threads + kT - 1) / kT;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// ---------------------------------------------------------------- eq tables
// Doubling in one block: out[x] for x < 2^n (n <= 14), index bit i <-> r[i]
// (src/poly/utils.rs build_eq_x_r_vec).
__global__ void k_eq_small(const uint4* __restrict__ r, int n, uint4* __restrict__ out) {
    if (threadIdx.x == 0) out[0] = make_uint4(1, 0, 0, 0);
    __syncthreads();
    for (int i = 0; i < n; i++) {
        gf128 ri = gf_load(&r[i]);
        uint32_t len = 1u << i;
        for (uint32_t x = threadIdx.x; x < len; x += blockDim.x) {
            gf128 p = gf_load(&out[x]);
            gf128 hi = gf_mul(p, ri);
            gf_store(&out[x + len], hi);
            gf_store(&out[x], gf_add(p, hi));
        }
        __syncthreads();
    }
}

// out[x] = lo[x & (2^nlo - 1)] * hi[x >> nlo] for runs of at least 2^10
// entries per hi factor: block b takes entries b 2^pbl .. (b + 1) 2^pbl - 1
// (pbl <= nlo, so one hi factor), whose product is a window-table product.
__global__ __launch_bounds__(kT) void k_eq_combine_tab(const uint4* __restrict__ lo, int nlo,
                                                       const uint4* __restrict__ hi, int pbl,
                                                       uint4* __restrict__ out) {
    __shared__ uint4 tab[512];
    const uint64_t base = (uint64_t)blockIdx.x << pbl, lmask = (1ull << nlo) - 1;
    gf_table_build(tab, gf_load(&hi[base >> nlo]), threadIdx.x, blockDim.x);
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < (1ull << pbl); i += kT) {
        const uint64_t x = base + i;
        gf_store(&out[x], gf_mul_tab(tab, gf_load(&lo[x & lmask])));
    }
}

// out[x] = lo[x & (2^nlo - 1)] * hi[x >> nlo]
__global__ void k_eq_combine(const uint4* __restrict__ lo, int nlo, const uint4* __restrict__ hi,
                             uint64_t total, uint4* __restrict__ out) {
    for (uint64_t x = grid_tid(); x < total; x += grid_threads()) {
        gf_store(&out[x], gf_mul(gf_load(&lo[x & ((1ull << nlo) - 1)]), gf_load(&hi[x >> nlo])));
    }
}

// Coordinate sources of k_tensor: by value (a kernel argument, no upload) or
// in device memory.
struct ArgCoords {
    TensorArgs c;
    __device__ __forceinline__ uint4 a(int i) const { return c.a[i]; }
    __device__ __forceinline__ uint4 b(int i) const { return c.b[i]; }
};
struct DevCoords {
    const uint4* pa;
    const uint4* pb;
    __device__ __forceinline__ uint4 a(int i) const { return pa[i]; }
    __device__ __forceinline__ uint4 b(int i) const { return pb[i]; }
};

// prod_{i < cnt} (bit i of x ? b(base + i) : a(base + i)) for cnt <= 4, as
// a two-level tree (three products at most)
template <class C>
__device__ __forceinline__ gf128 tensor_prod4(const C& c, int base, int cnt, uint32_t x) {
    gf128 f[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (i < cnt) {
            const uint4 v = ((x >> i) & 1u) ? c.b(base + i) : c.a(base + i);
            f[i] = gf_load(&v);
        } else {
            f[i] = gf_one();
        }
    }
    gf128 p = cnt > 1 ? gf_mul(f[0], f[1]) : f[0];
    if (cnt > 2) p = gf_mul(p, cnt > 3 ? gf_mul(f[2], f[3]) : f[2]);
    return p;
}

// out[x] = prod_i (bit i of x ? b(i) : a(i)) for x < 2^n. Block k writes
// entries 256k..256k+255 (the low 8 coordinates; the high ones are fixed by
// k). Three waves work side by side: wave 0 the table A of coordinates 0-3,
// wave 1 the table B of coordinates 4-7, wave 2 the product H of the high
// coordinates' factors by a shuffle tree; then entry x = A[x & 15] B[x >> 4]
// H. The longest chain is five products.
template <class C>
__device__ __forceinline__ void tensor_block(const C& c, int n, uint32_t blk, uint4* __restrict__ out) {
    __shared__ uint4 ta[16], tb[16], th[1];
    const uint32_t t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int nlo = n < 8 ? n : 8, la = nlo < 4 ? nlo : 4, lb = nlo - la, nh = n - nlo;
    if (wave == 0) {
        if (lane < (1u << la)) gf_store(&ta[lane], tensor_prod4(c, 0, la, lane));
    } else if (wave == 1) {
        if (lane < (1u << lb)) gf_store(&tb[lane], tensor_prod4(c, la, lb, lane));
    } else if (wave == 2) {
        gf128 v = gf_one();
        if ((int)lane < nh) {
            const uint4 f = ((blk >> lane) & 1u) ? c.b(nlo + lane) : c.a(nlo + lane);
            v = gf_load(&f);
        }
        for (int off = 1; off < nh; off <<= 1) {  // lanes past nh hold one
            gf128 o;
#pragma unroll
            for (int w = 0; w < 4; w++) o.w[w] = __shfl_down(v.w[w], off, 64);
            if ((int)(lane + off) >= 64) o = gf_one();
            v = gf_mul(v, o);
        }
        if (lane == 0) gf_store(&th[0], v);
    }
    __syncthreads();
    if (t < (1u << nlo)) {
        gf128 v = gf_load(&ta[t & ((1u << la) - 1)]);
        if (lb > 0) v = gf_mul(v, gf_load(&tb[t >> la]));
        if (nh > 0) v = gf_mul(v, gf_load(&th[0]));
        gf_store(&out[((uint64_t)blk << nlo) + t], v);
    }
}

template <class C>
__global__ __launch_bounds__(256) void k_tensor(C c, int n, uint4* __restrict__ out) {
    tensor_block(c, n, blockIdx.x, out);
}

// Several tables in one launch: job j owns blocks first[j]..first[j+1]-1.
struct TensorJobs {
    uint4 a[kTensorJobs][kTensorJobN];
    uint4 b[kTensorJobs][kTensorJobN];
    uint4* out[kTensorJobs];
    int n[kTensorJobs];
    uint32_t first[kTensorJobs + 1];
    int njobs;
};
struct JobCoords {
    const uint4* pa;
    const uint4* pb;
    __device__ __forceinline__ uint4 a(int i) const { return pa[i]; }
    __device__ __forceinline__ uint4 b(int i) const { return pb[i]; }
};
__global__ __launch_bounds__(256) void k_tensor_jobs(TensorJobs J) {
    int j = 0;
    while (j + 1 < J.njobs && blockIdx.x >= J.first[j + 1]) j++;
    tensor_block(JobCoords{J.a[j], J.b[j]}, J.n[j], blockIdx.x - J.first[j], J.out[j]);
}

// ------------------------------------------------------------- column folds
// eta_c = sum_b w_b bit(c, b) as wrapping u128 (src/bitz/fold.rs:58-170).
// Grid: (column, word-chunk); partial sums per chunk, then a second pass.
__device__ __forceinline__ void add128(uint64_t& lo, uint64_t& hi, uint64_t blo, uint64_t bhi) {
    uint64_t n = lo + blo;
    hi += bhi + (n < lo ? 1 : 0);
    lo = n;
}

__global__ void k_fold_columns_partial(const uint64_t* __restrict__ rows, int t, int s,
                                       const uint4* __restrict__ w, uint32_t chunk_words,
                                       uint4* __restrict__ partial /* [chunks][2^s] */) {
    const uint64_t words = 1ull << (t - 6);
    const uint64_t cols = 1ull << s;
    const uint64_t chunks = words / chunk_words;
    uint64_t idx = grid_tid();
    if (idx >= cols * chunks) return;
    uint64_t c = idx % cols, ch = idx / cols;
    uint64_t lo = 0, hi = 0;
    for (uint32_t k = 0; k < chunk_words; k++) {
        uint64_t wi = ch * chunk_words + k;
        uint64_t x = rows[c * words + wi];
        while (x) {
            int b = __builtin_ctzll(x);
            uint4 e = w[wi * 64 + b];
            add128(lo, hi, (uint64_t)e.x | ((uint64_t)e.y << 32), (uint64_t)e.z | ((uint64_t)e.w << 32));
            x &= x - 1;
        }
    }
    partial[ch * cols + c] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}

// The same with the chunk's 4096 row weights staged in LDS (64 KiB) so the
// per-set-bit gathers hit LDS: block = 256 consecutive columns of one chunk
// of 64 words (needs 2^s a multiple of 256 and 64-word chunks).
// The same partial sums through nibble tables (four Russians): block (chunk
// of 16 words = 1024 rows, 256 columns) builds T[k][n][v] = sum of the
// weights of rows 64k + 4n + j over the set bits j of v (16 words x 16
// nibbles x 16 entries, 64 KiB of LDS, one 128-bit add per entry), then each
// column adds one entry per nibble of its words: 16 lookups per word, no
// branch on the bits.
__global__ __launch_bounds__(256) void k_fold_columns_partial_lds(const uint64_t* __restrict__ rows, int t, int s,
                                                                  const uint4* __restrict__ w,
                                                                  uint4* __restrict__ partial) {
    __shared__ uint4 tab[16 * 16 * 16];
    const uint64_t words = 1ull << (t - 6), cols = 1ull << s;
    const uint64_t blocks_per_chunk = cols / 256;
    const uint64_t ch = blockIdx.x / blocks_per_chunk;
    const uint64_t c = (blockIdx.x % blocks_per_chunk) * 256 + threadIdx.x;
    {
        // thread (k, n) = threadIdx.x builds the 16 sums of rows 64k + 4n + j
        const uint4* wn = w + ch * 1024 + 4 * threadIdx.x;
        uint4* tn = tab + 16 * threadIdx.x;
        uint64_t wl[4], wh[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint4 e = wn[j];
            wl[j] = (uint64_t)e.x | ((uint64_t)e.y << 32);
            wh[j] = (uint64_t)e.z | ((uint64_t)e.w << 32);
        }
        uint64_t sl[16], sh[16];
        sl[0] = sh[0] = 0;
#pragma unroll
        for (int v = 1; v < 16; v++) {
            const int j = __builtin_ctz(v), u = v & (v - 1);
            sl[v] = sl[u];
            sh[v] = sh[u];
            add128(sl[v], sh[v], wl[j], wh[j]);
        }
#pragma unroll
        for (int v = 0; v < 16; v++)
            tn[v] = make_uint4((uint32_t)sl[v], (uint32_t)(sl[v] >> 32), (uint32_t)sh[v], (uint32_t)(sh[v] >> 32));
    }
    __syncthreads();
    uint64_t lo = 0, hi = 0;
    const uint64_t* col = rows + c * words + ch * 16;
    for (uint32_t k = 0; k < 16; k++) {
        const uint64_t x = col[k];
        const uint4* tk = tab + 256 * k;
#pragma unroll
        for (int n = 0; n < 16; n++) {
            const uint4 e = tk[16 * n + (uint32_t)((x >> (4 * n)) & 15u)];
            add128(lo, hi, (uint64_t)e.x | ((uint64_t)e.y << 32), (uint64_t)e.z | ((uint64_t)e.w << 32));
        }
    }
    partial[ch * cols + c] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}

__global__ void k_fold_columns_sum(const uint4* __restrict__ partial, uint64_t chunks, uint64_t cols,
                                   uint4* __restrict__ out) {
    uint64_t c = grid_tid();
    if (c >= cols) return;
    uint64_t lo = 0, hi = 0;
    for (uint64_t ch = 0; ch < chunks; ch++) {
        uint4 e = partial[ch * cols + c];
        add128(lo, hi, (uint64_t)e.x | ((uint64_t)e.y << 32), (uint64_t)e.z | ((uint64_t)e.w << 32));
    }
    out[c] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}

// ------------------------------------------------------- fixed-base powers
// comb[i*256 + v] = g^(v * 2^(8i)); g^e = prod_i comb[i][byte_i(e)]
// (vendor/field/src/preparation.rs:177-263: a 16 x 256 window table).
__global__ void k_comb_rows(uint4* __restrict__ comb, uint64_t g_lo, uint64_t g_hi) {
    // row i base: g^(2^(8i)) computed by thread i by repeated squaring
    int i = threadIdx.x;
    if (i >= 16) return;
    gf128 b = gf_from_u64(g_lo, g_hi);
    for (int k = 0; k < 8 * i; k++) b = gf_square(b);
    // powers b^v, v in 0..255, by a running product
    gf128 acc = gf_one();
    for (int v = 0; v < 256; v++) {
        gf_store(&comb[i * 256 + v], acc);
        acc = gf_mul(acc, b);
    }
}

__global__ void k_pow_comb(const uint4* __restrict__ comb, const uint4* __restrict__ exps, uint64_t n,
                           uint4* __restrict__ out) {
    uint64_t j = grid_tid();
    if (j >= n) return;
    uint4 e = exps[j];
    uint32_t ew[4] = {e.x, e.y, e.z, e.w};
    gf128 f[16];
#pragma unroll
    for (int i = 0; i < 16; i++) f[i] = gf_load(&comb[i * 256 + ((ew[i >> 2] >> (8 * (i & 3))) & 255u)]);
    // balanced product tree
#pragma unroll
    for (int w = 1; w < 16; w <<= 1)
#pragma unroll
        for (int i = 0; i < 16; i += 2 * w) f[i] = gf_mul(f[i], f[i + w]);
    gf_store(&out[j], f[0]);
}

// -------------------------------------------------- bit layout transposes
// packed[g * 2^t + b] = bit b of columns 64g..64g+63, from column-major rows.
// One wavefront per (group g, row word w): lane j loads word w of column
// 64g + j, the wave transposes the 64 x 64 block and lane k stores the bits
// of row 64w + k (a coalesced 512-byte store). Consecutive waves take
// consecutive w, so the column loads of 16 waves share each cache line.
__global__ __launch_bounds__(256) void k_pack_columns(const uint64_t* __restrict__ rows, int t, int s,
                                                      uint64_t* __restrict__ packed) {
    const uint64_t words = 1ull << (t - 6), cols = 1ull << s;
    const uint64_t groups = (cols + 63) >> 6;
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wid >= groups * words) return;  // whole waves (blockDim is a multiple of 64)
    const uint64_t g = wid >> (t - 6), w = wid & (words - 1);
    const uint64_t c = (g << 6) + (uint64_t)lane;
    const uint64_t v = c < cols ? rows[c * words + w] : 0;
    packed[(g << t) + (w << 6) + (uint64_t)lane] = wave_transpose64(v, lane);
}

// ------------------------------------------------------------- GKR rounds
// Weight of entry x of a round's half: eq_y[x >> s_c] * eq_c[x & (2^s_c - 1)].
// A block takes chunks of kT * R consecutive slots, thread i the slots
// base + i + kT k (k < R): loads coalesce across the wave. When R > 1 a chunk
// lies inside one row of eq_c (kT * R divides 2^s_c), so the thread adds its
// R weighted products unreduced and multiplies by eq_y once.
struct RoundVals {
    gf128 el, eh, ol, oh;
};

// ColW = false: the right operand already carries eq_c (it was multiplied
// in once per layer), only eq_y weighs the chunk.
template <class Src, bool ColW = true>
__global__ __launch_bounds__(kT) void k_gkr_round_sums(Src src, uint64_t h, const uint4* __restrict__ eq_c,
                                                       int s_c, const uint4* __restrict__ eq_y, int send_one,
                                                       uint32_t R, uint4* __restrict__ partials,
                                                       const uint4* __restrict__ eq_lo, uint32_t lo_mask) {
    __shared__ uint4 tab[Src::kTable ? 512 : 1];
    if constexpr (Src::kTable) {
        src.build(tab);
        __syncthreads();
    }
    gf128 acc[2] = {gf_zero(), gf_zero()};
    const uint64_t chunk = (uint64_t)kT * R;
    const uint64_t nchunks = (h + chunk - 1) / chunk;
    const uint64_t cmask = (1ull << s_c) - 1;
    // column weights of the low index bits, left out of the right half:
    // chunks start at multiples of kT, so every slot x of this thread has
    // x & (kT - 1) = threadIdx.x and one factor weighs all of them, taken
    // into each chunk's row weight (Src::kLoPerChunk: one product per chunk)
    // or into the thread's two sums at the end (two products per thread;
    // measured the faster form for DenseSrc, the other for FoldSrc)
    const gf128 wl = eq_lo ? gf_load(&eq_lo[threadIdx.x & lo_mask]) : gf_one();
    for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const uint64_t x0 = ch * chunk + threadIdx.x;
        if (x0 >= h) continue;
        gf256 w_end = gf256_zero(), w_inf = gf256_zero();
        for (uint32_t k = 0; k < R; k++) {
            uint64_t x = x0 + (uint64_t)k * kT;
            if (x >= h) break;
            RoundVals v = src.at(x, h, tab);
            gf128 e_end = send_one ? v.eh : v.el;
            gf128 o_end = send_one ? v.oh : v.ol;
            if constexpr (ColW) {
                gf128 p_end = gf_mul(e_end, o_end);
                gf128 p_inf = gf_mul(gf_add(v.eh, v.el), gf_add(v.oh, v.ol));
                gf128 wc = gf_load(&eq_c[x & cmask]);
                gf256_acc(&w_end, gf_mul_wide(p_end, wc));
                gf256_acc(&w_inf, gf_mul_wide(p_inf, wc));
            } else {
                gf256_acc(&w_end, gf_mul_wide(e_end, o_end));
                gf256_acc(&w_inf, gf_mul_wide(gf_add(v.eh, v.el), gf_add(v.oh, v.ol)));
            }
        }
        gf128 wy = gf_load(&eq_y[x0 >> s_c]);
        if (Src::kLoPerChunk && eq_lo) {
            // opaque per chunk, so the product's operand masks of wl are not
            // hoisted out of the loop (they would hold ~36 more VGPRs)
            gf128 w = wl;
            asm volatile("" : "+v"(w.w[0]), "+v"(w.w[1]), "+v"(w.w[2]), "+v"(w.w[3]));
            wy = gf_mul(wy, w);
        }
        acc[0] = gf_add(acc[0], gf_mul(gf_reduce(w_end), wy));
        acc[1] = gf_add(acc[1], gf_mul(gf_reduce(w_inf), wy));
    }
    if (!Src::kLoPerChunk && eq_lo) {
        acc[0] = gf_mul(acc[0], wl);
        acc[1] = gf_mul(acc[1], wl);
    }
    block_xor_store<2>(acc, partials + (size_t)blockIdx.x * 2);
}

// Slots per thread for a round of h slots with eq_c rows of 2^s_c: up to 32,
// the chunk kT * R inside one row, and enough chunks to fill the chip.
static uint32_t round_chunk_r(const Tune& tn, uint64_t h, int s_c) {
    uint64_t row = 1ull << s_c;
    if (row < (uint64_t)kT) return 1;
    uint64_t r = std::min<uint64_t>(tn.round_r_max, row / kT);
    while (r > 1 && h / ((uint64_t)kT * r) < tn.round_min_chunks) r >>= 1;
    return (uint32_t)std::max<uint64_t>(r, 1);
}
static unsigned round_blocks(const Tune& tn, uint64_t h, uint32_t R) {
    uint64_t nch = (h + (uint64_t)kT * R - 1) / ((uint64_t)kT * R);
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(nch, tn.round_block_cap));
}

// materialised halves l, r (each 2h entries)
struct DenseSrc {
    static constexpr bool kLoPerChunk = false;
    static constexpr bool kTable = false;
    const uint4* l;
    const uint4* r;
    __device__ __forceinline__ RoundVals at(uint64_t x, uint64_t h, const uint4*) const {
        return RoundVals{gf_load(&l[x]), gf_load(&l[x + h]), gf_load(&r[x]), gf_load(&r[x + h])};
    }
};

// Round 0 of a layer whose right half gets eq_c multiplied in: slot x
// replaces r[x], r[x+h] by eq_c[x & cmask] * r[x], eq_c[.] * r[x+h] and
// returns those (each r entry is read and written by its own slot only).
struct PremulSrc {
    static constexpr bool kLoPerChunk = false;
    static constexpr bool kTable = false;
    const uint4* l;
    uint4* r;
    const uint4* eq_c;
    uint64_t cmask;
    __device__ __forceinline__ RoundVals at(uint64_t x, uint64_t h, const uint4*) const {
        gf128 wc = gf_load(&eq_c[x & cmask]);
        gf128 ol = gf_mul(gf_load(&r[x]), wc), oh = gf_mul(gf_load(&r[x + h]), wc);
        gf_store(&r[x], ol);
        gf_store(&r[x + h], oh);
        return RoundVals{gf_load(&l[x]), gf_load(&l[x + h]), ol, oh};
    }
};

// The previous round's fold fused in: l, r still hold that round's halves
// (4h entries each); slot x folds entries x, x+2h into x and x+h, x+3h into
// x+h by rho (v + rho (v' - v)), stores them and returns them. Slot x writes
// only x and x+h and reads only x, x+h, x+2h, x+3h (x < h): no other slot
// touches those, so folding in place is safe.
struct FoldSrc {
    static constexpr bool kLoPerChunk = true;
    static constexpr bool kTable = true;
    uint4* l;
    uint4* r;
    uint64_t rho_lo, rho_hi;
    const uint4* rho_p;
    __device__ __forceinline__ void build(uint4* tab) const {
        gf_table_build(tab, challenge_arg(rho_lo, rho_hi, rho_p), threadIdx.x, blockDim.x);
    }
    __device__ __forceinline__ gf128 fold(const uint4* tab, uint4* v, uint64_t i, uint64_t h) const {
        gf128 a = gf_load(&v[i]), b = gf_load(&v[i + 2 * h]);
        gf128 o = gf_add(a, gf_mul_tab(tab, gf_add(a, b)));
        gf_store_nt(&v[i], o);
        return o;
    }
    __device__ __forceinline__ RoundVals at(uint64_t x, uint64_t h, const uint4* tab) const {
        gf128 el = fold(tab, l, x, h), eh = fold(tab, l, x + h, h);
        gf128 ol = fold(tab, r, x, h), oh = fold(tab, r, x + h, h);
        return RoundVals{el, eh, ol, oh};
    }
};

// leaf layer: leaf(row, c) = bit(c, row) ? A[row] : 1, halves of 2^(t+s-1)
// entries (l = rows < 2^(t-1)), so round 0 has h = 2^(t+s-2).
struct LeafSrc {
    static constexpr bool kLoPerChunk = false;
    static constexpr bool kTable = false;
    const uint64_t* packed;  // row-major packed columns
    const uint4* A;
    int t, s;
    __device__ __forceinline__ gf128 leaf(uint64_t row, uint64_t c) const {
        uint64_t word = packed[(c >> 6) * (1ull << t) + row];
        return ((word >> (c & 63)) & 1) ? gf_load(&A[row]) : gf_one();
    }
    __device__ __forceinline__ RoundVals at(uint64_t x, uint64_t h, const uint4* = nullptr) const {
        uint64_t y = x >> s, c = x & ((1ull << s) - 1);
        uint64_t q = 1ull << (t - 2);
        return RoundVals{leaf(y, c), leaf(y + q, c), leaf(y + 2 * q, c), leaf(y + 3 * q, c)};
    }
};

// v[i] = v[i] + rho (v[i+h] - v[i]) for i < h, with rho's window table in LDS
__global__ __launch_bounds__(kT) void k_fold_halves(uint4* __restrict__ v, uint64_t h, uint64_t rho_lo,
                                                    uint64_t rho_hi) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, gf_from_u64(rho_lo, rho_hi), threadIdx.x, blockDim.x);
    __syncthreads();
    for (uint64_t i = grid_tid(); i < h; i += grid_threads()) {
        gf128 a = gf_load(&v[i]), b = gf_load(&v[i + h]);
        gf_store(&v[i], gf_add(a, gf_mul_tab(tab, gf_add(a, b))));
    }
}

// ---- leaf layer by bit patterns (src/bitz/forest.rs bit_rounds) ----
// Leaf (row, c) is 1 or A[row] according to one bit. For the rows
// y, y+q, y+2q, y+3q (q = 2^(t-2)) that meet in one round-0 slot, every
// product of the round takes one of 16 values, indexed by the column's
// 4-bit pattern (bit k = bit(c, y + kq)).
__device__ __forceinline__ uint32_t leaf_pattern4(const uint64_t* col, uint64_t y, uint64_t q, int bit) {
    return (uint32_t)(((col[y] >> bit) & 1ull) | (((col[y + q] >> bit) & 1ull) << 1) |
                      (((col[y + 2 * q] >> bit) & 1ull) << 2) | (((col[y + 3 * q] >> bit) & 1ull) << 3));
}

// qtab[(y*16 + pat)*2 + {0,1}] = eq_y[y] * {p_end, p_inf}(pat)
__global__ void k_leaf_qtab(const uint4* __restrict__ A, int t, const uint4* __restrict__ eq_y, int send_one,
                            uint4* __restrict__ qtab) {
    const uint64_t q = 1ull << (t - 2);
    uint64_t idx = grid_tid();
    if (idx >= q * 16) return;
    uint64_t y = idx >> 4;
    uint32_t pat = (uint32_t)(idx & 15);
    gf128 one = gf_one();
    gf128 el = (pat & 1) ? gf_load(&A[y]) : one;
    gf128 eh = (pat & 2) ? gf_load(&A[y + q]) : one;
    gf128 ol = (pat & 4) ? gf_load(&A[y + 2 * q]) : one;
    gf128 oh = (pat & 8) ? gf_load(&A[y + 3 * q]) : one;
    gf128 wy = gf_load(&eq_y[y]);
    gf128 pe = gf_mul(send_one ? eh : el, send_one ? oh : ol);
    gf128 pi = gf_mul(gf_add(el, eh), gf_add(ol, oh));
    gf_store(&qtab[idx * 2], gf_mul(pe, wy));
    gf_store(&qtab[idx * 2 + 1], gf_mul(pi, wy));
}

// Pattern sums of the leaf layer with weights eq_y[y] * eq_c[c]: the slot
// (y, c) reads K bits of column c at rows y + k*stride (k < K) and the table
// entry qtab[((y << K) + pattern)*2 + {0,1}]; thread (c, ylane) XORs the
// entries over its rows y < ny, then multiplies by eq_c[c].
// one thread's XOR of its rows' table entries (the loop of k_leaf_pattern_sums)
template <int K>
__device__ __forceinline__ void leaf_pattern_rows(const uint64_t* __restrict__ col, int bit, uint64_t yl, uint64_t ny,
                                                  uint64_t ylanes, uint64_t stride,
                                                  const uint4* __restrict__ qtab, uint32_t* e, uint32_t* i) {
    for (uint64_t y = yl; y < ny; y += ylanes) {
        uint32_t pat = 0;
#pragma unroll
        for (int k = 0; k < K; k++) pat |= (uint32_t)((col[y + k * stride] >> bit) & 1ull) << k;
        const uint4* q = &qtab[((y << K) + pat) * 2];
        const uint4 a = q[0], b = q[1];
        e[0] ^= a.x, e[1] ^= a.y, e[2] ^= a.z, e[3] ^= a.w;
        i[0] ^= b.x, i[1] ^= b.y, i[2] ^= b.z, i[3] ^= b.w;
    }
}

template <int K>
__global__ __launch_bounds__(kT) void k_leaf_pattern_sums(const uint64_t* __restrict__ packed, int t, int s,
                                                          uint64_t ny, uint64_t stride,
                                                          const uint4* __restrict__ qtab,
                                                          const uint4* __restrict__ eq_c, uint64_t ylanes,
                                                          uint4* __restrict__ partials) {
    const uint64_t cols = 1ull << s, nrows = 1ull << t;
    gf128 acc[2] = {gf_zero(), gf_zero()};
    uint64_t gt = grid_tid();
    if (gt < cols * ylanes) {
        const uint64_t c = gt & (cols - 1), yl = gt >> s;
        const int bit = (int)(c & 63);
        uint32_t e[4] = {0, 0, 0, 0}, i[4] = {0, 0, 0, 0};
        leaf_pattern_rows<K>(packed + (c >> 6) * nrows, bit, yl, ny, ylanes, stride, qtab, e, i);
        const gf128 wc = gf_load(&eq_c[c]);
        acc[0] = gf_mul(gf128{{e[0], e[1], e[2], e[3]}}, wc);
        acc[1] = gf_mul(gf128{{i[0], i[1], i[2], i[3]}}, wc);
    }
    block_xor_store<2>(acc, partials + (size_t)blockIdx.x * 2);
}

// The same sums from the 16-bit row words (nr[(y' << s) | c], y' < Y4 =
// 2^(t-4): column c's bits at rows y' + w Y4, even w in the low byte, odd w
// in the high byte). Block y' scatters eq_c[c] into 256 buckets per byte of
// its words (LDS XOR), then contracts the buckets with its rows' tables, so
// each table entry is read once:
//   K = 8, ny = 2 Y4: row y = y' + h Y4 reads byte h (pattern bit k <-> row
//     y + k ny = y' + Y4 (h + 2k), word bit h + 2k);
//   K = 4, ny = 4 Y4: row y = y' + v Y4 reads bits (v >> 1) + 2k of byte
//     v & 1 (row y' + Y4 (v + 4k)); the 256 buckets of that byte are first
//     summed down to the 16 patterns.
template <int K>
__global__ __launch_bounds__(kT) void k_byte_bucket_sums(const uint16_t* __restrict__ nr, int t, int s,
                                                         const uint4* __restrict__ qtab,
                                                         const uint4* __restrict__ eq_c,
                                                         uint4* __restrict__ partials) {
    static_assert(K == 4 || K == 8, "K is 4 or 8");
    __shared__ uint32_t B[4][2][256];  // [word][byte][index]
    const uint64_t yp = blockIdx.x, y4 = 1ull << (t - 4), cols = 1ull << s;
    for (uint32_t i = threadIdx.x; i < 4 * 2 * 256; i += kT) (&B[0][0][0])[i] = 0u;
    __syncthreads();
    const uint16_t* wp = nr + (yp << s);
    for (uint64_t c = threadIdx.x; c < cols; c += kT) {
        const uint32_t w = wp[c], lo = w & 255u, hi = w >> 8;
        const uint4 e = eq_c[c];
        atomicXor(&B[0][0][lo], e.x);
        atomicXor(&B[1][0][lo], e.y);
        atomicXor(&B[2][0][lo], e.z);
        atomicXor(&B[3][0][lo], e.w);
        atomicXor(&B[0][1][hi], e.x);
        atomicXor(&B[1][1][hi], e.y);
        atomicXor(&B[2][1][hi], e.z);
        atomicXor(&B[3][1][hi], e.w);
    }
    __syncthreads();
    gf256 end = gf256_zero(), inf = gf256_zero();
    if constexpr (K == 8) {
        const uint32_t i = threadIdx.x;
#pragma unroll
        for (uint32_t h = 0; h < 2; h++) {
            const gf128 bv = gf128{{B[0][h][i], B[1][h][i], B[2][h][i], B[3][h][i]}};
            const uint4* q = qtab + ((((yp + h * y4) << 8) + i) << 1);
            gf256_acc(&end, gf_mul_wide(gf_load(&q[0]), bv));
            gf256_acc(&inf, gf_mul_wide(gf_load(&q[1]), bv));
        }
    } else if (threadIdx.x < 64) {
        const uint32_t v = threadIdx.x >> 4, p = threadIdx.x & 15u, byte = v & 1u, sh = v >> 1;
        uint32_t fixed = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) fixed |= ((p >> k) & 1u) << (sh + 2 * k);
        uint32_t m[4] = {0, 0, 0, 0};
        for (uint32_t f = 0; f < 16; f++) {
            uint32_t b = fixed;
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) b |= ((f >> k) & 1u) << ((1 - sh) + 2 * k);
#pragma unroll
            for (int wd = 0; wd < 4; wd++) m[wd] ^= B[wd][byte][b];
        }
        const gf128 bv = gf128{{m[0], m[1], m[2], m[3]}};
        const uint4* q = qtab + ((((yp + v * y4) << 4) + p) << 1);
        gf256_acc(&end, gf_mul_wide(gf_load(&q[0]), bv));
        gf256_acc(&inf, gf_mul_wide(gf_load(&q[1]), bv));
    }
    gf128 sums[2] = {gf_reduce(end), gf_reduce(inf)};
    block_xor_store<2>(sums, partials + (size_t)blockIdx.x * 2);
}

// Round 1 of the leaf layer. After round 0 folded by rho0, the halves are
// l'(y, c) = ftab[y*8 + pat2(y, y+q)] and r'(y, c) = ftab[y*8 + 4 + pat2(y+2q, y+3q)]
// (q = 2^(t-2)); a round-1 slot y1 < q8 = q/2 meets l'(y1), l'(y1+q8),
// r'(y1), r'(y1+q8): eight rows y1 + k q8, bit k of the 8-bit pattern.
// q1tab[(y1*256 + pat)*2 + {0,1}] = eq_y[y1] * {p_end, p_inf}(pat).
__device__ __forceinline__ void leaf_round1_vals(const uint4* ftab, uint64_t y1, uint64_t q8, uint32_t pat,
                                                 gf128* el, gf128* eh, gf128* ol, gf128* oh) {
    *el = gf_load(&ftab[y1 * 8 + ((pat & 1) | ((pat >> 1) & 2))]);
    *eh = gf_load(&ftab[(y1 + q8) * 8 + (((pat >> 1) & 1) | ((pat >> 2) & 2))]);
    *ol = gf_load(&ftab[y1 * 8 + 4 + (((pat >> 4) & 1) | ((pat >> 5) & 2))]);
    *oh = gf_load(&ftab[(y1 + q8) * 8 + 4 + (((pat >> 5) & 1) | ((pat >> 6) & 2))]);
}

__global__ void k_leaf_q1tab(const uint4* __restrict__ ftab, int t, const uint4* __restrict__ eq_y, int send_one,
                             uint4* __restrict__ q1tab) {
    const uint64_t q8 = 1ull << (t - 3);
    uint64_t idx = grid_tid();
    if (idx >= q8 * 256) return;
    uint64_t y1 = idx >> 8;
    gf128 el, eh, ol, oh;
    leaf_round1_vals(ftab, y1, q8, (uint32_t)(idx & 255), &el, &eh, &ol, &oh);
    gf128 wy = gf_load(&eq_y[y1]);
    gf128 pe = send_one ? gf_mul(eh, oh) : gf_mul(el, ol);
    gf128 pi = gf_mul(gf_add(el, eh), gf_add(ol, oh));
    gf_store(&q1tab[idx * 2], gf_mul(pe, wy));
    gf_store(&q1tab[idx * 2 + 1], gf_mul(pi, wy));
}

// The leaf layer after two folds (rho0, then rho1): gtab[y1*32 + half*16 + sub]
// with sub the half's 4 pattern bits (l: bits 0-3, r: bits 4-7).
__global__ __launch_bounds__(kT) void k_leaf_gtab(const uint4* __restrict__ ftab, int t, uint64_t rho_lo,
                                                  uint64_t rho_hi, const uint4* rho_p, uint4* __restrict__ gtab) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, challenge_arg(rho_lo, rho_hi, rho_p), threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t q8 = 1ull << (t - 3);
    uint64_t idx = grid_tid();
    if (idx >= q8 * 32) return;
    uint64_t y1 = idx >> 5;
    uint32_t half = (uint32_t)((idx >> 4) & 1), sub = (uint32_t)(idx & 15);
    gf128 el, eh, ol, oh;
    leaf_round1_vals(ftab, y1, q8, sub << (4 * half), &el, &eh, &ol, &oh);
    gf128 lo = half ? ol : el, hi = half ? oh : eh;
    gf_store(&gtab[idx], gf_add(lo, gf_mul_tab(tab, gf_add(lo, hi))));
}

// Folded halves written from 8-bit row patterns (rows y + k q8, q8 =
// 2^(t-3)): l_out[x] = gtab[y*32 + lsub], r_out[x] = gtab[y*32 + 16 + rsub].
// Layout 0 (leaf layer after two folds): lsub = bits 0-3, rsub = bits 4-7.
// Layout 1 (level-1 layer after one fold): lsub = bits 0,1,4,5 and
// rsub = bits 2,3,6,7. eq_c != nullptr: the right half is written multiplied
// by eq_c[c].
template <int Layout>
__global__ __launch_bounds__(kT) void k_pattern_fold(const uint64_t* __restrict__ packed, int t, int s,
                                                     const uint4* __restrict__ gtab, const uint4* __restrict__ eq_c,
                                                     uint4* __restrict__ l_out, uint4* __restrict__ r_out) {
    const uint64_t q8 = 1ull << (t - 3), nrows = 1ull << t, h = 1ull << (t + s - 3);
    for (uint64_t x = grid_tid(); x < h; x += grid_threads()) {
        uint64_t y1 = x >> s, c = x & ((1ull << s) - 1);
        const uint64_t* col = packed + (c >> 6) * nrows;
        int bit = (int)(c & 63);
        uint32_t pat = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) pat |= (uint32_t)((col[y1 + k * q8] >> bit) & 1ull) << k;
        uint32_t lsub = Layout == 0 ? (pat & 15) : ((pat & 3) | ((pat >> 2) & 12));
        uint32_t rsub = Layout == 0 ? (pat >> 4) : (((pat >> 2) & 3) | ((pat >> 4) & 12));
        l_out[x] = gtab[y1 * 32 + lsub];
        if (eq_c)
            gf_store(&r_out[x], gf_mul(gf_load(&gtab[y1 * 32 + 16 + rsub]), gf_load(&eq_c[c])));
        else
            r_out[x] = gtab[y1 * 32 + 16 + rsub];
    }
}

// Column weights split as eq_c[c] = eq_lo[c & 255] * eq_hi[c >> 8] (the eq
// table is a tensor product over the bits of c): a 256-column run of one row
// shares its eq_hi factor, so a block that keeps to one run multiplies by a
// constant through its window table (~125 VALU) where eq_c[c] would take a
// variable product (~430); eq_lo goes to the row-only round sums, one
// factor per thread. Block b takes run g = b mod 2^(s-8) and rows
// b >> (s-8), + row_lanes.
template <int Layout>
__global__ __launch_bounds__(kT) void k_pattern_fold_hi(const uint64_t* __restrict__ packed, int t, int s,
                                                        const uint4* __restrict__ gtab,
                                                        const uint4* __restrict__ eq_hi, uint32_t row_lanes,
                                                        uint4* __restrict__ l_out, uint4* __restrict__ r_out) {
    __shared__ uint4 tab[512];
    const uint64_t q8 = 1ull << (t - 3), nrows = 1ull << t;
    const uint32_t g = blockIdx.x & ((1u << (s - 8)) - 1);
    gf_table_build(tab, gf_load(&eq_hi[g]), threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t c = ((uint64_t)g << 8) + threadIdx.x;
    const uint64_t* col = packed + (c >> 6) * nrows;
    const int bit = (int)(c & 63);
    for (uint64_t y1 = blockIdx.x >> (s - 8); y1 < q8; y1 += row_lanes) {
        const uint64_t x = (y1 << s) + c;
        uint32_t pat = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) pat |= (uint32_t)((col[y1 + k * q8] >> bit) & 1ull) << k;
        uint32_t lsub = Layout == 0 ? (pat & 15) : ((pat & 3) | ((pat >> 2) & 12));
        uint32_t rsub = Layout == 0 ? (pat >> 4) : (((pat >> 2) & 3) | ((pat >> 4) & 12));
        gf_store_nt(&l_out[x], gf_load(&gtab[y1 * 32 + lsub]));
        gf_store_nt(&r_out[x], gf_mul_tab(tab, gf_load(&gtab[y1 * 32 + 16 + rsub])));
    }
}

// v[x] *= eq_hi[(x & (2^s - 1)) >> 8] for x < rows * 2^s, blocks as above
__global__ __launch_bounds__(kT) void k_scale_cols_hi(uint4* __restrict__ v, uint64_t rows, int s,
                                                      const uint4* __restrict__ eq_hi, uint32_t row_lanes) {
    __shared__ uint4 tab[512];
    const uint32_t g = blockIdx.x & ((1u << (s - 8)) - 1);
    gf_table_build(tab, gf_load(&eq_hi[g]), threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t c = ((uint64_t)g << 8) + threadIdx.x;
    for (uint64_t y = blockIdx.x >> (s - 8); y < rows; y += row_lanes) {
        uint4* p = &v[(y << s) + c];
        gf_store_nt(p, gf_mul_tab(tab, gf_load(p)));
    }
}

// ---- level 1 by bit patterns (src/bitz/forest.rs prove_jit_level) ----
// Level-1 entry (y, c), y < Y1 = 2^(t-1), is one of vtab[y*4 + p]: 1, A[y],
// A[y+Y1], A[y] A[y+Y1] for p = bit(c, y) | bit(c, y+Y1) << 1.
__global__ void k_level1_vtab(const uint4* __restrict__ A, int t, uint4* __restrict__ vtab) {
    const uint64_t Y1 = 1ull << (t - 1);
    uint64_t y = grid_tid();
    if (y >= Y1) return;
    gf128 a = gf_load(&A[y]), b = gf_load(&A[y + Y1]);
    gf_store(&vtab[y * 4], gf_one());
    gf_store(&vtab[y * 4 + 1], a);
    gf_store(&vtab[y * 4 + 2], b);
    gf_store(&vtab[y * 4 + 3], gf_mul(a, b));
}

// The layer over level 1 (halves: rows [0, Y2) and [Y2, Y1), Y2 = 2^(t-2)):
// a round-0 slot y < Y3 = 2^(t-3) meets level-1 rows y, y+Y3 (left) and
// y+Y2, y+Y2+Y3 (right); with pattern bit k = bit(c, y + k Y3) their
// values are vtab[row][bits (k, k+4)] for k = 0, 1, 2, 3.
__device__ __forceinline__ void level1_round0_vals(const uint4* vtab, uint64_t y, uint64_t Y3, uint32_t pat,
                                                   gf128* el, gf128* eh, gf128* ol, gf128* oh) {
    *el = gf_load(&vtab[y * 4 + ((pat & 1) | ((pat >> 3) & 2))]);
    *eh = gf_load(&vtab[(y + Y3) * 4 + (((pat >> 1) & 1) | ((pat >> 4) & 2))]);
    *ol = gf_load(&vtab[(y + 2 * Y3) * 4 + (((pat >> 2) & 1) | ((pat >> 5) & 2))]);
    *oh = gf_load(&vtab[(y + 3 * Y3) * 4 + (((pat >> 3) & 1) | ((pat >> 6) & 2))]);
}

__global__ void k_level1_qtab(const uint4* __restrict__ vtab, int t, const uint4* __restrict__ eq_y, int send_one,
                              uint4* __restrict__ qtab) {
    const uint64_t Y3 = 1ull << (t - 3);
    uint64_t idx = grid_tid();
    if (idx >= Y3 * 256) return;
    uint64_t y = idx >> 8;
    gf128 el, eh, ol, oh;
    level1_round0_vals(vtab, y, Y3, (uint32_t)(idx & 255), &el, &eh, &ol, &oh);
    gf128 wy = gf_load(&eq_y[y]);
    gf128 pe = send_one ? gf_mul(eh, oh) : gf_mul(el, ol);
    gf128 pi = gf_mul(gf_add(el, eh), gf_add(ol, oh));
    gf_store(&qtab[idx * 2], gf_mul(pe, wy));
    gf_store(&qtab[idx * 2 + 1], gf_mul(pi, wy));
}

// gtab[y*32 + half*16 + sub] = lo + rho (lo + hi) for the level-1 layer's
// round-0 fold; sub = (s0, s1, s2, s3) are pattern bits (k, k+1, k+4, k+5)
// with k = 2*half, lo = vtab[row][s0 | s2 << 1], hi = vtab[row + Y3][s1 | s3 << 1].
__global__ __launch_bounds__(kT) void k_level1_gtab(const uint4* __restrict__ vtab, int t, uint64_t rho_lo,
                                                    uint64_t rho_hi, const uint4* rho_p, uint4* __restrict__ gtab) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, challenge_arg(rho_lo, rho_hi, rho_p), threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t Y3 = 1ull << (t - 3);
    uint64_t idx = grid_tid();
    if (idx >= Y3 * 32) return;
    uint64_t y = idx >> 5;
    uint32_t half = (uint32_t)((idx >> 4) & 1), sub = (uint32_t)(idx & 15);
    uint64_t row = y + 2 * half * Y3;
    gf128 lo = gf_load(&vtab[row * 4 + ((sub & 1) | ((sub >> 1) & 2))]);
    gf128 hi = gf_load(&vtab[(row + Y3) * 4 + (((sub >> 1) & 1) | ((sub >> 2) & 2))]);
    gf_store(&gtab[idx], gf_add(lo, gf_mul_tab(tab, gf_add(lo, hi))));
}

// Level 2 from the bits: entry (y, c), y < Y2, is ltab[y*16 + p] =
// vtab[y][p & 3] * vtab[y + Y2][p >> 2] with p = bits of rows y, y+Y1,
// y+Y2, y+Y2+Y1.
__global__ void k_level2_ltab(const uint4* __restrict__ vtab, int t, uint4* __restrict__ ltab) {
    const uint64_t Y2 = 1ull << (t - 2);
    uint64_t idx = grid_tid();
    if (idx >= Y2 * 16) return;
    uint64_t y = idx >> 4;
    uint32_t p = (uint32_t)(idx & 15);
    gf_store(&ltab[idx], gf_mul(gf_load(&vtab[y * 4 + (p & 3)]), gf_load(&vtab[(y + Y2) * 4 + (p >> 2)])));
}

// Level 3 from the bits: entry (y, c), y < Y3, is l3tab[y*256 + p] with p
// bit k = bit(c, y + k Y3); as level3(y) = level2(y) level2(y + Y3) and a
// level-2 entry reads rows y', y'+Y1, y'+Y2, y'+Y2+Y1 (k offsets 0, 4, 2, 6),
// l3tab[y][p] = ltab[y][b0 b4 b2 b6] * ltab[y + Y3][b1 b5 b3 b7].
__global__ void k_level3_tab(const uint4* __restrict__ ltab, int t, uint4* __restrict__ l3tab) {
    const uint64_t Y3 = 1ull << (t - 3);
    uint64_t idx = grid_tid();
    if (idx >= Y3 * 256) return;
    uint64_t y = idx >> 8;
    uint32_t p = (uint32_t)(idx & 255);
    uint32_t lo = (p & 1) | ((p >> 3) & 2) | ((p << 0) & 4) | ((p >> 3) & 8);
    uint32_t hi = ((p >> 1) & 1) | ((p >> 4) & 2) | ((p >> 1) & 4) | ((p >> 4) & 8);
    gf_store(&l3tab[idx], gf_mul(gf_load(&ltab[y * 16 + lo]), gf_load(&ltab[(y + Y3) * 16 + hi])));
}

__global__ __launch_bounds__(kT) void k_level3_pattern(const uint64_t* __restrict__ packed, int t, int s,
                                                       const uint4* __restrict__ l3tab, uint4* __restrict__ out) {
    const uint64_t nrows = 1ull << t, Y3 = nrows >> 3, n = 1ull << (t + s - 3);
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) {
        uint64_t y = x >> s, c = x & ((1ull << s) - 1);
        const uint64_t* col = packed + (c >> 6) * nrows;
        int bit = (int)(c & 63);
        uint32_t p = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) p |= (uint32_t)((col[y + k * Y3] >> bit) & 1ull) << k;
        out[x] = l3tab[y * 256 + p];
    }
}

__global__ __launch_bounds__(kT) void k_level2_pattern(const uint64_t* __restrict__ packed, int t, int s,
                                                       const uint4* __restrict__ ltab, uint4* __restrict__ out) {
    const uint64_t nrows = 1ull << t, Y1 = nrows >> 1, Y2 = nrows >> 2, n = 1ull << (t + s - 2);
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) {
        uint64_t y = x >> s, c = x & ((1ull << s) - 1);
        const uint64_t* col = packed + (c >> 6) * nrows;
        int bit = (int)(c & 63);
        uint32_t p = (uint32_t)(((col[y] >> bit) & 1ull) | (((col[y + Y1] >> bit) & 1ull) << 1) |
                                (((col[y + Y2] >> bit) & 1ull) << 2) | (((col[y + Y2 + Y1] >> bit) & 1ull) << 3));
        out[x] = ltab[y * 16 + p];
    }
}

// The same two levels with the row's table staged in LDS, for rows of at
// least kT entries (s >= 8): block b takes 2^cs consecutive entries of row
// y = b >> (s - cs), so the 16 (level 2) or 256 (level 3) table entries are
// read from HBM once per block and every entry is an LDS read instead of a
// 16-byte gather.
template <int K>
__global__ __launch_bounds__(kT) void k_level_pattern_lds(const uint64_t* __restrict__ packed, int t, int s, int cs,
                                                          const uint4* __restrict__ tab_g, uint4* __restrict__ out) {
    __shared__ uint4 tab[1 << K];
    const uint64_t nrows = 1ull << t;
    const uint64_t y = (uint64_t)blockIdx.x >> (s - cs);
    const uint64_t c0 = ((uint64_t)blockIdx.x & ((1ull << (s - cs)) - 1)) << cs;
    for (int i = threadIdx.x; i < (1 << K); i += kT) tab[i] = tab_g[(y << K) + i];
    __syncthreads();
    uint64_t off[K];
    if constexpr (K == 4) {
        const uint64_t Y1 = nrows >> 1, Y2 = nrows >> 2;
        off[0] = y;
        off[1] = y + Y1;
        off[2] = y + Y2;
        off[3] = y + Y2 + Y1;
    } else {
        const uint64_t Y3 = nrows >> 3;
#pragma unroll
        for (int k = 0; k < K; k++) off[k] = y + k * Y3;
    }
    for (uint64_t i = threadIdx.x; i < (1ull << cs); i += kT) {
        const uint64_t c = c0 + i;
        const uint64_t* col = packed + (c >> 6) * nrows;
        const int bit = (int)(c & 63);
        uint32_t p = 0;
#pragma unroll
        for (int k = 0; k < K; k++) p |= (uint32_t)((col[off[k]] >> bit) & 1ull) << k;
        gf_store_nt(&out[(y << s) + c], gf_load(&tab[p]));
    }
}

// ftab[y*8 + half*4 + pat] = lo + rho (lo + hi) with lo/hi the leaves of rows
// y + 2*half*q and y + (2*half+1)*q selected by the 2-bit pattern.
__global__ __launch_bounds__(kT) void k_leaf_ftab(const uint4* __restrict__ A, int t, uint64_t rho_lo,
                                                  uint64_t rho_hi, const uint4* rho_p, uint4* __restrict__ ftab) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, challenge_arg(rho_lo, rho_hi, rho_p), threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t q = 1ull << (t - 2);
    uint64_t idx = grid_tid();
    if (idx >= q * 8) return;
    uint64_t y = idx >> 3;
    uint32_t half = (uint32_t)((idx >> 2) & 1), pat = (uint32_t)(idx & 3);
    gf128 lo = (pat & 1) ? gf_load(&A[y + 2 * half * q]) : gf_one();
    gf128 hi = (pat & 2) ? gf_load(&A[y + (2 * half + 1) * q]) : gf_one();
    gf_store(&ftab[idx], gf_add(lo, gf_mul_tab(tab, gf_add(lo, hi))));
}

__global__ __launch_bounds__(kT) void k_leaf_fold_pattern(const uint64_t* __restrict__ packed, int t, int s,
                                                          const uint4* __restrict__ ftab, uint4* __restrict__ l_out,
                                                          uint4* __restrict__ r_out) {
    const uint64_t q = 1ull << (t - 2), nrows = 1ull << t, h = 1ull << (t + s - 2);
    for (uint64_t x = grid_tid(); x < h; x += grid_threads()) {
        uint64_t y = x >> s, c = x & ((1ull << s) - 1);
        uint32_t pat = leaf_pattern4(packed + (c >> 6) * nrows, y, q, (int)(c & 63));
        l_out[x] = ftab[y * 8 + (pat & 3)];
        r_out[x] = ftab[y * 8 + 4 + (pat >> 2)];
    }
}

// Level 1 from the bits: entry (y, c) is 1, A[y], A[y + half] or
// pa[y] = A[y] * A[y + half] by the two bits.
__global__ __launch_bounds__(kT) void k_level1_pattern(const uint64_t* __restrict__ packed, int t, int s,
                                                       const uint4* __restrict__ A, const uint4* __restrict__ pa,
                                                       uint4* __restrict__ out) {
    const uint64_t nrows = 1ull << t, half = nrows >> 1, n = 1ull << (t + s - 1);
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) {
        uint64_t y = x >> s, c = x & ((1ull << s) - 1);
        const uint64_t* col = packed + (c >> 6) * nrows;
        int bit = (int)(c & 63);
        uint32_t b0 = (uint32_t)((col[y] >> bit) & 1ull), b1 = (uint32_t)((col[y + half] >> bit) & 1ull);
        uint4 v = make_uint4(1u, 0u, 0u, 0u);
        if (b0 & b1)
            v = pa[y];
        else if (b0)
            v = A[y];
        else if (b1)
            v = A[y + half];
        out[x] = v;
    }
}

// Both halves l, r (2h entries each) folded by rho in one launch.
__global__ __launch_bounds__(kT) void k_fold_pair(uint4* __restrict__ l, uint4* __restrict__ r, uint64_t h,
                                                  uint64_t rho_lo, uint64_t rho_hi, const uint4* rho_p) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, challenge_arg(rho_lo, rho_hi, rho_p), threadIdx.x, blockDim.x);
    __syncthreads();
    for (uint64_t i = grid_tid(); i < 2 * h; i += grid_threads()) {
        uint4* v = i < h ? l : r;
        uint64_t j = i < h ? i : i - h;
        gf128 a = gf_load(&v[j]), b = gf_load(&v[j + h]);
        gf_store(&v[j], gf_add(a, gf_mul_tab(tab, gf_add(a, b))));
    }
}

// out[i] = in[i] * in[i + n/2]
__global__ void k_product_level(const uint4* __restrict__ in, uint64_t half, uint4* __restrict__ out) {
    for (uint64_t i = grid_tid(); i < half; i += grid_threads())
        gf_store(&out[i], gf_mul(gf_load(&in[i]), gf_load(&in[i + half])));
}

// ------------------------------------------------- inner-product sumcheck
// m[b] = sum_c w2[c] bit(c, b) from row-major packed columns, through 4-bit
// tables of w2 in LDS (four-Russians): block
// (column chunk of 1024 columns, 2048 rows) builds T[n][v] = sum of w2 over
// the set bits of v for the chunk's 256 nibble positions n (64 KiB), then each
// row XORs one table entry per nibble of its 16 packed words: 256 LDS reads
// per row and chunk in place of a dependent global gather per set bit. Chunk
// k's sums go to partial[k][b] (out itself when there is one chunk).
constexpr int kXiChunkWords = 16, kXiRows = 2048;
__global__ __launch_bounds__(kT) void k_xi_nibbles(const uint64_t* __restrict__ packed, int t, int s,
                                                   const uint4* __restrict__ w2, uint4* __restrict__ partial) {
    __shared__ uint4 tab[kXiChunkWords * 16 * 16];
    const uint64_t nrows = 1ull << t, cols = 1ull << s, groups = (cols + 63) >> 6;
    const uint64_t g0 = (uint64_t)blockIdx.x * kXiChunkWords;
    const uint32_t nw = (uint32_t)std::min<uint64_t>(kXiChunkWords, groups - g0);
    for (uint32_t e = threadIdx.x; e < nw * 256; e += kT) {
        const uint64_t c0 = g0 * 64 + (e >> 4) * 4;  // first column of nibble position e >> 4
        const uint32_t v = e & 15;
        uint32_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;
        for (int k = 0; k < 4; k++)
            if (((v >> k) & 1) && c0 + k < cols) {
                const uint4 q = w2[c0 + k];
                x0 ^= q.x, x1 ^= q.y, x2 ^= q.z, x3 ^= q.w;
            }
        tab[e] = make_uint4(x0, x1, x2, x3);
    }
    __syncthreads();
    const uint64_t row0 = (uint64_t)blockIdx.y * kXiRows;
    const uint64_t rows_here = std::min<uint64_t>(kXiRows, nrows - row0);
    for (uint64_t i = threadIdx.x; i < rows_here; i += kT) {
        const uint64_t b = row0 + i;
        uint32_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;
        for (uint32_t w = 0; w < nw; w++) {
            const uint64_t word = packed[(g0 + w) * nrows + b];
#pragma unroll
            for (int nib = 0; nib < 16; nib++) {
                const uint4 q = tab[(w * 16 + nib) * 16 + (uint32_t)((word >> (4 * nib)) & 15)];
                x0 ^= q.x, x1 ^= q.y, x2 ^= q.z, x3 ^= q.w;
            }
        }
        partial[(uint64_t)blockIdx.x * nrows + b] = make_uint4(x0, x1, x2, x3);
    }
}

// out[b] = XOR over k < chunks of partial[k][b]
__global__ void k_xor_chunks(const uint4* __restrict__ partial, uint32_t chunks, uint64_t n, uint4* __restrict__ out) {
    for (uint64_t b = grid_tid(); b < n; b += grid_threads()) {
        uint4 a = partial[b];
        for (uint32_t k = 1; k < chunks; k++) {
            const uint4 q = partial[(uint64_t)k * n + b];
            a.x ^= q.x, a.y ^= q.y, a.z ^= q.z, a.w ^= q.w;
        }
        out[b] = a;
    }
}

// partial[chunk][c] = sum over the chunk's rows b of eq[b] bit(c, b), from
// column-major rows, through nibble tables (four Russians): block (group of
// 64 columns, chunk of kRowsEq = 1024 rows) builds the 16 XOR sums of each
// 4-row nibble's eq weights in LDS (256 nibbles x 16 entries, 64 KiB), and
// the four threads of a column (quarters of its words, adjacent lanes) XOR
// one entry per nibble of their words; the quarters merge by lane XOR and
// each chunk's column sums go to `partial` (reduce_columns adds the chunks).
constexpr int kRowsEq = 1024;
__global__ __launch_bounds__(kT) void k_fold_rows_eq(const uint64_t* __restrict__ rows, int t, int s,
                                                     const uint4* __restrict__ eq, uint4* __restrict__ partial) {
    __shared__ uint4 tab[kRowsEq / 4 * 16];
    const uint64_t words = 1ull << (t - 6), cols = 1ull << s;
    const uint32_t chunk_rows = (uint32_t)std::min<uint64_t>(kRowsEq, 1ull << t);
    const uint64_t row0 = (uint64_t)blockIdx.y * chunk_rows;
    if (threadIdx.x < chunk_rows / 4) {
        const uint4* e = eq + row0 + 4 * threadIdx.x;
        gf128 w[4], acc[16];
#pragma unroll
        for (int j = 0; j < 4; j++) w[j] = gf_load(&e[j]);
        acc[0] = gf_zero();
#pragma unroll
        for (int v = 1; v < 16; v++) acc[v] = gf_add(acc[v & (v - 1)], w[__builtin_ctz(v)]);
        uint4* tn = tab + 16 * threadIdx.x;
#pragma unroll
        for (int v = 0; v < 16; v++) gf_store(&tn[v], acc[v]);
    }
    __syncthreads();
    const uint64_t c = (uint64_t)blockIdx.x * 64 + (threadIdx.x >> 2);
    const uint32_t q = threadIdx.x & 3, cw = chunk_rows / 64;
    uint32_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;
    if (c < cols) {
        const uint64_t* col = rows + c * words + row0 / 64;
        for (uint32_t w = q; w < cw; w += 4) {
            const uint64_t x = col[w];
            const uint4* tw = tab + 256 * w;
#pragma unroll
            for (int n = 0; n < 16; n++) {
                const uint4 e = tw[16 * n + (uint32_t)((x >> (4 * n)) & 15u)];
                x0 ^= e.x, x1 ^= e.y, x2 ^= e.z, x3 ^= e.w;
            }
        }
    }
    // the four quarters of a column are adjacent lanes
#pragma unroll
    for (int off = 1; off < 4; off <<= 1) {
        x0 ^= __shfl_xor(x0, off);
        x1 ^= __shfl_xor(x1, off);
        x2 ^= __shfl_xor(x2, off);
        x3 ^= __shfl_xor(x3, off);
    }
    if (q == 0 && c < cols) partial[(uint64_t)blockIdx.y * cols + c] = make_uint4(x0, x1, x2, x3);
}

// [sum m0 w0, sum m1 w1, sum (m0+m1)(w0+w1)] over pairs (2k, 2k+1)
__global__ __launch_bounds__(kT) void k_round_poly(const uint4* __restrict__ m, const uint4* __restrict__ w,
                                                   uint64_t pairs, uint4* __restrict__ partials) {
    gf256 a0 = gf256_zero(), a1 = gf256_zero(), q = gf256_zero();
    for (uint64_t k = grid_tid(); k < pairs; k += grid_threads()) {
        gf128 m0 = gf_load(&m[2 * k]), m1 = gf_load(&m[2 * k + 1]);
        gf128 w0 = gf_load(&w[2 * k]), w1 = gf_load(&w[2 * k + 1]);
        gf256_acc(&a0, gf_mul_wide(m0, w0));
        gf256_acc(&a1, gf_mul_wide(m1, w1));
        gf256_acc(&q, gf_mul_wide(gf_add(m0, m1), gf_add(w0, w1)));
    }
    gf128 acc[3] = {gf_reduce(a0), gf_reduce(a1), gf_reduce(q)};
    block_xor_store<3>(acc, partials + (size_t)blockIdx.x * 3);
}

// out[i] = v[2i] + r (v[2i] + v[2i+1]); out may alias v (reads precede the
// write of the same index and out index i <= 2i).
__global__ __launch_bounds__(kT) void k_fold_lsb(const uint4* v, uint64_t half, uint64_t r_lo, uint64_t r_hi,
                                                 uint4* out) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, gf_from_u64(r_lo, r_hi), threadIdx.x, blockDim.x);
    __syncthreads();
    for (uint64_t i = grid_tid(); i < half; i += grid_threads()) {
        gf128 v0 = gf_load(&v[2 * i]), v1 = gf_load(&v[2 * i + 1]);
        gf_store(&out[i], gf_add(v0, gf_mul_tab(tab, gf_add(v0, v1))));
    }
}

// ------------------------------------------------------------ ring switch
// hat_s[i] = sum_x bit_i(p[x]) suffix[x] for the 128 bits i of p
// (fold_1b_rows_naive, ring_switch.rs:1332-1381). 128 accumulators do not
// fit a thread, so block (j, chunk) owns bits 8j..8j+7 over a contiguous
// chunk of elements; the 16 blocks of a chunk are consecutive in the grid,
// run together and share the chunk's loads through the caches (the data is
// read from HBM about once, not 16 times).
__global__ __launch_bounds__(kT) void k_fold_1b_rows(const uint4* __restrict__ p, const uint4* __restrict__ suffix,
                                                     uint64_t n, uint64_t per, uint4* __restrict__ partials /* [chunks][128] */) {
    const int j = blockIdx.x & 15;
    const uint64_t chunk = blockIdx.x >> 4;
    const uint64_t x0 = chunk * per, x1 = std::min<uint64_t>(n, x0 + per);
    gf128 acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = gf_zero();
    for (uint64_t x = x0 + threadIdx.x; x < x1; x += kT) {
        const uint4 pv = p[x];
        const uint32_t word = (&pv.x)[j >> 2];
        const uint32_t byte = (word >> ((j & 3) * 8)) & 255u;
        if (!byte) continue;
        const gf128 sv = gf_load(&suffix[x]);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t m = 0u - ((byte >> k) & 1u);
            acc[k].w[0] ^= sv.w[0] & m;
            acc[k].w[1] ^= sv.w[1] & m;
            acc[k].w[2] ^= sv.w[2] & m;
            acc[k].w[3] ^= sv.w[3] & m;
        }
    }
    block_xor_store<8>(acc, partials + (size_t)chunk * 128 + 8 * j);
}

// basis[x] = sum over set bits i of suffix[x] of w[i], via 16 byte tables
// T[j][v] = sum_{bits k of v} w[8j+k] in LDS (ring_switch.rs:1486-1560).
__global__ __launch_bounds__(kT) void k_fold_b128(const uint4* __restrict__ suffix, uint64_t n,
                                                  const uint4* __restrict__ w, uint4* __restrict__ out) {
    __shared__ uint4 tab[16 * 256];
    for (int e = threadIdx.x; e < 16 * 256; e += blockDim.x) {
        int j = e >> 8, v = e & 255;
        uint32_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;
        for (int k = 0; k < 8; k++)
            if ((v >> k) & 1) {
                uint4 q = w[8 * j + k];
                x0 ^= q.x;
                x1 ^= q.y;
                x2 ^= q.z;
                x3 ^= q.w;
            }
        tab[e] = make_uint4(x0, x1, x2, x3);
    }
    __syncthreads();
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) {
        uint4 sv = suffix[x];
        uint32_t words[4] = {sv.x, sv.y, sv.z, sv.w};
        uint32_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            uint4 q = tab[j * 256 + ((words[j >> 2] >> (8 * (j & 3))) & 255u)];
            x0 ^= q.x;
            x1 ^= q.y;
            x2 ^= q.z;
            x3 ^= q.w;
        }
        out[x] = make_uint4(x0, x1, x2, x3);
    }
}

// acc[x] += eta * e[x]
__global__ __launch_bounds__(kT) void k_axpy(uint4* __restrict__ acc, const uint4* __restrict__ e, uint64_t n,
                                             uint64_t eta_lo, uint64_t eta_hi) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, gf_from_u64(eta_lo, eta_hi), threadIdx.x, blockDim.x);
    __syncthreads();
    for (uint64_t x = grid_tid(); x < n; x += grid_threads())
        gf_store(&acc[x], gf_add(gf_load(&acc[x]), gf_mul_tab(tab, gf_load(&e[x]))));
}

// ------------------------------------------------------ Ligerito sumcheck
// (u0, u2[, y]) over pairs: u0 = sum f0 b0, u2 = sum (f0+f1)(b0+b1),
// y = sum f0 b0 + f1 b1 (the inner product, for OOD introduce).
template <int K>
__global__ __launch_bounds__(kT) void k_lig_msg(const uint4* __restrict__ f, const uint4* __restrict__ b,
                                                uint64_t pairs, uint4* __restrict__ partials) {
    gf256 u0 = gf256_zero(), u2 = gf256_zero(), y1 = gf256_zero();
    for (uint64_t j = grid_tid(); j < pairs; j += grid_threads()) {
        gf128 f0 = gf_load(&f[2 * j]), f1 = gf_load(&f[2 * j + 1]);
        gf128 b0 = gf_load(&b[2 * j]), b1 = gf_load(&b[2 * j + 1]);
        gf256_acc(&u0, gf_mul_wide(f0, b0));
        gf256_acc(&u2, gf_mul_wide(gf_add(f0, f1), gf_add(b0, b1)));
        if (K == 3) gf256_acc(&y1, gf_mul_wide(f1, b1));
    }
    gf128 acc[K];
    acc[0] = gf_reduce(u0);
    acc[1] = gf_reduce(u2);
    if (K == 3) acc[K - 1] = gf_add(gf_reduce(u0), gf_reduce(y1));
    block_xor_store<K>(acc, partials + (size_t)blockIdx.x * K);
}

// f'[j] = f0 + r (f0 + f1), b' likewise; outputs may alias inputs.
__global__ __launch_bounds__(kT) void k_lig_fold(const uint4* f, const uint4* b, uint64_t half, uint64_t r_lo,
                                                 uint64_t r_hi, uint4* fo, uint4* bo) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, gf_from_u64(r_lo, r_hi), threadIdx.x, blockDim.x);
    __syncthreads();
    for (uint64_t j = grid_tid(); j < half; j += grid_threads()) {
        gf128 f0 = gf_load(&f[2 * j]), f1 = gf_load(&f[2 * j + 1]);
        gf128 b0 = gf_load(&b[2 * j]), b1 = gf_load(&b[2 * j + 1]);
        gf_store(&fo[j], gf_add(f0, gf_mul_tab(tab, gf_add(f0, f1))));
        gf_store(&bo[j], gf_add(b0, gf_mul_tab(tab, gf_add(b0, b1))));
    }
}

// Induced basis: out[j] = sum_i aw[i] prod_{k in bits(j)} what[i][k]
// (pcs/ligerito.rs:1968-2070). With j = (j_hi, j_lo) split at bit `lo`,
// the product factors into plo[i][j_lo] * phi[i][j_hi] (aw folded into phi),
// so out is a product of an (2^lo x nq) and an (nq x 2^hi) matrix.
// k_induce_factors: plo[i][j] for j < 2^lo, then phi[i][j] for j < 2^hi.
__global__ void k_induce_factors(const uint4* __restrict__ what /* [nq][log_n] */, const uint4* __restrict__ aw,
                                 uint32_t nq, int log_n, int lo, uint4* __restrict__ plo, uint4* __restrict__ phi) {
    const uint64_t nlo = 1ull << lo, nhi = 1ull << (log_n - lo);
    uint64_t idx = grid_tid();
    if (idx >= (uint64_t)nq * (nlo + nhi)) return;
    bool is_lo = idx < (uint64_t)nq * nlo;
    uint64_t k = is_lo ? idx : idx - (uint64_t)nq * nlo;
    uint64_t n = is_lo ? nlo : nhi;
    uint32_t i = (uint32_t)(k / n);
    uint64_t j = k % n;
    const uint4* w = what + (size_t)i * log_n + (is_lo ? 0 : lo);
    gf128 p = is_lo ? gf_one() : gf_load(&aw[i]);
    for (int b = 0; b < (is_lo ? lo : log_n - lo); b++)
        if ((j >> b) & 1) p = gf_mul(p, gf_load(&w[b]));
    gf_store(is_lo ? &plo[(size_t)i * nlo + j] : &phi[(size_t)i * nhi + j], p);
}

__global__ __launch_bounds__(kT) void k_induce_basis(const uint4* __restrict__ plo, const uint4* __restrict__ phi,
                                                     uint32_t nq, int log_n, int lo, uint4* __restrict__ out) {
    const uint64_t n = 1ull << log_n, nlo = 1ull << lo, nhi = 1ull << (log_n - lo);
    for (uint64_t j = grid_tid(); j < n; j += grid_threads()) {
        uint64_t jl = j & (nlo - 1), jh = j >> lo;
        gf256 acc = gf256_zero();
        for (uint32_t i = 0; i < nq; i++)
            gf256_acc(&acc, gf_mul_wide(gf_load(&plo[(size_t)i * nlo + jl]), gf_load(&phi[(size_t)i * nhi + jh])));
        gf_store(&out[j], gf_reduce(acc));
    }
}

// What_k(q_i) for all queries: s_0(x) = x, s_k = s_{k-1}^2 + s_{k-1}(v_{k-1}) s_{k-1},
// normalized by inv_sks[k] (ligerito.rs:1760-1790).
__global__ void k_induce_what(const uint64_t* __restrict__ queries, uint32_t nq, int log_n,
                              const uint4* __restrict__ sks, const uint4* __restrict__ inv_sks,
                              uint4* __restrict__ what) {
    uint32_t i = grid_tid();
    if (i >= nq) return;
    gf128 s = gf128{{(uint32_t)queries[i], (uint32_t)(queries[i] >> 32), 0u, 0u}};
    for (int k = 0; k < log_n; k++) {
        if (k > 0) s = gf_add(gf_square(s), gf_mul(gf_load(&sks[k - 1]), s));
        gf_store(&what[(size_t)i * log_n + k], gf_mul(s, gf_load(&inv_sks[k])));
    }
}

// sum_x a[x] b[x]
__global__ __launch_bounds__(kT) void k_dot(const uint4* __restrict__ a, const uint4* __restrict__ b, uint64_t n,
                                            uint4* __restrict__ partials) {
    gf256 acc = gf256_zero();
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) gf256_acc(&acc, gf_mul_wide(gf_load(&a[x]), gf_load(&b[x])));
    gf128 v[1] = {gf_reduce(acc)};
    block_xor_store<1>(v, partials + blockIdx.x);
}

// out[x] = a[x] * b[x]
__global__ void k_mul_elem(const uint4* __restrict__ a, const uint4* __restrict__ b, uint64_t n,
                           uint4* __restrict__ out) {
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) gf_store(&out[x], gf_mul(gf_load(&a[x]), gf_load(&b[x])));
}

// v[x] *= b[x] (each thread reads and writes only its own index)
// out[x] = bit x of the flat bit stream (the committed bits as one column)
__global__ void k_bits_to_gf(const uint64_t* __restrict__ bits, uint64_t n, uint4* __restrict__ out) {
    for (uint64_t x = grid_tid(); x < n; x += grid_threads())
        gf_store_nt(&out[x], gf128{{(uint32_t)((bits[x >> 6] >> (x & 63)) & 1ull), 0u, 0u, 0u}});
}

// Inner-product sumcheck rounds J = 0, 1, 2 over a single column of bits
// (src/bitz/sumcheck.rs:146-198, the table m being the committed bits). After
// J folds the table is A_J[k] = sum_c e_c bit(2^J k + c), e the eq table of
// the J challenges, so each round sum splits by c into XORs of the weights
// at the set bits: acc[s 2^J + c] for s = (m0 w0, m1 w1, (m0 + m1)(w0 + w1));
// the host weighs them by e_c. For J > 0 the kernel first folds the weights
// by the last challenge: W_J[k] = W[2k] + r (W[2k] + W[2k + 1]) (written to
// w_out, 2 per pair), so A is never materialised and the round needs no
// variable product. Thread pair i covers bits 2^(J+1) i .. 2^(J+1) (i + 1).
template <int J>
__global__ __launch_bounds__(kT) void k_bits_round(const uint64_t* __restrict__ bits, const uint4* __restrict__ w,
                                                   uint64_t pairs, uint64_t r_lo, uint64_t r_hi,
                                                   uint4* __restrict__ w_out, uint4* __restrict__ partials) {
    constexpr int C = 1 << J, K = 3 * C, NB = 2 * C;
    __shared__ uint4 tab[512];
    if (J > 0) {
        gf_table_build(tab, gf_from_u64(r_lo, r_hi), threadIdx.x, blockDim.x);
        __syncthreads();
    }
    gf128 acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = gf_zero();
    for (uint64_t i = grid_tid(); i < pairs; i += grid_threads()) {
        gf128 w0, w1;
        if (J == 0) {
            w0 = gf_load(&w[2 * i]);
            w1 = gf_load(&w[2 * i + 1]);
        } else {
            const gf128 a0 = gf_load(&w[4 * i]), a1 = gf_load(&w[4 * i + 1]);
            const gf128 b0 = gf_load(&w[4 * i + 2]), b1 = gf_load(&w[4 * i + 3]);
            w0 = gf_add(a0, gf_mul_tab(tab, gf_add(a0, a1)));
            w1 = gf_add(b0, gf_mul_tab(tab, gf_add(b0, b1)));
            gf_store(&w_out[2 * i], w0);
            gf_store(&w_out[2 * i + 1], w1);
        }
        const gf128 d = gf_add(w0, w1);
        const uint64_t bit0 = (uint64_t)NB * i;
        const uint32_t pat = (uint32_t)(bits[bit0 >> 6] >> (bit0 & 63)) & ((1u << NB) - 1);
#pragma unroll
        for (int c = 0; c < C; c++) {
            const uint32_t e = (pat >> c) & 1u, o = (pat >> (C + c)) & 1u;
            const uint32_t me = 0u - e, mo = 0u - o, mx = 0u - (e ^ o);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                acc[c].w[q] ^= w0.w[q] & me;
                acc[C + c].w[q] ^= w1.w[q] & mo;
                acc[2 * C + c].w[q] ^= d.w[q] & mx;
            }
        }
    }
    block_xor_store<K>(acc, partials + (size_t)blockIdx.x * K);
}

// Round 3's start for the single bit column: W_3 = fold of W_2 by r (as in
// k_bits_round) and A_3[k] = T[byte k of the bits], T[v] = sum_{c in v} e_c
// over the eq table e of the first three challenges (built per block).
struct E8 {
    uint4 e[8];
};
__global__ __launch_bounds__(kT) void k_bits_materialize(const uint64_t* __restrict__ bits,
                                                         const uint4* __restrict__ w, uint64_t n_out, uint64_t r_lo,
                                                         uint64_t r_hi, E8 e8, uint4* __restrict__ a_out,
                                                         uint4* __restrict__ w_out) {
    __shared__ uint4 tab[512];
    __shared__ uint4 pt[256];
    gf_table_build(tab, gf_from_u64(r_lo, r_hi), threadIdx.x, blockDim.x);
    for (int v = threadIdx.x; v < 256; v += blockDim.x) {
        uint4 s = make_uint4(0, 0, 0, 0);
        for (int c = 0; c < 8; c++)
            if ((v >> c) & 1) {
                s.x ^= e8.e[c].x;
                s.y ^= e8.e[c].y;
                s.z ^= e8.e[c].z;
                s.w ^= e8.e[c].w;
            }
        pt[v] = s;
    }
    __syncthreads();
    for (uint64_t k = grid_tid(); k < n_out; k += grid_threads()) {
        const gf128 a0 = gf_load(&w[2 * k]), a1 = gf_load(&w[2 * k + 1]);
        gf_store(&w_out[k], gf_add(a0, gf_mul_tab(tab, gf_add(a0, a1))));
        const uint32_t byte = (uint32_t)(bits[k >> 3] >> ((k & 7) * 8)) & 255u;
        a_out[k] = pt[byte];
    }
}

__global__ void k_mul_inplace(uint4* v, const uint4* __restrict__ b, uint64_t n) {
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) gf_store(&v[x], gf_mul(gf_load(&v[x]), gf_load(&b[x])));
}

// out[x] = (a[x] + 1) * b[x]
__global__ void k_mul_elem_plus1(const uint4* __restrict__ a, const uint4* __restrict__ b, uint64_t n,
                                 uint4* __restrict__ out) {
    for (uint64_t x = grid_tid(); x < n; x += grid_threads())
        gf_store(&out[x], gf_mul(gf_add(gf_load(&a[x]), gf_one()), gf_load(&b[x])));
}

// out[x] = a[x] * c
__global__ __launch_bounds__(kT) void k_scale(const uint4* __restrict__ a, uint64_t n, uint64_t c_lo,
                                              uint64_t c_hi, uint4* __restrict__ out) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, gf_from_u64(c_lo, c_hi), threadIdx.x, blockDim.x);
    __syncthreads();
    for (uint64_t x = grid_tid(); x < n; x += grid_threads()) gf_store(&out[x], gf_mul_tab(tab, gf_load(&a[x])));
}

// ----------------------------------------------------------- proof of work
// Smallest nonce in [base, base + span) with >= bits leading zero bits of
// BLAKE3(prefix16 || seed16 || nonce_le) (src/bitz/pcs.rs:1252-1298);
// result via atomicMin on a u64 (stable: the minimum is unique).
struct PowPrefix {
    uint32_t w[8];  // prefix16 || seed16 as little-endian words
};
__global__ __launch_bounds__(kT) void k_pow_scan(PowPrefix pre, uint64_t base, uint64_t span, uint32_t bits,
                                                 unsigned long long* __restrict__ best) {
    uint64_t i = grid_tid();
    if (i >= span) return;
    uint64_t nonce = base + i;
    if ((unsigned long long)nonce >= *((volatile unsigned long long*)best)) return;
    uint32_t m[16] = {pre.w[0], pre.w[1], pre.w[2], pre.w[3], pre.w[4], pre.w[5], pre.w[6], pre.w[7],
                      (uint32_t)nonce, (uint32_t)(nonce >> 32), 0, 0, 0, 0, 0, 0};
    uint32_t cv[8];
    b3::iv_init(cv);
    b3::compress(cv, m, 0, 40, b3::CHUNK_START | b3::CHUNK_END | b3::ROOT);
    // leading zero bits over bytes in order (byte 0 = low byte of cv[0])
    uint32_t zeros = 0;
    for (int wd = 0; wd < 8; wd++) {
        uint32_t x = cv[wd];
        // bytes little-endian within the word; count MSB-first per byte
        bool done = false;
        for (int by = 0; by < 4; by++) {
            uint32_t byte = (x >> (8 * by)) & 255u;
            uint32_t z = byte ? (__clz(byte) - 24) : 8;
            zeros += z;
            if (z != 8) {
                done = true;
                break;
            }
        }
        if (done || zeros >= bits) break;
    }
    if (zeros >= bits) atomicMin(best, (unsigned long long)nonce);
}

__global__ void k_pow_init(unsigned long long* best) {
    if (threadIdx.x == 0) *best = ~0ull;
}

// the smallest valid nonce so far (all ones: none) to the mailbox; with
// reset, the cell goes back to all ones for the next search
__global__ void k_pow_report(unsigned long long* best, uint4* mbox_vals, uint32_t* mbox_flag, uint32_t seq,
                             int reset) {
    if (threadIdx.x == 0) {
        const unsigned long long b = *(volatile unsigned long long*)best;
        if (reset) *best = ~0ull;
        mbox_vals[0] = make_uint4((uint32_t)b, (uint32_t)(b >> 32), 0, 0);
        __threadfence_system();
        *(volatile uint32_t*)mbox_flag = seq;
    }
}

// out[j] = XOR over b < nblocks of partials[b * width + j]
// out[j] = XOR over b < nblocks of partials[b][j]. Block: 16 columns x 16
// slices of the b range (loads of 16 consecutive columns coalesce), the
// slices combined in LDS.
__global__ __launch_bounds__(256) void k_reduce_columns(const uint4* __restrict__ partials, uint32_t nblocks,
                                                        uint32_t width, uint4* __restrict__ out) {
    __shared__ uint4 red[16][16];
    const uint32_t cj = threadIdx.x & 15, rs = threadIdx.x >> 4;
    const uint32_t j = blockIdx.x * 16 + cj;
    uint4 a = make_uint4(0, 0, 0, 0);
    if (j < width)
        for (uint32_t b = rs; b < nblocks; b += 16) {
            const uint4 q = partials[(size_t)b * width + j];
            a.x ^= q.x, a.y ^= q.y, a.z ^= q.z, a.w ^= q.w;
        }
    red[rs][cj] = a;
    __syncthreads();
    if (rs == 0 && j < width) {
        for (int k = 1; k < 16; k++) {
            const uint4 q = red[k][cj];
            a.x ^= q.x, a.y ^= q.y, a.z ^= q.z, a.w ^= q.w;
        }
        out[j] = a;
    }
}

// gather rows[q] (lanes elements each) for sorted query positions
__global__ void k_gather_rows(const uint4* __restrict__ cw, const uint64_t* __restrict__ q, uint32_t nq,
                              uint32_t lanes, uint4* __restrict__ out) {
    uint64_t idx = grid_tid();
    if (idx >= (uint64_t)nq * lanes) return;
    uint64_t i = idx / lanes, l = idx % lanes;
    out[idx] = cw[q[i] * lanes + l];
}

__global__ void k_gather_nodes(const uint4* __restrict__ tree, const uint64_t* __restrict__ idx, uint32_t n,
                               uint4* __restrict__ out) {
    uint32_t i = grid_tid();
    if (i >= n) return;
    out[2 * i] = tree[2 * idx[i]];
    out[2 * i + 1] = tree[2 * idx[i] + 1];
}

// ------------------------------------------------ device Fiat-Shamir steps

// One GKR round (gkr.rs:301-310 prover_message / verifier_message) on one
// wavefront, after the block partials were reduced into sums[0..2):
// message = factor * corr * sums, absorb it,
// squeeze the challenge, factor *= eq_factor(challenge, z)
// (src/bitz/mod.rs:328-331). The hasher state is worked on in LDS.
constexpr uint32_t kFsThreads = 64;
__global__ __launch_bounds__(kFsThreads) void k_gkr_fs_round(const uint4* __restrict__ sums, bitz::B3State* st,
                                                             uint4* factor, const uint4* internal, const uint4* corr,
                                                             uint4* rho, uint4* msgs, uint32_t round) {
    __shared__ bitz::B3State sst;
    __shared__ __align__(16) uint8_t stage[64];
    __shared__ __align__(16) uint8_t msg[32];
    __shared__ __align__(16) uint8_t rb[16];
    constexpr uint32_t kWords = sizeof(bitz::B3State) / 4;
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < kWords; i += blockDim.x) ((uint32_t*)&sst)[i] = ((const uint32_t*)st)[i];
    __syncthreads();
    gf128 f = gf_mul(gf_load(factor), gf_load(&corr[round]));
    gf128 m0 = gf_mul(f, gf_load(&sums[0])), m1 = gf_mul(f, gf_load(&sums[1]));
    if (lane == 0) {
        for (int i = 0; i < 4; i++) {
            ((uint32_t*)msg)[i] = m0.w[i];
            ((uint32_t*)msg)[4 + i] = m1.w[i];
        }
        gf_store(&msgs[2 * round], m0);
        gf_store(&msgs[2 * round + 1], m1);
    }
    __syncthreads();
    b3s::absorb_slice(&sst, msg, 32, stage);
    b3s::challenge16(&sst, rb, stage);
    gf128 r = gf_load((const uint4*)rb), z = gf_load(&internal[round]);
    gf128 e = gf_add(gf_mul(r, z), gf_mul(gf_add(r, gf_one()), gf_add(z, gf_one())));
    gf128 fnew = gf_mul(gf_load(factor), e);
    __syncthreads();
    if (lane == 0) {
        gf_store(&rho[round], r);
        gf_store(factor, fnew);
    }
    for (uint32_t i = lane; i < kWords; i += blockDim.x) ((uint32_t*)st)[i] = ((const uint32_t*)&sst)[i];
}

__global__ __launch_bounds__(kFsThreads) void k_transcript_challenges(bitz::B3State* st, const uint8_t* __restrict__ msgs,
                                                                      uint32_t n, uint32_t msg_len,
                                                                      uint8_t* __restrict__ out) {
    __shared__ bitz::B3State sst;
    __shared__ __align__(16) uint8_t stage[1600];
    __shared__ __align__(16) uint8_t data[1536];
    __shared__ __align__(16) uint8_t rb[16];
    constexpr uint32_t kWords = sizeof(bitz::B3State) / 4;
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < kWords; i += blockDim.x) ((uint32_t*)&sst)[i] = ((const uint32_t*)st)[i];
    __syncthreads();
    for (uint32_t m = 0; m < n; m++) {
        for (uint32_t i = lane; i < msg_len; i += blockDim.x) data[i] = msgs[(size_t)m * msg_len + i];
        __syncthreads();
        b3s::absorb_slice(&sst, data, msg_len, stage);
        b3s::challenge16(&sst, rb, stage);
        if (lane < 16) out[16 * m + lane] = rb[lane];
        __syncthreads();
    }
}

// l[0..n) then r[0..n) into host-mapped memory, then the sequence flag the
// host polls on: every block fences its stores at system scope and counts
// itself in `done` (device memory, zero between launches); the last block to
// arrive resets the count and writes the flag.
__global__ __launch_bounds__(kT) void k_pair_to_host(const uint4* __restrict__ l, const uint4* __restrict__ r,
                                                     uint64_t n, uint4* dst, uint32_t* done, uint32_t* flag,
                                                     uint32_t seq) {
    for (uint64_t i = grid_tid(); i < n; i += grid_threads()) {
        dst[i] = l[i];
        dst[n + i] = r[i];
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        if (atomicAdd(done, 1u) == gridDim.x - 1) {
            atomicExch(done, 0u);
            __threadfence_system();
            *(volatile uint32_t*)flag = seq;
        }
    }
}

}  // namespace

// ================================================================ host side

int Scratch::ensure(size_t partial_bytes) {
    if (!partials.alloc(partial_bytes)) return BITZ_ERR_ALLOC;
    if (!result.alloc(64 * 16)) return BITZ_ERR_ALLOC;
    if (!mbox) {
        if (hipHostMalloc(&mbox, 4096, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
            hipHostGetDevicePointer(&mbox_dev, mbox, 0) != hipSuccess) {
            if (mbox) (void)hipHostFree(mbox);
            mbox = mbox_dev = nullptr;
            return BITZ_ERR_ALLOC;
        }
        memset(mbox, 0, 4096);
    }
    if (!host) {
        if (hipHostMalloc(&host, 64 * 16, hipHostMallocDefault) != hipSuccess) {
            host = nullptr;
            return BITZ_ERR_ALLOC;
        }
    }
    return BITZ_OK;
}
Scratch::~Scratch() {
    if (host) (void)hipHostFree(host);
    if (tail) (void)hipHostFree(tail);
    if (mbox) (void)hipHostFree(mbox);
    if (stage) (void)hipHostFree(stage);
    if (up) (void)hipHostFree(up);
}

int upload_async(bitz_ctx* ctx, const void* host, size_t bytes, void* dev) {
    if (bytes == 0) return BITZ_OK;
    Scratch& sc = ctx->scratch;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (need > sc.up_bytes) {
        // grow (at least 1 MiB): earlier copies from the old ring finish first
        BITZ_HIP(hipStreamSynchronize(ctx->stream));
        if (sc.up) (void)hipHostFree(sc.up);
        sc.up = nullptr;
        sc.up_bytes = sc.up_off = 0;
        const size_t cap = std::max<size_t>(need, 1u << 20);
        if (hipHostMalloc(&sc.up, cap, hipHostMallocDefault) != hipSuccess) {
            sc.up = nullptr;
            set_error("upload ring: hipHostMalloc failed");
            return BITZ_ERR_ALLOC;
        }
        sc.up_bytes = cap;
    }
    if (sc.up_off + need > sc.up_bytes) {
        BITZ_HIP(hipStreamSynchronize(ctx->stream));  // every copy out of the ring is done
        sc.up_off = 0;
    }
    uint8_t* slot = (uint8_t*)sc.up + sc.up_off;
    memcpy(slot, host, bytes);
    sc.up_off += need;
    BITZ_HIP(hipMemcpyAsync(dev, slot, bytes, hipMemcpyHostToDevice, ctx->stream));
    return BITZ_OK;
}

int Scratch::ensure_stage(size_t bytes) {
    if (bytes <= stage_bytes) return BITZ_OK;
    if (stage) (void)hipHostFree(stage);
    stage = nullptr;
    stage_bytes = 0;
    if (hipHostMalloc(&stage, bytes, hipHostMallocDefault) != hipSuccess) {
        stage = nullptr;
        set_error("host staging: hipHostMalloc failed");
        return BITZ_ERR_ALLOC;
    }
    stage_bytes = bytes;
    return BITZ_OK;
}

static void split16(const uint8_t* p, uint64_t* lo, uint64_t* hi) {
    if (!p) {  // the challenge comes from device memory instead
        *lo = *hi = 0;
        return;
    }
    memcpy(lo, p, 8);
    memcpy(hi, p + 8, 8);
}

int sums_finish(bitz_ctx* ctx, unsigned nblocks, int K, void* out_dev) {
    if (ctx->defer_sums) {  // the caller reduces the block partials itself
        ctx->deferred_nb = nblocks;
        return BITZ_OK;
    }
    Scratch& sc = ctx->scratch;
    uint4* parts = (uint4*)sc.partials.ptr;
    if (!out_dev && sc.mbox_dev) {
        // the host reads these sums next: deliver them to the mailbox
        uint32_t seq = ++sc.seq;
        uint32_t* flag = (uint32_t*)sc.mbox_dev;
        uint4* vals = (uint4*)((uint8_t*)sc.mbox_dev + 64);
        uint4* dst = (uint4*)sc.result.ptr;
        switch (K) {
            case 1: hipLaunchKernelGGL(k_xor_partials_mbox<1>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst, vals, flag, seq); break;
            case 2: hipLaunchKernelGGL(k_xor_partials_mbox<2>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst, vals, flag, seq); break;
            case 3: hipLaunchKernelGGL(k_xor_partials_mbox<3>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst, vals, flag, seq); break;
            case 6: hipLaunchKernelGGL(k_xor_partials_mbox<6>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst, vals, flag, seq); break;
            case 12: hipLaunchKernelGGL(k_xor_partials_mbox<12>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst, vals, flag, seq); break;
            default: return BITZ_ERR_ARG;
        }
        BITZ_LAUNCH_CHECK();
        sc.mbox_pending = true;
        return BITZ_OK;
    }
    uint4* dst = out_dev ? (uint4*)out_dev : (uint4*)sc.result.ptr;
    switch (K) {
        case 1: hipLaunchKernelGGL(k_xor_partials<1>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst); break;
        case 2: hipLaunchKernelGGL(k_xor_partials<2>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst); break;
        case 3: hipLaunchKernelGGL(k_xor_partials<3>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst); break;
        case 6: hipLaunchKernelGGL(k_xor_partials<6>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst); break;
        case 12: hipLaunchKernelGGL(k_xor_partials<12>, dim3(1), dim3(kT), 0, ctx->stream, parts, nblocks, dst); break;
        default: return BITZ_ERR_ARG;
    }
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

// Spins on the mailbox flag until it reads the pending sequence number;
// every 1024 polls ask whether the stream has drained, so a launch that
// never delivers ends in an error, not a hang.
static int wait_mbox(bitz_ctx* ctx, const char* what) {
    Scratch& sc = ctx->scratch;
    volatile uint32_t* flag = (volatile uint32_t*)sc.mbox;
    for (uint64_t spin = 1;; spin++) {
        if (*flag == sc.seq) break;
        if (spin > (1u << 16)) std::this_thread::yield();  // a long kernel: share the core
        if ((spin & 1023) == 0) {
            hipError_t q = hipStreamQuery(ctx->stream);
            if (q == hipSuccess) {
                if (*flag == sc.seq) break;
                set_error(std::string(what) + ": the stream finished without delivering its values");
                return BITZ_ERR_DEVICE;
            }
            if (q != hipErrorNotReady) {
                set_error(std::string(what) + ": " + hipGetErrorString(q));
                return BITZ_ERR_DEVICE;
            }
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    sc.mbox_pending = false;
    return BITZ_OK;
}

int read_result(bitz_ctx* ctx, int K, uint8_t* host_out) {
    Scratch& sc = ctx->scratch;
    if (sc.mbox_pending) {
        if (int rc = wait_mbox(ctx, "read_result")) return rc;
        memcpy(host_out, (const uint8_t*)sc.mbox + 64, K * 16);
        return BITZ_OK;
    }
    BITZ_HIP(hipMemcpyAsync(sc.host, sc.result.ptr, K * 16, hipMemcpyDeviceToHost, ctx->stream));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(host_out, sc.host, K * 16);
    return BITZ_OK;
}

int download_pair_mapped(bitz_ctx* ctx, const void* l, const void* r, uint64_t n, void** host_out) {
    Scratch& sc = ctx->scratch;
    if (int rc = sc.ensure(16)) return rc;
    BITZ_REQUIRE(!sc.mbox_pending, BITZ_ERR_ARG, "download_pair_mapped: a mailbox result is still pending");
    const size_t bytes = std::max<size_t>(2 * n * 16, 1u << 20);  // at least 1 MiB: grown rarely
    if (2 * n * 16 > sc.tail_bytes) {
        BITZ_HIP(hipStreamSynchronize(ctx->stream));  // no kernel still writes the old buffer
        if (sc.tail) (void)hipHostFree(sc.tail);
        sc.tail = sc.tail_dev = nullptr;
        sc.tail_bytes = 0;
        if (hipHostMalloc(&sc.tail, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
            hipHostGetDevicePointer(&sc.tail_dev, sc.tail, 0) != hipSuccess) {
            if (sc.tail) (void)hipHostFree(sc.tail);
            sc.tail = sc.tail_dev = nullptr;
            set_error("download_pair_mapped: hipHostMalloc failed");
            return BITZ_ERR_ALLOC;
        }
        sc.tail_bytes = bytes;
    }
    if (!sc.tail_done.ptr) {
        if (!sc.tail_done.alloc(256)) return BITZ_ERR_ALLOC;
        BITZ_HIP(hipMemsetAsync(sc.tail_done.ptr, 0, 256, ctx->stream));
    }
    const uint32_t seq = ++sc.seq;
    hipLaunchKernelGGL(k_pair_to_host, dim3(blocks_for(n, 2, 64)), dim3(kT), 0, ctx->stream, (const uint4*)l,
                       (const uint4*)r, n, (uint4*)sc.tail_dev, (uint32_t*)sc.tail_done.ptr, (uint32_t*)sc.mbox_dev,
                       seq);
    BITZ_LAUNCH_CHECK();
    sc.mbox_pending = true;
    if (int rc = wait_mbox(ctx, "download_pair_mapped")) return rc;
    *host_out = sc.tail;
    return BITZ_OK;
}

// out = lo (2^nlo) (x) hi (2^nhi): table products for long runs of one hi factor
static int eq_combine(bitz_ctx* ctx, const uint4* lo, int nlo, const uint4* hi, int nhi, uint4* out) {
    const uint64_t total = 1ull << (nlo + nhi);
    if (nlo >= 10) {
        const int pbl = std::min(nlo, 12);
        hipLaunchKernelGGL(k_eq_combine_tab, dim3((unsigned)(total >> pbl)), dim3(kT), 0, ctx->stream, lo, nlo, hi,
                           pbl, out);
    } else {
        hipLaunchKernelGGL(k_eq_combine, dim3(blocks_for(total, 4)), dim3(kT), 0, ctx->stream, lo, nlo, hi, total,
                           out);
    }
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int eq_table(bitz_ctx* ctx, const void* r_dev, int n, void* out) {
    BITZ_REQUIRE(n >= 0 && n <= 40, BITZ_ERR_SHAPE, "eq_table: n out of range");
    if (n <= 12) {
        hipLaunchKernelGGL(k_eq_small, dim3(1), dim3(kT), 0, ctx->stream, (const uint4*)r_dev, n, (uint4*)out);
        BITZ_LAUNCH_CHECK();
        return BITZ_OK;
    }
    int nlo = n / 2, nhi = n - nlo;
    int rc = ctx->scratch.ensure(4096 * 16 * 3);
    if (rc) return rc;
    if (!ctx->scratch.eq_lo.alloc((16ull << nlo) + (16ull << nhi))) return BITZ_ERR_ALLOC;
    uint4* lo = (uint4*)ctx->scratch.eq_lo.ptr;
    uint4* hi = lo + (1ull << nlo);
    BITZ_REQUIRE(nhi <= 14, BITZ_ERR_SHAPE, "eq_table: n too large");
    hipLaunchKernelGGL(k_eq_small, dim3(1), dim3(kT), 0, ctx->stream, (const uint4*)r_dev, nlo, lo);
    hipLaunchKernelGGL(k_eq_small, dim3(1), dim3(kT), 0, ctx->stream, (const uint4*)r_dev + nlo, nhi, hi);
    BITZ_LAUNCH_CHECK();
    return eq_combine(ctx, lo, nlo, hi, nhi, (uint4*)out);
}

int fold_columns(bitz_ctx* ctx, const uint64_t* rows, int t, int s, const void* w, void* out) {
    BITZ_REQUIRE(t >= 6 && s >= 0, BITZ_ERR_SHAPE, "fold_columns: bad shape");
    uint64_t words = 1ull << (t - 6), cols = 1ull << s;
    const bool tables = words >= 16 && cols % 256 == 0;  // the nibble-table kernel: chunks of 16 words
    uint32_t chunk = tables ? 16 : words >= 64 ? 64 : (uint32_t)words;
    uint64_t chunks = words / chunk;
    if (!ctx->scratch.fold_partial.alloc(chunks * cols * 16)) return BITZ_ERR_ALLOC;
    uint4* part = (uint4*)ctx->scratch.fold_partial.ptr;
    uint64_t items = cols * chunks;
    if (tables)
        hipLaunchKernelGGL(k_fold_columns_partial_lds, dim3((unsigned)(items / 256)), dim3(256), 0, ctx->stream, rows,
                           t, s, (const uint4*)w, part);
    else
        hipLaunchKernelGGL(k_fold_columns_partial, dim3((unsigned)((items + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                           rows, t, s, (const uint4*)w, chunk, part);
    hipLaunchKernelGGL(k_fold_columns_sum, dim3((unsigned)((cols + kT - 1) / kT)), dim3(kT), 0, ctx->stream, part,
                       chunks, cols, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int pow_images(bitz_ctx* ctx, const uint8_t g[16], const void* exps, uint64_t n, void* out) {
    if (!ctx->scratch.comb.alloc(16 * 256 * 16)) return BITZ_ERR_ALLOC;
    uint64_t g_lo, g_hi;
    memcpy(&g_lo, g, 8);
    memcpy(&g_hi, g + 8, 8);
    if (memcmp(ctx->scratch.comb_base, g, 16) != 0 || !ctx->scratch.comb_ready) {
        hipLaunchKernelGGL(k_comb_rows, dim3(1), dim3(64), 0, ctx->stream, (uint4*)ctx->scratch.comb.ptr, g_lo, g_hi);
        BITZ_LAUNCH_CHECK();
        memcpy(ctx->scratch.comb_base, g, 16);
        ctx->scratch.comb_ready = true;
    }
    if (n == 0) return BITZ_OK;
    hipLaunchKernelGGL(k_pow_comb, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)ctx->scratch.comb.ptr, (const uint4*)exps, n, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int pack_columns(bitz_ctx* ctx, const uint64_t* rows, int t, int s, uint64_t* packed) {
    BITZ_REQUIRE(t >= 6, BITZ_ERR_SHAPE, "pack_columns: t < 6");
    const uint64_t groups = ((1ull << s) + 63) / 64, n = (groups << (t - 6)) * 64;  // one wave per word
    TimedLaunch tl(ctx, "pack_columns", 2.0 * 8.0 * (double)(groups << t), 0.0);
    hipLaunchKernelGGL(k_pack_columns, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, rows, t, s,
                       packed);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}


int gkr_round_sums(bitz_ctx* ctx, const void* l, const void* r, uint64_t h, const void* eq_c, int s_c,
                   const void* eq_y, int send_one, void* out_dev) {
    uint32_t R = round_chunk_r(ctx->tune, h, s_c);
    unsigned nb = round_blocks(ctx->tune, h, R);
    int rc = ctx->scratch.ensure((size_t)nb * 2 * 16);
    if (rc) return rc;
    TimedLaunch tl(ctx, "gkr_round_sums", 4.0 * 16.0 * h, 4.0 * h * kGfMulOps);
    DenseSrc src{(const uint4*)l, (const uint4*)r};
    hipLaunchKernelGGL(k_gkr_round_sums<DenseSrc>, dim3(nb), dim3(kT), 0, ctx->stream, src, h,
                       (const uint4*)eq_c, s_c, (const uint4*)eq_y, send_one, R, (uint4*)ctx->scratch.partials.ptr, nullptr, 0u);
    BITZ_LAUNCH_CHECK();
    return sums_finish(ctx, nb, 2, out_dev);
}

// One size for the leaf-layer tables, so the buffer is never reallocated
// between launches that share it: round 0 qtab (2^(t+3) entries) or ftab
// (2^(t+1)) or the level-1 row products (2^(t-1)) at the start; the round-1
// q1tab (2^(t+6)) or gtab (2^(t+2)) after 2^(t+1) entries.
static size_t leaf_tab_bytes(int t) { return (size_t)16 * ((1ull << (t + 1)) + (1ull << (t + 6))); }
static uint4* leaf_tab_hi(bitz_ctx* ctx, int t) { return (uint4*)ctx->scratch.leaf_tab.ptr + (1ull << (t + 1)); }

// threads of a gather pattern-sums launch: (column, ylane) pairs, about 2^20
static uint64_t pattern_ylanes(int s, uint64_t ny) {
    return s >= 20 ? 1 : std::min<uint64_t>(ny, (1ull << 20) >> s);
}

// A K-bit pattern-sums launch over ny = 2^(K/4 - 1) Y4... rows (row stride
// ny): off the row words nr (built into the scratch buffer when not given)
// for t >= 4, one block per row-word row; the per-(column, row) gather
// kernel for smaller t. Returns the block count.
template <int K>
static int launch_pattern_sums(bitz_ctx* ctx, const uint64_t* packed, const void* nr, int t, int s, uint64_t ny,
                               const uint4* tab, const void* eq_c, unsigned* nb_out) {
    if (t >= 4) {
        const uint64_t y4 = 1ull << (t - 4);
        BITZ_REQUIRE(ny == (K == 8 ? 2 * y4 : 4 * y4) && y4 <= 0x7fffffffull, BITZ_ERR_SHAPE,
                     "pattern sums: rows do not match the row words");
        if (!nr) {
            if (!ctx->scratch.nib_rows.alloc((y4 << s) * 2)) return BITZ_ERR_ALLOC;
            int rc = gkr_nibble_rows(ctx, packed, t, s, ctx->scratch.nib_rows.ptr);
            if (rc) return rc;
            nr = ctx->scratch.nib_rows.ptr;
        }
        const unsigned nb = (unsigned)y4;
        int rc = ctx->scratch.ensure((size_t)nb * 2 * 16);
        if (rc) return rc;
        hipLaunchKernelGGL(k_byte_bucket_sums<K>, dim3(nb), dim3(kT), 0, ctx->stream, (const uint16_t*)nr, t, s, tab,
                           (const uint4*)eq_c, (uint4*)ctx->scratch.partials.ptr);
        BITZ_LAUNCH_CHECK();
        *nb_out = nb;
        return BITZ_OK;
    }
    const uint64_t ylanes = pattern_ylanes(s, ny);
    const unsigned nb = (unsigned)(((1ull << s) * ylanes + kT - 1) / kT);
    int rc = ctx->scratch.ensure((size_t)nb * 2 * 16);
    if (rc) return rc;
    hipLaunchKernelGGL(k_leaf_pattern_sums<K>, dim3(nb), dim3(kT), 0, ctx->stream, packed, t, s, ny, ny, tab,
                       (const uint4*)eq_c, ylanes, (uint4*)ctx->scratch.partials.ptr);
    BITZ_LAUNCH_CHECK();
    *nb_out = nb;
    return BITZ_OK;
}

int gkr_premul_round_sums(bitz_ctx* ctx, const void* l, void* r, uint64_t h, const void* eq_c, int s,
                          const void* eq_y, int send_one, void* out_dev) {
    uint32_t R = round_chunk_r(ctx->tune, h, s);
    unsigned nb = round_blocks(ctx->tune, h, R);
    int rc = ctx->scratch.ensure((size_t)nb * 2 * 16);
    if (rc) return rc;
    TimedLaunch tl(ctx, "gkr_premul_round_sums", 6.0 * 16.0 * h, 4.0 * h * kGfMulOps);
    PremulSrc src{(const uint4*)l, (uint4*)r, (const uint4*)eq_c, (1ull << s) - 1};
    hipLaunchKernelGGL((k_gkr_round_sums<PremulSrc, false>), dim3(nb), dim3(kT), 0, ctx->stream, src, h,
                       (const uint4*)eq_c, s, (const uint4*)eq_y, send_one, R, (uint4*)ctx->scratch.partials.ptr,
                       nullptr, 0u);
    BITZ_LAUNCH_CHECK();
    return sums_finish(ctx, nb, 2, out_dev);
}

int gkr_round_sums_rowonly(bitz_ctx* ctx, const void* l, const void* r, uint64_t h, int s, const void* eq_y,
                           int send_one, void* out_dev, const void* eq_lo, int lo_bits) {
    BITZ_REQUIRE(!eq_lo || (lo_bits >= 0 && lo_bits <= 8), BITZ_ERR_SHAPE, "gkr_round_sums_rowonly: lo_bits > 8");
    uint32_t R = round_chunk_r(ctx->tune, h, s);
    unsigned nb = round_blocks(ctx->tune, h, R);
    int rc = ctx->scratch.ensure((size_t)nb * 2 * 16);
    if (rc) return rc;
    TimedLaunch tl(ctx, "gkr_round_sums_rowonly", 4.0 * 16.0 * h, 2.0 * h * kGfMulOps);
    DenseSrc src{(const uint4*)l, (const uint4*)r};
    hipLaunchKernelGGL((k_gkr_round_sums<DenseSrc, false>), dim3(nb), dim3(kT), 0, ctx->stream, src, h,
                       (const uint4*)nullptr, s, (const uint4*)eq_y, send_one, R, (uint4*)ctx->scratch.partials.ptr,
                       (const uint4*)eq_lo, (1u << lo_bits) - 1);
    BITZ_LAUNCH_CHECK();
    return sums_finish(ctx, nb, 2, out_dev);
}

int tensor_args(bitz_ctx* ctx, const hgf* a, const hgf* b, int n, void* out) {
    BITZ_REQUIRE(n >= 0 && n <= kMaxTensorN, BITZ_ERR_SHAPE, "tensor_args: n out of range");
    TensorArgs c;
    for (int i = 0; i < n; i++) {
        c.a[i] = make_uint4((uint32_t)a[i].lo, (uint32_t)(a[i].lo >> 32), (uint32_t)a[i].hi, (uint32_t)(a[i].hi >> 32));
        c.b[i] = make_uint4((uint32_t)b[i].lo, (uint32_t)(b[i].lo >> 32), (uint32_t)b[i].hi, (uint32_t)(b[i].hi >> 32));
    }
    const unsigned blocks = n > 8 ? 1u << (n - 8) : 1u;
    hipLaunchKernelGGL(k_tensor<ArgCoords>, dim3(blocks), dim3(256), 0, ctx->stream, ArgCoords{c}, n, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int eq_args(bitz_ctx* ctx, const hgf* r, int n, void* out) {
    BITZ_REQUIRE(n >= 0 && n <= kMaxTensorN, BITZ_ERR_SHAPE, "eq_args: n out of range");
    hgf a[kMaxTensorN];
    for (int i = 0; i < n; i++) a[i] = hgf_add(r[i], hgf{1, 0});
    if (n <= 14) return tensor_args(ctx, a, r, n, out);
    // large tables: the tables of the low and high halves of the
    // coordinates, then one product per entry (a pass at memory speed)
    const int nlo = n / 2, nhi = n - nlo;
    if (!ctx->scratch.eq_lo.alloc((16ull << nlo) + (16ull << nhi))) return BITZ_ERR_ALLOC;
    uint4* lo = (uint4*)ctx->scratch.eq_lo.ptr;
    uint4* hi = lo + (1ull << nlo);
    int rc = tensor_args(ctx, a, r, nlo, lo);
    if (rc) return rc;
    if ((rc = tensor_args(ctx, a + nlo, r + nlo, nhi, hi))) return rc;
    return eq_combine(ctx, lo, nlo, hi, nhi, (uint4*)out);
}

int tensor_args_multi(bitz_ctx* ctx, int njobs, const hgf* const* a, const hgf* const* b, const int* n,
                      void* const* out) {
    BITZ_REQUIRE(njobs >= 1 && njobs <= kTensorJobs, BITZ_ERR_ARG, "tensor_args_multi: 1..4 tables");
    TensorJobs J{};
    uint32_t blocks = 0;
    for (int j = 0; j < njobs; j++) {
        BITZ_REQUIRE(n[j] >= 0 && n[j] <= kTensorJobN && out[j], BITZ_ERR_SHAPE, "tensor_args_multi: bad table");
        for (int i = 0; i < n[j]; i++) {
            J.a[j][i] = make_uint4((uint32_t)a[j][i].lo, (uint32_t)(a[j][i].lo >> 32), (uint32_t)a[j][i].hi,
                                   (uint32_t)(a[j][i].hi >> 32));
            J.b[j][i] = make_uint4((uint32_t)b[j][i].lo, (uint32_t)(b[j][i].lo >> 32), (uint32_t)b[j][i].hi,
                                   (uint32_t)(b[j][i].hi >> 32));
        }
        J.out[j] = (uint4*)out[j];
        J.n[j] = n[j];
        J.first[j] = blocks;
        blocks += n[j] > 8 ? 1u << (n[j] - 8) : 1u;
    }
    J.first[njobs] = blocks;
    J.njobs = njobs;
    hipLaunchKernelGGL(k_tensor_jobs, dim3(blocks), dim3(256), 0, ctx->stream, J);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int tensor_table(bitz_ctx* ctx, const void* a, const void* b, int n, void* out) {
    BITZ_REQUIRE(n >= 0 && n <= 14, BITZ_ERR_SHAPE, "tensor_table: n out of range");
    const unsigned blocks = n > 8 ? 1u << (n - 8) : 1u;
    hipLaunchKernelGGL(k_tensor<DevCoords>, dim3(blocks), dim3(256), 0, ctx->stream,
                       DevCoords{(const uint4*)a, (const uint4*)b}, n, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_fold_round_sums(bitz_ctx* ctx, void* l, void* r, uint64_t h, const uint8_t rho[16], const void* eq_c,
                        int s_c, const void* eq_y, int send_one, void* out_dev, const void* rho_dev,
                        const void* eq_lo, int lo_bits) {
    BITZ_REQUIRE(!eq_lo || (!eq_c && lo_bits >= 0 && lo_bits <= 8), BITZ_ERR_SHAPE,
                 "gkr_fold_round_sums: eq_lo needs eq_c null and lo_bits <= 8");
    uint32_t R = round_chunk_r(ctx->tune, h, s_c);
    unsigned nb = round_blocks(ctx->tune, h, R);
    int rc = ctx->scratch.ensure((size_t)nb * 2 * 16);
    if (rc) return rc;
    uint64_t lo, hi;
    split16(rho, &lo, &hi);
    FoldSrc src{(uint4*)l, (uint4*)r, lo, hi, (const uint4*)rho_dev};
    if (eq_c) {
        TimedLaunch tl(ctx, "gkr_fold_round_sums", 12.0 * 16.0 * h, h * (4.0 * kGfMulOps + 4.0 * kGfTabMulOps));
        hipLaunchKernelGGL(k_gkr_round_sums<FoldSrc>, dim3(nb), dim3(kT), 0, ctx->stream, src, h, (const uint4*)eq_c,
                           s_c, (const uint4*)eq_y, send_one, R, (uint4*)ctx->scratch.partials.ptr, nullptr, 0u);
    } else {
        TimedLaunch tl(ctx, "gkr_fold_round_sums_rowonly", 12.0 * 16.0 * h, h * (2.0 * kGfMulOps + 4.0 * kGfTabMulOps));
        hipLaunchKernelGGL((k_gkr_round_sums<FoldSrc, false>), dim3(nb), dim3(kT), 0, ctx->stream, src, h, nullptr,
                           s_c, (const uint4*)eq_y, send_one, R, (uint4*)ctx->scratch.partials.ptr,
                           (const uint4*)eq_lo, (1u << lo_bits) - 1);
    }
    BITZ_LAUNCH_CHECK();
    return sums_finish(ctx, nb, 2, out_dev);
}

int gkr_leaf_round_sums(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const void* eq_c,
                        int s_c, const void* eq_y, int send_one, void* out_dev, const void* nr) {
    uint64_t h = 1ull << (t + s - 2);
    if (s_c == s && t >= 2) {
        // weights eq_y[y] * eq_c[c]: the bit-pattern path
        const uint64_t q = 1ull << (t - 2), cols = 1ull << s;
        if (!ctx->scratch.leaf_tab.alloc(leaf_tab_bytes(t))) return BITZ_ERR_ALLOC;
        uint4* qtab = (uint4*)ctx->scratch.leaf_tab.ptr;
        TimedLaunch tl(ctx, "gkr_leaf_round_sums", 2.0 * (double)(q << s) / 4.0 + 16.0 * (double)cols + q * 512.0,
                       q * 64.0 * kGfMulOps);
        hipLaunchKernelGGL(k_leaf_qtab, dim3((unsigned)((q * 16 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                           (const uint4*)A, t, (const uint4*)eq_y, send_one, qtab);
        unsigned nb;
        int rc = launch_pattern_sums<4>(ctx, packed, nr, t, s, q, (const uint4*)qtab, eq_c, &nb);
        if (rc) return rc;
        return sums_finish(ctx, nb, 2, out_dev);
    }
    uint32_t R = round_chunk_r(ctx->tune, h, s_c);
    unsigned nb = round_blocks(ctx->tune, h, R);
    int rc = ctx->scratch.ensure((size_t)nb * 2 * 16);
    if (rc) return rc;
    TimedLaunch tl(ctx, "gkr_leaf_round_sums", (double)(1ull << (t + s - 3)), 4.0 * h * kGfMulOps);
    LeafSrc src{packed, (const uint4*)A, t, s};
    hipLaunchKernelGGL(k_gkr_round_sums<LeafSrc>, dim3(nb), dim3(kT), 0, ctx->stream, src, h,
                       (const uint4*)eq_c, s_c, (const uint4*)eq_y, send_one, R, (uint4*)ctx->scratch.partials.ptr, nullptr, 0u);
    BITZ_LAUNCH_CHECK();
    return sums_finish(ctx, nb, 2, out_dev);
}

int fold_halves(bitz_ctx* ctx, void* v, uint64_t h, const uint8_t rho[16]) {
    uint64_t lo, hi;
    split16(rho, &lo, &hi);
    TimedLaunch tl(ctx, "fold_halves", 3.0 * 16.0 * h, h * kGfTabMulOps);
    hipLaunchKernelGGL(k_fold_halves, dim3(blocks_for(h, 4)), dim3(kT), 0, ctx->stream, (uint4*)v, h, lo, hi);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int fold_pair(bitz_ctx* ctx, void* l, void* r, uint64_t h, const uint8_t rho[16], const void* rho_dev) {
    uint64_t lo, hi;
    split16(rho, &lo, &hi);
    TimedLaunch tl(ctx, "fold_pair", 6.0 * 16.0 * h, 2.0 * h * kGfTabMulOps);
    hipLaunchKernelGGL(k_fold_pair, dim3(blocks_for(2 * h, 4)), dim3(kT), 0, ctx->stream, (uint4*)l, (uint4*)r, h,
                       lo, hi, (const uint4*)rho_dev);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_leaf_fold(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const uint8_t rho[16],
                  void* l_out, void* r_out) {
    BITZ_REQUIRE(t >= 2, BITZ_ERR_SHAPE, "gkr_leaf_fold: t < 2");
    uint64_t lo, hi;
    split16(rho, &lo, &hi);
    const uint64_t q = 1ull << (t - 2), h = 1ull << (t + s - 2);
    if (!ctx->scratch.leaf_tab.alloc(leaf_tab_bytes(t))) return BITZ_ERR_ALLOC;
    uint4* ftab = (uint4*)ctx->scratch.leaf_tab.ptr;
    TimedLaunch tl(ctx, "gkr_leaf_fold", 2.0 * 16.0 * h + (double)(1ull << (t + s - 3)), q * 8.0 * kGfTabMulOps);
    hipLaunchKernelGGL(k_leaf_ftab, dim3((unsigned)((q * 8 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)A, t, lo, hi, (const uint4*)nullptr, ftab);
    hipLaunchKernelGGL(k_leaf_fold_pattern, dim3(blocks_for(h, 4)), dim3(kT), 0, ctx->stream, packed, t, s,
                       (const uint4*)ftab, (uint4*)l_out, (uint4*)r_out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_leaf_round1_sums(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const uint8_t rho0[16],
                         const void* eq_c, const void* eq_y, int send_one, void* out_dev, const void* rho0_dev,
                         const void* nr) {
    BITZ_REQUIRE(t >= 3, BITZ_ERR_SHAPE, "gkr_leaf_round1_sums: t < 3");
    uint64_t lo, hi;
    split16(rho0, &lo, &hi);
    const uint64_t q = 1ull << (t - 2), q8 = q >> 1, cols = 1ull << s;
    if (!ctx->scratch.leaf_tab.alloc(leaf_tab_bytes(t))) return BITZ_ERR_ALLOC;
    uint4* ftab = (uint4*)ctx->scratch.leaf_tab.ptr;
    uint4* q1tab = leaf_tab_hi(ctx, t);
    TimedLaunch tl(ctx, "gkr_leaf_round1_sums", (double)(q8 << s) + 16.0 * (double)cols + q8 * 8192.0,
                   (q * 8.0 + q8 * 1024.0) * kGfMulOps);
    hipLaunchKernelGGL(k_leaf_ftab, dim3((unsigned)((q * 8 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)A, t, lo, hi, (const uint4*)rho0_dev, ftab);
    hipLaunchKernelGGL(k_leaf_q1tab, dim3((unsigned)((q8 * 256 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)ftab, t, (const uint4*)eq_y, send_one, q1tab);
    unsigned nb;
    int rc = launch_pattern_sums<8>(ctx, packed, nr, t, s, q8, (const uint4*)q1tab, eq_c, &nb);
    if (rc) return rc;
    return sums_finish(ctx, nb, 2, out_dev);
}

// Row lanes of a k_pattern_fold_hi / k_scale_cols_hi launch over `rows` rows
// of 2^(s-8) runs: about 2048 blocks in all.
static uint32_t hi_row_lanes(uint64_t rows, int s) {
    const uint64_t want = std::max<uint64_t>(1, 2048ull >> std::min(s - 8, 11));
    return (uint32_t)std::min<uint64_t>(rows, want);
}

int scale_cols_hi(bitz_ctx* ctx, void* v, uint64_t n, int s, const void* eq_hi) {
    BITZ_REQUIRE(s > 8 && s <= 28 && n % (1ull << s) == 0 && eq_hi, BITZ_ERR_SHAPE,
                 "scale_cols_hi: needs 8 < s <= 28 and n a multiple of 2^s");
    const uint64_t rows = n >> s;
    if (rows == 0) return BITZ_OK;
    const uint32_t lanes = hi_row_lanes(rows, s);
    TimedLaunch tl(ctx, "scale_cols_hi", 2.0 * 16.0 * n, (double)n * kGfTabMulOps);
    hipLaunchKernelGGL(k_scale_cols_hi, dim3(lanes << (s - 8)), dim3(kT), 0, ctx->stream, (uint4*)v, rows, s,
                       (const uint4*)eq_hi, lanes);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_leaf_fold2(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const uint8_t rho0[16],
                   const uint8_t rho1[16], const void* eq_c, void* l_out, void* r_out, const void* rho0_dev,
                   const void* rho1_dev, const void* eq_hi) {
    BITZ_REQUIRE(t >= 3, BITZ_ERR_SHAPE, "gkr_leaf_fold2: t < 3");
    BITZ_REQUIRE(!eq_hi || (!eq_c && s > 8 && s <= 28), BITZ_ERR_SHAPE,
                 "gkr_leaf_fold2: eq_hi needs eq_c null and 8 < s <= 28");
    uint64_t lo0, hi0, lo1, hi1;
    split16(rho0, &lo0, &hi0);
    split16(rho1, &lo1, &hi1);
    const uint64_t q = 1ull << (t - 2), q8 = q >> 1, h = 1ull << (t + s - 3);
    if (!ctx->scratch.leaf_tab.alloc(leaf_tab_bytes(t))) return BITZ_ERR_ALLOC;
    uint4* ftab = (uint4*)ctx->scratch.leaf_tab.ptr;
    uint4* gtab = leaf_tab_hi(ctx, t);
    TimedLaunch tl(ctx, "gkr_leaf_fold2", 2.0 * 16.0 * h + (double)(h / 8),
                   (q * 8.0 + q8 * 32.0 + (eq_hi ? (double)h : 0.0)) * kGfTabMulOps +
                       (eq_c ? (double)h : 0.0) * kGfMulOps);
    hipLaunchKernelGGL(k_leaf_ftab, dim3((unsigned)((q * 8 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)A, t, lo0, hi0, (const uint4*)rho0_dev, ftab);
    hipLaunchKernelGGL(k_leaf_gtab, dim3((unsigned)((q8 * 32 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)ftab, t, lo1, hi1, (const uint4*)rho1_dev, gtab);
    if (eq_hi) {
        const uint32_t lanes = hi_row_lanes(q8, s);
        hipLaunchKernelGGL(k_pattern_fold_hi<0>, dim3(lanes << (s - 8)), dim3(kT), 0, ctx->stream, packed, t, s,
                           (const uint4*)gtab, (const uint4*)eq_hi, lanes, (uint4*)l_out, (uint4*)r_out);
    } else {
        hipLaunchKernelGGL(k_pattern_fold<0>, dim3(blocks_for(h, 4)), dim3(kT), 0, ctx->stream, packed, t, s,
                           (const uint4*)gtab, (const uint4*)eq_c, (uint4*)l_out, (uint4*)r_out);
    }
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_level1_round0_sums(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const void* eq_c,
                           const void* eq_y, int send_one, void* out_dev, const void* nr) {
    BITZ_REQUIRE(t >= 3, BITZ_ERR_SHAPE, "gkr_level1_round0_sums: t < 3");
    const uint64_t Y1 = 1ull << (t - 1), Y3 = 1ull << (t - 3), cols = 1ull << s;
    if (!ctx->scratch.leaf_tab.alloc(leaf_tab_bytes(t))) return BITZ_ERR_ALLOC;
    uint4* vtab = (uint4*)ctx->scratch.leaf_tab.ptr;
    uint4* qtab = leaf_tab_hi(ctx, t);
    TimedLaunch tl(ctx, "gkr_level1_round0_sums", (double)(Y3 << s) + 16.0 * (double)cols + Y3 * 8192.0,
                   (Y1 + Y3 * 1024.0) * kGfMulOps);
    hipLaunchKernelGGL(k_level1_vtab, dim3((unsigned)((Y1 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)A, t, vtab);
    hipLaunchKernelGGL(k_level1_qtab, dim3((unsigned)((Y3 * 256 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)vtab, t, (const uint4*)eq_y, send_one, qtab);
    unsigned nb;
    int rc = launch_pattern_sums<8>(ctx, packed, nr, t, s, Y3, (const uint4*)qtab, eq_c, &nb);
    if (rc) return rc;
    return sums_finish(ctx, nb, 2, out_dev);
}

int gkr_level1_fold(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const uint8_t rho[16],
                    const void* eq_c, void* l_out, void* r_out, const void* rho_dev, const void* eq_hi) {
    BITZ_REQUIRE(t >= 3, BITZ_ERR_SHAPE, "gkr_level1_fold: t < 3");
    BITZ_REQUIRE(!eq_hi || (!eq_c && s > 8 && s <= 28), BITZ_ERR_SHAPE,
                 "gkr_level1_fold: eq_hi needs eq_c null and 8 < s <= 28");
    uint64_t lo, hi;
    split16(rho, &lo, &hi);
    const uint64_t Y1 = 1ull << (t - 1), Y3 = 1ull << (t - 3), h = 1ull << (t + s - 3);
    if (!ctx->scratch.leaf_tab.alloc(leaf_tab_bytes(t))) return BITZ_ERR_ALLOC;
    uint4* vtab = (uint4*)ctx->scratch.leaf_tab.ptr;
    uint4* gtab = leaf_tab_hi(ctx, t);
    TimedLaunch tl(ctx, "gkr_level1_fold", 2.0 * 16.0 * h + (double)(h / 8),
                   (Y1 + (eq_c ? (double)h : 0.0)) * kGfMulOps + (Y3 * 32.0 + (eq_hi ? (double)h : 0.0)) * kGfTabMulOps);
    hipLaunchKernelGGL(k_level1_vtab, dim3((unsigned)((Y1 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)A, t, vtab);
    hipLaunchKernelGGL(k_level1_gtab, dim3((unsigned)((Y3 * 32 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)vtab, t, lo, hi, (const uint4*)rho_dev, gtab);
    if (eq_hi) {
        const uint32_t lanes = hi_row_lanes(Y3, s);
        hipLaunchKernelGGL(k_pattern_fold_hi<1>, dim3(lanes << (s - 8)), dim3(kT), 0, ctx->stream, packed, t, s,
                           (const uint4*)gtab, (const uint4*)eq_hi, lanes, (uint4*)l_out, (uint4*)r_out);
    } else {
        hipLaunchKernelGGL(k_pattern_fold<1>, dim3(blocks_for(h, 4)), dim3(kT), 0, ctx->stream, packed, t, s,
                           (const uint4*)gtab, (const uint4*)eq_c, (uint4*)l_out, (uint4*)r_out);
    }
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_level2(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, void* out) {
    BITZ_REQUIRE(t >= 2, BITZ_ERR_SHAPE, "gkr_level2: t < 2");
    const uint64_t Y1 = 1ull << (t - 1), Y2 = 1ull << (t - 2), n = 1ull << (t + s - 2);
    if (!ctx->scratch.leaf_tab.alloc(leaf_tab_bytes(t))) return BITZ_ERR_ALLOC;
    uint4* vtab = (uint4*)ctx->scratch.leaf_tab.ptr;
    uint4* ltab = leaf_tab_hi(ctx, t);
    TimedLaunch tl(ctx, "gkr_level2", 16.0 * n + (double)(n / 8), (Y1 + Y2 * 16.0) * kGfMulOps);
    hipLaunchKernelGGL(k_level1_vtab, dim3((unsigned)((Y1 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)A, t, vtab);
    hipLaunchKernelGGL(k_level2_ltab, dim3((unsigned)((Y2 * 16 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)vtab, t, ltab);
    if (s >= 8) {
        const int cs = std::min(s, 12);
        hipLaunchKernelGGL(k_level_pattern_lds<4>, dim3((unsigned)(Y2 << (s - cs))), dim3(kT), 0, ctx->stream,
                           packed, t, s, cs, (const uint4*)ltab, (uint4*)out);
    } else {
        hipLaunchKernelGGL(k_level2_pattern, dim3(blocks_for(n, 4)), dim3(kT), 0, ctx->stream, packed, t, s,
                           (const uint4*)ltab, (uint4*)out);
    }
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_level3(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, void* out) {
    BITZ_REQUIRE(t >= 3, BITZ_ERR_SHAPE, "gkr_level3: t < 3");
    const uint64_t Y1 = 1ull << (t - 1), Y2 = 1ull << (t - 2), Y3 = 1ull << (t - 3), n = 1ull << (t + s - 3);
    if (!ctx->scratch.leaf_tab.alloc(leaf_tab_bytes(t))) return BITZ_ERR_ALLOC;
    uint4* vtab = (uint4*)ctx->scratch.leaf_tab.ptr;
    uint4* ltab = leaf_tab_hi(ctx, t);
    uint4* l3tab = ltab + Y2 * 16;
    TimedLaunch tl(ctx, "gkr_level3", 16.0 * n + (double)(n / 8), (Y1 + Y2 * 16.0 + Y3 * 256.0) * kGfMulOps);
    hipLaunchKernelGGL(k_level1_vtab, dim3((unsigned)((Y1 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)A, t, vtab);
    hipLaunchKernelGGL(k_level2_ltab, dim3((unsigned)((Y2 * 16 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)vtab, t, ltab);
    hipLaunchKernelGGL(k_level3_tab, dim3((unsigned)((Y3 * 256 + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)ltab, t, l3tab);
    if (s >= 8) {
        const int cs = std::min(s, 12);
        hipLaunchKernelGGL(k_level_pattern_lds<8>, dim3((unsigned)(Y3 << (s - cs))), dim3(kT), 0, ctx->stream,
                           packed, t, s, cs, (const uint4*)l3tab, (uint4*)out);
    } else {
        hipLaunchKernelGGL(k_level3_pattern, dim3(blocks_for(n, 4)), dim3(kT), 0, ctx->stream, packed, t, s,
                           (const uint4*)l3tab, (uint4*)out);
    }
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gkr_level1(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, void* out) {
    const uint64_t half = 1ull << (t - 1), n = 1ull << (t + s - 1);
    if (!ctx->scratch.leaf_tab.alloc(leaf_tab_bytes(t))) return BITZ_ERR_ALLOC;
    uint4* pa = (uint4*)ctx->scratch.leaf_tab.ptr;
    TimedLaunch tl(ctx, "gkr_level1", 16.0 * n + (double)(n / 8), half * kGfMulOps);
    hipLaunchKernelGGL(k_mul_elem, dim3(blocks_for(half, 4)), dim3(kT), 0, ctx->stream, (const uint4*)A,
                       (const uint4*)A + half, half, pa);
    hipLaunchKernelGGL(k_level1_pattern, dim3(blocks_for(n, 4)), dim3(kT), 0, ctx->stream, packed, t, s,
                       (const uint4*)A, (const uint4*)pa, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int product_level(bitz_ctx* ctx, const void* in, uint64_t n, void* out) {
    TimedLaunch tl(ctx, "product_level", 1.5 * 16.0 * n, (n / 2) * kGfMulOps);
    hipLaunchKernelGGL(k_product_level, dim3(blocks_for(n / 2, 4)), dim3(kT), 0, ctx->stream, (const uint4*)in,
                       n / 2, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int xi_combined(bitz_ctx* ctx, const uint64_t* packed, int t, int s, const void* w2, void* out) {
    const uint64_t n = 1ull << t, groups = ((1ull << s) + 63) / 64;
    const uint64_t chunks = (groups + kXiChunkWords - 1) / kXiChunkWords;
    const uint64_t row_blocks = (n + kXiRows - 1) / kXiRows;
    TimedLaunch tl(ctx, "xi_combined", 8.0 * (double)(groups << t) + 16.0 * n * (2 * chunks + 1), 0.0);
    uint4* dst = (uint4*)out;
    if (chunks > 1) {
        if (!ctx->scratch.fold_partial.alloc(chunks * n * 16)) return BITZ_ERR_ALLOC;
        dst = (uint4*)ctx->scratch.fold_partial.ptr;
    }
    hipLaunchKernelGGL(k_xi_nibbles, dim3((unsigned)chunks, (unsigned)row_blocks), dim3(kT), 0, ctx->stream, packed,
                       t, s, (const uint4*)w2, dst);
    if (chunks > 1)
        hipLaunchKernelGGL(k_xor_chunks, dim3(blocks_for(n, 1)), dim3(kT), 0, ctx->stream, (const uint4*)dst,
                           (uint32_t)chunks, n, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int fold_rows_eq(bitz_ctx* ctx, const uint64_t* rows, int t, int s, const void* eq, void* out) {
    BITZ_REQUIRE(t >= 6, BITZ_ERR_SHAPE, "fold_rows_eq: t < 6");
    const uint64_t cols = 1ull << s, chunk_rows = std::min<uint64_t>(kRowsEq, 1ull << t);
    const uint64_t chunks = (1ull << t) / chunk_rows, groups = (cols + 63) / 64;
    if (!ctx->scratch.fold_partial.alloc(chunks * cols * 16)) return BITZ_ERR_ALLOC;
    TimedLaunch tl(ctx, "fold_rows_eq", (double)((1ull << (t + s)) / 8) + 16.0 * (1ull << t), 0.0);
    hipLaunchKernelGGL(k_fold_rows_eq, dim3((unsigned)groups, (unsigned)chunks), dim3(kT), 0, ctx->stream, rows, t, s,
                       (const uint4*)eq, (uint4*)ctx->scratch.fold_partial.ptr);
    BITZ_LAUNCH_CHECK();
    return reduce_columns(ctx, (const uint4*)ctx->scratch.fold_partial.ptr, (uint32_t)chunks, (uint32_t)cols,
                          (uint4*)out);
}

int round_poly(bitz_ctx* ctx, const void* m, const void* w, uint64_t n, void* out_dev) {
    uint64_t pairs = n / 2;
    unsigned nb = blocks_for(pairs, 4, 2048);
    int rc = ctx->scratch.ensure((size_t)nb * 3 * 16);
    if (rc) return rc;
    hipLaunchKernelGGL(k_round_poly, dim3(nb), dim3(kT), 0, ctx->stream, (const uint4*)m, (const uint4*)w, pairs,
                       (uint4*)ctx->scratch.partials.ptr);
    BITZ_LAUNCH_CHECK();
    return sums_finish(ctx, nb, 3, out_dev);
}

int fold_lsb(bitz_ctx* ctx, const void* v, uint64_t n, const uint8_t r[16], void* out) {
    uint64_t lo, hi;
    split16(r, &lo, &hi);
    // in place needs a single pass order guarantee: out index i reads 2i, 2i+1;
    // a block may overwrite index i while another block still reads it, so
    // in-place folds go through the ping-pong buffer of the caller.
    BITZ_REQUIRE(v != out, BITZ_ERR_ARG, "fold_lsb: out must not alias v");
    hipLaunchKernelGGL(k_fold_lsb, dim3(blocks_for(n / 2, 4)), dim3(kT), 0, ctx->stream, (const uint4*)v, n / 2, lo,
                       hi, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

unsigned fold_1b_rows_chunks(uint64_t n) { return blocks_for(n, 8, kFold1bMaxChunks); }

int fold_1b_rows_into(bitz_ctx* ctx, const void* p, const void* suffix, uint64_t n, void* out_dev, void* partials) {
    const unsigned nb = fold_1b_rows_chunks(n);  // element chunks
    const uint64_t per = (n + nb - 1) / nb;
    // every block writes its chunk's 8 partial sums (zero for an empty
    // chunk), so the partials need no clearing
    hipLaunchKernelGGL(k_fold_1b_rows, dim3(nb * 16), dim3(kT), 0, ctx->stream, (const uint4*)p, (const uint4*)suffix,
                       n, per, (uint4*)partials);
    BITZ_LAUNCH_CHECK();
    // 128 independent sums over nb blocks: one column reduction per bit
    return reduce_columns(ctx, (const uint4*)partials, nb, 128, (uint4*)out_dev);
}

int fold_1b_rows(bitz_ctx* ctx, const void* p, const void* suffix, uint64_t n, void* out_dev) {
    const unsigned nb = fold_1b_rows_chunks(n);
    int rc = ctx->scratch.ensure((size_t)nb * 128 * 16);
    if (rc) return rc;
    TimedLaunch tl(ctx, "fold_1b_rows", 32.0 * (double)n, 0.0);
    return fold_1b_rows_into(ctx, p, suffix, n, out_dev, ctx->scratch.partials.ptr);
}

int reduce_columns(bitz_ctx* ctx, const uint4* partials, uint32_t nblocks, uint32_t width, uint4* out) {
    hipLaunchKernelGGL(k_reduce_columns, dim3((width + 15) / 16), dim3(256), 0, ctx->stream, partials, nblocks,
                       width, out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int fold_b128(bitz_ctx* ctx, const void* suffix, uint64_t n, const void* w, void* out) {
    hipLaunchKernelGGL(k_fold_b128, dim3(blocks_for(n, 8)), dim3(kT), 0, ctx->stream, (const uint4*)suffix, n,
                       (const uint4*)w, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int axpy(bitz_ctx* ctx, void* acc, const void* e, uint64_t n, const uint8_t eta[16]) {
    uint64_t lo, hi;
    split16(eta, &lo, &hi);
    hipLaunchKernelGGL(k_axpy, dim3(blocks_for(n, 4)), dim3(kT), 0, ctx->stream, (uint4*)acc, (const uint4*)e, n, lo,
                       hi);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int lig_msg(bitz_ctx* ctx, const void* f, const void* b, uint64_t n, int with_eval, void* out_dev) {
    uint64_t pairs = n / 2;
    unsigned nb = blocks_for(pairs, 4, 2048);
    int K = with_eval ? 3 : 2;
    int rc = ctx->scratch.ensure((size_t)nb * K * 16);
    if (rc) return rc;
    TimedLaunch tl(ctx, "lig_msg", 2.0 * 16.0 * n, (with_eval ? 3.0 : 2.0) * pairs * kGfMulOps);
    if (with_eval)
        hipLaunchKernelGGL(k_lig_msg<3>, dim3(nb), dim3(kT), 0, ctx->stream, (const uint4*)f, (const uint4*)b, pairs,
                           (uint4*)ctx->scratch.partials.ptr);
    else
        hipLaunchKernelGGL(k_lig_msg<2>, dim3(nb), dim3(kT), 0, ctx->stream, (const uint4*)f, (const uint4*)b, pairs,
                           (uint4*)ctx->scratch.partials.ptr);
    BITZ_LAUNCH_CHECK();
    return sums_finish(ctx, nb, K, out_dev);
}

int lig_fold(bitz_ctx* ctx, const void* f, const void* b, uint64_t n, const uint8_t r[16], void* fo, void* bo) {
    BITZ_REQUIRE(f != fo && b != bo, BITZ_ERR_ARG, "lig_fold: outputs must not alias inputs");
    uint64_t lo, hi;
    split16(r, &lo, &hi);
    TimedLaunch tl(ctx, "lig_fold", 3.0 * 16.0 * n, n * kGfTabMulOps);
    hipLaunchKernelGGL(k_lig_fold, dim3(blocks_for(n / 2, 4)), dim3(kT), 0, ctx->stream, (const uint4*)f,
                       (const uint4*)b, n / 2, lo, hi, (uint4*)fo, (uint4*)bo);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int induce_basis(bitz_ctx* ctx, const uint64_t* queries_dev, uint32_t nq, int log_n, const void* sks,
                 const void* inv_sks, const void* aw, void* out) {
    const int lo = std::min(log_n, 10);
    const uint64_t nlo = 1ull << lo, nhi = 1ull << (log_n - lo);
    const size_t what_elems = (size_t)std::max<uint32_t>(nq, 1) * std::max(log_n, 1);
    const size_t fac_elems = (size_t)nq * (nlo + nhi);
    if (!ctx->scratch.what.alloc((what_elems + fac_elems) * 16)) return BITZ_ERR_ALLOC;
    uint4* what = (uint4*)ctx->scratch.what.ptr;
    uint4* plo = what + what_elems;
    uint4* phi = plo + (size_t)nq * nlo;
    TimedLaunch tl(ctx, "induce_basis", 16.0 * (1ull << log_n), (double)nq * (1ull << log_n) * kGfMulOps);
    hipLaunchKernelGGL(k_induce_what, dim3((nq + kT - 1) / kT), dim3(kT), 0, ctx->stream, queries_dev, nq, log_n,
                       (const uint4*)sks, (const uint4*)inv_sks, what);
    if (nq > 0)
        hipLaunchKernelGGL(k_induce_factors, dim3((unsigned)((fac_elems + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                           (const uint4*)what, (const uint4*)aw, nq, log_n, lo, plo, phi);
    hipLaunchKernelGGL(k_induce_basis, dim3(blocks_for(1ull << log_n, 1)), dim3(kT), 0, ctx->stream,
                       (const uint4*)plo, (const uint4*)phi, nq, log_n, lo, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int pow_find(bitz_ctx* ctx, const uint8_t prefix[16], const uint8_t seed[16], uint32_t bits, uint64_t* nonce) {
    if (bits == 0) {
        *nonce = 0;
        return BITZ_OK;
    }
    BITZ_REQUIRE(bits <= 40, BITZ_ERR_ARG, "pow_find: bits too large");
    int rc = ctx->scratch.ensure(4096);
    if (rc) return rc;
    Scratch& sc = ctx->scratch;
    if (!sc.pow_best.alloc(256)) return BITZ_ERR_ALLOC;
    unsigned long long* best = (unsigned long long*)sc.pow_best.ptr;
    PowPrefix pre;
    memcpy(pre.w, prefix, 16);
    memcpy(pre.w + 4, seed, 16);
    // the cell is all ones unless this is the first search or the last one
    // ended without a mailbox report
    if (!sc.pow_best_ready || !sc.mbox_dev) hipLaunchKernelGGL(k_pow_init, dim3(1), dim3(64), 0, ctx->stream, best);
    sc.pow_best_ready = false;
    // spans of 16 x 2^bits nonces (at least 2^14, at most 2^20): the first
    // holds a solution with probability 1 - e^-16, and scanning spans in
    // order keeps the smallest valid nonce. After each span the best nonce
    // goes to the mailbox (no copy, no stream synchronisation).
    const uint64_t span = std::min<uint64_t>(1ull << 20, std::max<uint64_t>(1ull << 14, 16ull << bits));
    for (uint64_t base = 0;; base += span) {
        hipLaunchKernelGGL(k_pow_scan, dim3((unsigned)(span / kT)), dim3(kT), 0, ctx->stream, pre, base, span,
                           bits, best);
        unsigned long long h;
        if (sc.mbox_dev) {
            const uint32_t seq = ++sc.seq;
            // reset the cell with this report; a span without a solution
            // keeps it (the next span continues the search)
            hipLaunchKernelGGL(k_pow_report, dim3(1), dim3(64), 0, ctx->stream, best,
                               (uint4*)((uint8_t*)sc.mbox_dev + 64), (uint32_t*)sc.mbox_dev, seq, 1);
            BITZ_LAUNCH_CHECK();
            sc.mbox_pending = true;
            uint8_t v[16];
            if ((rc = read_result(ctx, 1, v))) return rc;
            memcpy(&h, v, 8);
            sc.pow_best_ready = true;  // the report reset the cell to all ones
        } else {
            BITZ_LAUNCH_CHECK();
            BITZ_HIP(hipMemcpyAsync(ctx->scratch.host, best, 8, hipMemcpyDeviceToHost, ctx->stream));
            BITZ_HIP(hipStreamSynchronize(ctx->stream));
            memcpy(&h, ctx->scratch.host, 8);
        }
        if (h != ~0ull) {
            *nonce = h;
            return BITZ_OK;
        }
        BITZ_REQUIRE(base < (1ull << 44), BITZ_ERR_DEVICE, "pow_find: search exhausted");
    }
}

int gkr_fs_round(bitz_ctx* ctx, unsigned nb, void* state, void* factor, const void* internal, const void* corr,
                 void* rho, void* msgs, uint32_t round) {
    int rc = ctx->scratch.ensure(64);
    if (rc) return rc;
    uint4* sums = (uint4*)ctx->scratch.result.ptr;
    hipLaunchKernelGGL(k_xor_partials<2>, dim3(1), dim3(kT), 0, ctx->stream, (const uint4*)ctx->scratch.partials.ptr,
                       nb, sums);
    hipLaunchKernelGGL(k_gkr_fs_round, dim3(1), dim3(kFsThreads), 0, ctx->stream, (const uint4*)sums,
                       (bitz::B3State*)state, (uint4*)factor, (const uint4*)internal, (const uint4*)corr,
                       (uint4*)rho, (uint4*)msgs, round);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int transcript_challenges(bitz_ctx* ctx, void* state, const void* msgs, uint32_t n, uint32_t msg_len, void* out) {
    BITZ_REQUIRE(msg_len <= 1536, BITZ_ERR_ARG, "transcript_challenges: messages of at most 1536 bytes");
    hipLaunchKernelGGL(k_transcript_challenges, dim3(1), dim3(kFsThreads), 0, ctx->stream, (bitz::B3State*)state,
                       (const uint8_t*)msgs, n, msg_len, (uint8_t*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

// ------------------------------------- transcript hashing of long messages
// Chaining values of m whole, non-final chunks of a BLAKE3 stream (chunk i
// has counter c0 + i), from 4-byte aligned data; one thread per chunk.
__global__ void k_b3_chunk_cvs(const uint32_t* __restrict__ data, uint64_t m, uint64_t c0,
                               uint32_t* __restrict__ cvs) {
    const uint64_t i = grid_tid();
    if (i >= m) return;
    const uint32_t* chunk = data + 256 * i;
    uint32_t cv[8];
    b3::iv_init(cv);
    for (int b = 0; b < 16; b++) {
        uint32_t w[16];
#pragma unroll
        for (int k = 0; k < 16; k++) w[k] = chunk[16 * b + k];
        b3::compress(cv, w, c0 + i, 64, (b == 0 ? b3::CHUNK_START : 0u) | (b == 15 ? b3::CHUNK_END : 0u));
    }
#pragma unroll
    for (int k = 0; k < 8; k++) cvs[8 * i + k] = cv[k];
}

// The runs' subtree chaining values: run r covers 2^ks[r] consecutive chunk
// CVs (in order from cvs[0]); one block merges every run level by level with
// PARENT compressions in place, then writes run r's root to roots[8r].
constexpr int kMaxB3Runs = 128;
struct B3Runs {
    uint8_t k[kMaxB3Runs];
    int n;
};
__global__ __launch_bounds__(256) void k_b3_runs(uint32_t* __restrict__ cvs, B3Runs runs,
                                                 uint32_t* __restrict__ roots) {
    int kmax = 0;
    for (int r = 0; r < runs.n; r++) kmax = max(kmax, (int)runs.k[r]);
    for (int level = 1; level <= kmax; level++) {
        // merges at this level: node j of run r joins children at
        // start_r + 2j 2^(level-1) and start_r + (2j+1) 2^(level-1)
        uint64_t start = 0;
        for (int r = 0; r < runs.n; r++) {
            const int k = runs.k[r];
            const uint64_t merges = k >= level ? 1ull << (k - level) : 0;
            for (uint64_t j = threadIdx.x; j < merges; j += blockDim.x) {
                const uint64_t left = start + (j << level), right = left + (1ull << (level - 1));
                uint32_t block[16], cv[8];
#pragma unroll
                for (int w = 0; w < 8; w++) {
                    block[w] = cvs[8 * left + w];
                    block[8 + w] = cvs[8 * right + w];
                }
                b3::iv_init(cv);
                b3::compress(cv, block, 0, 64, b3::PARENT);
#pragma unroll
                for (int w = 0; w < 8; w++) cvs[8 * left + w] = cv[w];
            }
            start += 1ull << k;
        }
        __syncthreads();
    }
    uint64_t start = 0;
    for (int r = 0; r < runs.n; r++) {
        if (threadIdx.x < 8) roots[8 * r + threadIdx.x] = cvs[8 * start + threadIdx.x];
        start += 1ull << runs.k[r];
    }
}

int blake3_runs(bitz_ctx* ctx, const void* data, uint64_t m, uint64_t c0, const std::vector<int>& ks, void* cvs,
                void* roots) {
    BITZ_REQUIRE(((uintptr_t)data & 3) == 0, BITZ_ERR_ARG, "blake3_runs: data must be 4-byte aligned");
    BITZ_REQUIRE((int)ks.size() <= kMaxB3Runs, BITZ_ERR_ARG, "blake3_runs: too many runs");
    uint64_t total = 0;
    B3Runs runs{};
    runs.n = (int)ks.size();
    for (int r = 0; r < runs.n; r++) {
        BITZ_REQUIRE(ks[r] >= 0 && ks[r] < 40, BITZ_ERR_ARG, "blake3_runs: run too long");
        runs.k[r] = (uint8_t)ks[r];
        total += 1ull << ks[r];
    }
    BITZ_REQUIRE(total == m, BITZ_ERR_ARG, "blake3_runs: runs do not tile the chunks");
    if (m == 0) return BITZ_OK;
    hipLaunchKernelGGL(k_b3_chunk_cvs, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, ctx->stream,
                       (const uint32_t*)data, m, c0, (uint32_t*)cvs);
    hipLaunchKernelGGL(k_b3_runs, dim3(1), dim3(256), 0, ctx->stream, (uint32_t*)cvs, runs, (uint32_t*)roots);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int dot(bitz_ctx* ctx, const void* a, const void* b, uint64_t n, void* out_dev) {
    unsigned nb = blocks_for(n, 4, 2048);
    int rc = ctx->scratch.ensure((size_t)nb * 16);
    if (rc) return rc;
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kT), 0, ctx->stream, (const uint4*)a, (const uint4*)b, n,
                       (uint4*)ctx->scratch.partials.ptr);
    BITZ_LAUNCH_CHECK();
    return sums_finish(ctx, nb, 1, out_dev);
}

int mul_elem(bitz_ctx* ctx, const void* a, const void* b, uint64_t n, void* out, int plus_one) {
    if (plus_one)
        hipLaunchKernelGGL(k_mul_elem_plus1, dim3(blocks_for(n, 4)), dim3(kT), 0, ctx->stream, (const uint4*)a,
                           (const uint4*)b, n, (uint4*)out);
    else
        hipLaunchKernelGGL(k_mul_elem, dim3(blocks_for(n, 4)), dim3(kT), 0, ctx->stream, (const uint4*)a,
                           (const uint4*)b, n, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int bits_round(bitz_ctx* ctx, int J, const uint64_t* bits, const void* w, uint64_t n, const uint8_t r[16],
               void* w_out) {
    BITZ_REQUIRE(J >= 0 && J <= 2 && n >= 2 && n % 2 == 0, BITZ_ERR_ARG, "bits_round: round or size out of range");
    BITZ_REQUIRE(J == 0 || (r && w_out && w_out != w), BITZ_ERR_ARG, "bits_round: fold needs a challenge and an output");
    uint64_t lo = 0, hi = 0;
    if (J > 0) split16(r, &lo, &hi);
    const uint64_t pairs = n / 2;  // pairs of the round's table (n entries after J folds)
    const int K = 3 << J;
    const unsigned nb = blocks_for(pairs, 4, 2048);
    int rc = ctx->scratch.ensure((size_t)nb * K * 16);
    if (rc) return rc;
    uint4* parts = (uint4*)ctx->scratch.partials.ptr;
    TimedLaunch tl(ctx, "bits_round", (double)pairs * (J ? 96.0 : 32.0) + (double)n / 8 * (1 << J), 0.0);
    switch (J) {
        case 0: hipLaunchKernelGGL(k_bits_round<0>, dim3(nb), dim3(kT), 0, ctx->stream, bits, (const uint4*)w, pairs, lo, hi, (uint4*)w_out, parts); break;
        case 1: hipLaunchKernelGGL(k_bits_round<1>, dim3(nb), dim3(kT), 0, ctx->stream, bits, (const uint4*)w, pairs, lo, hi, (uint4*)w_out, parts); break;
        default: hipLaunchKernelGGL(k_bits_round<2>, dim3(nb), dim3(kT), 0, ctx->stream, bits, (const uint4*)w, pairs, lo, hi, (uint4*)w_out, parts); break;
    }
    BITZ_LAUNCH_CHECK();
    return sums_finish(ctx, nb, K, nullptr);
}

int bits_materialize(bitz_ctx* ctx, const uint64_t* bits, const void* w, uint64_t n_out, const uint8_t r[16],
                     const uint8_t e8[128], void* a_out, void* w_out) {
    BITZ_REQUIRE(r && e8 && a_out && w_out && w_out != w, BITZ_ERR_ARG, "bits_materialize: bad arguments");
    uint64_t lo, hi;
    split16(r, &lo, &hi);
    E8 e;
    memcpy(&e, e8, sizeof(E8));
    TimedLaunch tl(ctx, "bits_materialize", 64.0 * (double)n_out + (double)n_out / 8, 0.0);
    hipLaunchKernelGGL(k_bits_materialize, dim3(blocks_for(n_out, 4)), dim3(kT), 0, ctx->stream, bits, (const uint4*)w,
                       n_out, lo, hi, e, (uint4*)a_out, (uint4*)w_out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int bits_to_gf(bitz_ctx* ctx, const uint64_t* bits, uint64_t n, void* out) {
    hipLaunchKernelGGL(k_bits_to_gf, dim3(blocks_for(n, 4)), dim3(kT), 0, ctx->stream, bits, n, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int mul_inplace(bitz_ctx* ctx, void* v, const void* b, uint64_t n) {
    hipLaunchKernelGGL(k_mul_inplace, dim3(blocks_for(n, 4)), dim3(kT), 0, ctx->stream, (uint4*)v, (const uint4*)b, n);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int scale(bitz_ctx* ctx, const void* a, uint64_t n, const uint8_t c[16], void* out) {
    uint64_t lo, hi;
    split16(c, &lo, &hi);
    hipLaunchKernelGGL(k_scale, dim3(blocks_for(n, 4)), dim3(kT), 0, ctx->stream, (const uint4*)a, n, lo, hi,
                       (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gather_rows(bitz_ctx* ctx, const void* cw, const uint64_t* q_dev, uint32_t nq, uint32_t lanes, void* out) {
    uint64_t n = (uint64_t)nq * lanes;
    if (n == 0) return BITZ_OK;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, ctx->stream,
                       (const uint4*)cw, q_dev, nq, lanes, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

int gather_nodes(bitz_ctx* ctx, const void* tree, const uint64_t* idx_dev, uint32_t n, void* out) {
    if (n == 0) return BITZ_OK;
    hipLaunchKernelGGL(k_gather_nodes, dim3((n + kT - 1) / kT), dim3(kT), 0, ctx->stream, (const uint4*)tree,
                       idx_dev, n, (uint4*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

}  // namespace bitz
