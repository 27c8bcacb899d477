// The SHA-256 chain proof: H_{i+1} = Compress(H_i, M_i) for N = 2^k blocks,
// proved as one linear relation over a derived bit grid and opened through
// the commitment to its source bits.
//
// prove_sha256_chain (src/piop/spartan/sha256/chain.rs:817-845) ->
// prove_linear (protocol/linear.rs:371-582):
//   the witness (chain.rs:635-794): the chain of states on the host, then
//     per instance on the device the compression's hint bits (the schedule
//     sums, the doubled round sums, the output sums), the committed source
//     cells [1 | block_i | hints_i ...] and the product tensor
//     D[local, instance] = (M f_bar_i)[local] with the terminal rows;
//   the commitment of the source rows (custom:1:4);
//   the statement frames (chain.rs:962-990: protocol, relation digest, map
//     digest, geometry, the public statement binding, the assignment
//     binding), the opener policy, Round 0, the runtime prime
//     (sha256/prime.rs:233-262), the relation frames, the local and instance
//     points, the constant-one and public-IO batching challenges
//     (linear.rs:306-340), the collapsed local rows beta = C^T eq(xi)
//     (sumcheck/bridge/repeated.rs:4-22), the rank-one product claim
//     (chain.rs:1080-1251) and its frame (linear.rs:324-368), the fork tag;
//   the virtual opening (bitz_opener.rs:1597-1624, src/bitz/virt.rs:
//     132-301): its statement, the fold and GKR over the product tensor, the
//     transposed weights through the chained map (binary.rs:831-876), and the
//     single-column inner product on the committed bits.
// Every grinding boundary of the lambda = 100 profile is at difficulty 0 for
// these shapes and absorbs nothing.
#include <string.h>

#include <algorithm>
#include <array>
#include <memory>

#include "chain_impl.h"
#include "common.h"
#include "gf128_dev.h"
#include "host_gf.h"
#include "host_proto.h"
#include "modq.h"
#include "open_impl.h"
#include "spartan_impl.h"

namespace bitz {
namespace {

using namespace chain;

constexpr char kProtocolDomain[] = "bitz/spartan-sha256-chain/product-linear/v1";
constexpr char kAssignmentDomain[] = "bitz/spartan-sha256-chain-assignment/runtime-prime/v1";
constexpr char kStatementDomain[] = "bitz/spartan-sha256-chain/public-statement/v1";
constexpr char kPublicIoDomain[] = "bitz/spartan-sha256-chain/public-io-batch/v1";
constexpr char kConstantOneDomain[] = "bitz/spartan-sha256-chain/constant-one-batch/v1";
constexpr char kLocalPointDomain[] = "bitz/spartan-sha256-chain/local-row-point/v1";
constexpr char kInstancePointDomain[] = "bitz/spartan-sha256-chain/instance-point/v1";
constexpr char kClaimDomain[] = "bitz/spartan-sha256-chain-opening/v1";
constexpr char kClaimTag[] = "sha256-chain-opening-claim";
constexpr char kPrimeDomain[] = "bitz/spartan-sha256/runtime-prime/v1";
constexpr char kPolicyLabel[] = "bitz/ligerito-policy/early-ood/v1";
constexpr char kVirtualSession[] = "bitz/bitz-opener/virtual/v1";
constexpr char kVirtualLabel[] = "bitz/virtual-statement/v2";
constexpr char kProfileName[] = "lambda100";
constexpr int kRate = 1, kBatch = 4;  // LigeritoSelection::JOHNSON (configuration.rs:24-27)

// FIPS 180-4 4.2.2 (device copy)
__constant__ uint32_t c_K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

__device__ __forceinline__ uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// the 64 instances' bit `b` of v, one LDS word per f_bar column (each wave
// holds the same 64 instances; only the wave that owns the call emits)
__device__ __forceinline__ void emit_bits(uint64_t* col, int base, uint64_t v, int nbits, uint32_t lane, bool mine) {
    if (!mine) return;
#pragma unroll 1
    for (int b = 0; b < nbits; b++) {
        const uint64_t m = __ballot((int)((v >> b) & 1ull));
        if (lane == 0) col[base + b] = m;
    }
}

// One workgroup of four wavefronts per 64 consecutive instances (n is a
// multiple of 64). Every wave replays the 64 compressions (lane j: instance
// 64 g + j) and emits the f_bar bits of every fourth call, in the circuit's
// column order ([1 | block | state | hints], hints in allocation order: 48
// schedule sums of 34 bits, per round the e then the a half-sums of 35 bits,
// 8 output sums of 33 bits); a ballot per bit makes col[c] the 64
// instances' bits of column c. Then:
//  * the product tensor: derived row r of the group is the XOR of the col
//    words of M's row r (m_off, m_cols over f_bar), stored as one 64-bit
//    row word of D (grid th x sh: cell = (r H + hi) 2^th + (instance mod
//    2^th), H = n >> th); the terminal rows hold the last instance's output
//    bits; rows past the chain's 20713 stay zero;
//  * the committed source stream: instance i's 6888 cells (block bits, then
//    hint bits) at bit 1 + 6888 i of src_bits, as 64-bit words shifted into
//    place, the wave w taking a quarter of them; words are OR-ed
//    (neighbouring instances share the end words); cell 0 (the constant) is
//    set once.
__global__ __launch_bounds__(256) void k_chain_witness(const uint32_t* __restrict__ blocks,
                                                       const uint32_t* __restrict__ states, uint64_t n,
                                                       const uint32_t* __restrict__ m_off,
                                                       const uint32_t* __restrict__ m_cols, int th,
                                                       uint64_t* __restrict__ D, uint64_t* __restrict__ src_bits) {
    __shared__ uint64_t col[kFBar];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t inst = blockIdx.x * 64ull + lane;
    uint32_t W[64];
    uint32_t st[8];
#pragma unroll
    for (int i = 0; i < 16; i++) W[i] = blocks[inst * 16 + i];
#pragma unroll
    for (int i = 0; i < 8; i++) st[i] = states[inst * 8 + i];
    if (threadIdx.x == 0) col[0] = ~0ull;
    uint32_t call = 0;  // emission calls, dealt round-robin to the waves
#pragma unroll
    for (int i = 0; i < 16; i++) emit_bits(col, 1 + 32 * i, W[i], 32, lane, (call++ & 3) == wave);
#pragma unroll
    for (int i = 0; i < 8; i++) emit_bits(col, 513 + 32 * i, st[i], 32, lane, (call++ & 3) == wave);
    int pos = 769;
#pragma unroll
    for (int i = 16; i < 64; i++) {
        const uint32_t s0 = rotr32(W[i - 15], 7) ^ rotr32(W[i - 15], 18) ^ (W[i - 15] >> 3);
        const uint32_t s1 = rotr32(W[i - 2], 17) ^ rotr32(W[i - 2], 19) ^ (W[i - 2] >> 10);
        const uint64_t total = (uint64_t)W[i - 16] + s0 + W[i - 7] + s1;
        emit_bits(col, pos, total, 34, lane, (call++ & 3) == wave);
        pos += 34;
        W[i] = (uint32_t)total;
    }
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3], e = st[4], f = st[5], g = st[6], h = st[7];
#pragma unroll
    for (int i = 0; i < 64; i++) {
        const uint32_t big1 = rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25);
        const uint32_t ch = (e & f) ^ (~e & g);
        const uint32_t big0 = rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22);
        const uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
        const uint64_t common = (uint64_t)h + big1 + c_K[i] + W[i] + ch;
        const uint64_t e_half = common + d;
        const uint64_t a_half = common + big0 + maj;
        emit_bits(col, pos, e_half, 35, lane, (call++ & 3) == wave);
        emit_bits(col, pos + 35, a_half, 35, lane, (call++ & 3) == wave);
        pos += 70;
        h = g;
        g = f;
        f = e;
        e = (uint32_t)e_half;
        d = c;
        c = b;
        b = a;
        a = (uint32_t)a_half;
    }
    const uint32_t wk[8] = {a, b, c, d, e, f, g, h};
#pragma unroll
    for (int i = 0; i < 8; i++) {
        emit_bits(col, pos, (uint64_t)st[i] + wk[i], 33, lane, (call++ & 3) == wave);
        pos += 33;
    }
    __syncthreads();
    // the product tensor's row words of this group
    const uint64_t first = blockIdx.x * 64ull;
    const uint64_t H = n >> th;
    const uint64_t hi = first >> th;
    const uint64_t wsel = (first & ((1ull << th) - 1)) >> 6;
    const uint64_t wpc = 1ull << (th - 6);
    const bool last_group = first + 64 == n;
    for (uint32_t r = threadIdx.x; r < (uint32_t)kChainHBar; r += blockDim.x) {
        uint64_t v = 0;
        if (r < (uint32_t)kHBar) {
            for (uint32_t k = m_off[r]; k < m_off[r + 1]; k++) v ^= col[m_cols[k]];
        } else if (last_group) {
            const int t = (int)r - kHBar;
            v = col[kOutputFBase + 33 * (t / 32) + t % 32] & (1ull << 63);
        }
        D[((uint64_t)r * H + hi) * wpc + wsel] = v;
    }
    // the committed source stream: instance-aligned local words q (cells
    // 64 q .. 64 q + 63 of the instance), shifted into place
    auto local_word = [&](int q) -> uint64_t {
        uint64_t cur = 0;
        for (int bb = 0; bb < 64; bb++) {
            const int j = q * 64 + bb;
            if (j >= kFInstance) break;
            const int cidx = j < 512 ? 1 + j : 769 + (j - 512);
            cur |= ((col[cidx] >> lane) & 1ull) << bb;
        }
        return cur;
    };
    constexpr int kWords = (kFInstance + 63) / 64, kPer = (kWords + 3) / 4;
    const uint64_t start = 1 + inst * (uint64_t)kFInstance, w0 = start >> 6;
    const int off = (int)(start & 63);
    const int q0 = (int)wave * kPer, q1 = min(kWords, q0 + kPer);
    uint64_t prev = q0 > 0 ? local_word(q0 - 1) : 0;
    for (int q = q0; q < q1; q++) {
        const uint64_t cur = local_word(q);
        const uint64_t word = off ? (cur << off) | (prev >> (64 - off)) : cur;
        if (word) atomicOr((unsigned long long*)&src_bits[w0 + q], (unsigned long long)word);
        prev = cur;
    }
    if (q1 == kWords && off && (prev >> (64 - off)))
        atomicOr((unsigned long long*)&src_bits[w0 + kWords], (unsigned long long)(prev >> (64 - off)));
    if (inst == 0 && wave == 0) atomicOr((unsigned long long*)&src_bits[0], 1ull);
}

// Transposed weights (VirtualStatement::transpose_query over the chained
// map): committed cell src = 1 + 6888 i + (lc - 1) (lc: chained f_bar
// column 1..6888) reads the derived rows of column lc of each map part at
// one instance: local at i, prev at i + 1, first at 0, last at n - 1. A
// derived cell d = r n + i' lies at row i' mod 2^th, column r H + (i' >> th),
// so each part contributes u1[i' mod 2^th] times the XOR of its rows' u2
// entries. Cell 0 (the constant) is set by the host; cells past the live
// ones are zero.
struct MapDev {
    const uint32_t* off;
    const uint32_t* rows;
};
// P[p][h][lc] (map part p = local, prev, first, last; h < H = n >> th; lc <
// kChainFBar) = the XOR of u2[row H + h] over the derived rows of column lc
// of part p: what every instance with i >> th = h reads for that column.
__global__ __launch_bounds__(256) void k_chain_part_sums(const uint4* __restrict__ u2, uint64_t H, MapDev loc,
                                                         MapDev prv, MapDev fst, MapDev lst, uint4* __restrict__ P) {
    const uint64_t per = H * (uint64_t)kChainFBar, total = 4 * per;
    for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        const int p = (int)(idx / per);
        const uint64_t rem = idx - p * per, h = rem / kChainFBar;
        const uint32_t lc = (uint32_t)(rem - h * kChainFBar);
        const MapDev m = p == 0 ? loc : p == 1 ? prv : p == 2 ? fst : lst;
        gf128 acc = gf_zero();
        for (uint32_t k = m.off[lc]; k < m.off[lc + 1]; k++) acc = gf_add(acc, gf_load(&u2[m.rows[k] * H + h]));
        gf_store(&P[idx], acc);
    }
}

// One block per instance i: its committed cells 1 + 6888 i + (lc - 1) take
// u1[i mod 2^th] times the local part's sum (and, at i = 0 and i = n - 1,
// the first and last parts') plus u1[(i + 1) mod 2^th] times the prev
// part's sum at instance i + 1. Both multipliers are constant over the
// block: window tables in LDS.
__global__ __launch_bounds__(256) void k_chain_transpose(const uint4* __restrict__ u1, const uint4* __restrict__ P,
                                                         uint64_t n, int th, uint4* __restrict__ wf) {
    __shared__ uint4 t0[512], t1[512];
    const uint64_t i = blockIdx.x, H = n >> th, mask = (1ull << th) - 1;
    const bool has_next = i + 1 < n;
    gf_table_build(t0, gf_load(&u1[i & mask]), threadIdx.x, blockDim.x);
    if (has_next) gf_table_build(t1, gf_load(&u1[(i + 1) & mask]), threadIdx.x, blockDim.x);
    __syncthreads();
    const uint64_t h0 = i >> th, h1 = (i + 1) >> th;
    const uint4* Ploc = P + h0 * kChainFBar;
    const uint4* Pprv = P + (H + (has_next ? h1 : 0)) * kChainFBar;
    const uint4* Pfst = P + 2 * H * kChainFBar;
    const uint4* Plst = P + (3 * H + h0) * kChainFBar;
    uint4* out = wf + 1 + (uint64_t)kFInstance * i - 1;
    for (uint32_t lc = 1 + threadIdx.x; lc < (uint32_t)kChainFBar; lc += blockDim.x) {
        gf128 acc = gf_mul_tab(t0, gf_load(&Ploc[lc]));
        if (has_next) {
            const gf128 sp = gf_load(&Pprv[lc]);
            if (!gf_is_zero(sp)) acc = gf_add(acc, gf_mul_tab(t1, sp));
        }
        if (i == 0) acc = gf_add(acc, gf_mul_tab(t0, gf_load(&Pfst[lc])));
        if (!has_next) acc = gf_add(acc, gf_mul_tab(t0, gf_load(&Plst[lc])));
        gf_store(&out[lc], acc);
    }
}

// wf[x] = 0 for x in [from, to): the cells past the live ones
__global__ __launch_bounds__(256) void k_chain_zero(uint4* __restrict__ wf, uint64_t from, uint64_t to) {
    for (uint64_t x = from + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < to;
         x += (uint64_t)gridDim.x * blockDim.x)
        wf[x] = make_uint4(0, 0, 0, 0);
}

// ---- host: the batch's device data ----
struct ChainDev {
    DevBuf m_off, m_cols, maps[4][2];
    MapDev md[4];
};
int upload_relation(bitz_ctx* ctx, const Relation& r, ChainDev* cd) {
    auto up = [&](const std::vector<uint32_t>& v, DevBuf& b) -> int {
        if (!b.alloc(std::max<size_t>(v.size(), 1) * 4)) return BITZ_ERR_ALLOC;
        BITZ_HIP(hipMemcpyAsync(b.ptr, v.data(), v.size() * 4, hipMemcpyHostToDevice, ctx->stream));
        return BITZ_OK;
    };
    int rc;
    if ((rc = up(r.m_off, cd->m_off)) || (rc = up(r.m_cols, cd->m_cols))) return rc;
    const Csc* parts[4] = {&r.local, &r.prev, &r.first, &r.last};
    for (int p = 0; p < 4; p++) {
        if ((rc = up(parts[p]->offsets, cd->maps[p][0])) || (rc = up(parts[p]->rows, cd->maps[p][1]))) return rc;
        cd->md[p] = MapDev{(const uint32_t*)cd->maps[p][0].ptr, (const uint32_t*)cd->maps[p][1].ptr};
    }
    // the host vectors must outlive the copies
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    return BITZ_OK;
}

void put64(std::vector<uint8_t>* o, uint64_t v) {
    for (int i = 0; i < 8; i++) o->push_back((uint8_t)(v >> (8 * i)));
}
void put128(std::vector<uint8_t>* o, u128 v) {
    put64(o, (uint64_t)v);
    put64(o, (uint64_t)(v >> 64));
}
void put32(std::vector<uint8_t>* o, uint32_t v) {
    for (int i = 0; i < 4; i++) o->push_back((uint8_t)(v >> (8 * i)));
}
void put_str(std::vector<uint8_t>* o, const char* s) { o->insert(o->end(), s, s + strlen(s)); }

// hash_security_params (sha256/proof.rs:2000-2034) at lambda 100
void security_bytes(const Batch& b, int ood_bits, std::vector<uint8_t>* o) {
    put64(o, strlen(kProfileName));
    put_str(o, kProfileName);
    put32(o, 100);
    put128(o, b.pmin);
    put128(o, b.pmax);
    o->push_back(0);  // not the full-width fingerprint
    put32(o, 0);      // initial, PIOP round and terminal grinding
    put32(o, 0);
    put32(o, 0);
    o->push_back(0);  // no reduction draw
    put32(o, 0);      // forest and ring-switch grinding
    put32(o, 0);
    o->push_back(1);  // Round 0 runs
    put32(o, (uint32_t)ood_bits);
    put64(o, 100);
}

// hash_ligerito_config (proof.rs:2036-2067) of the ladder's prover config
// (pcs/ligerito.rs:1441-1475)
void ligerito_bytes(const LigConfig& cfg, std::vector<uint8_t>* o) {
    const auto& lv = cfg.levels;
    put64(o, lv.size() - 1);
    put64(o, lv[0].log_msg_cols);
    put64(o, cfg.initial_k);
    put64(o, cfg.initial_k);
    auto vec = [&](auto field, size_t from) {
        put64(o, lv.size() - from);
        for (size_t i = from; i < lv.size(); i++) put64(o, (uint64_t)field(lv[i]));
    };
    vec([](const LigLevel& l) { return l.log_inv_rate; }, 0);
    vec([](const LigLevel& l) { return l.log_msg_cols; }, 1);
    vec([](const LigLevel& l) { return l.k_recursive; }, 1);
    vec([](const LigLevel& l) { return l.queries; }, 0);
    vec([](const LigLevel& l) { return l.grinding_bits; }, 0);
    vec([](const LigLevel& l) { return l.fold_grinding_bits; }, 0);
    vec([](const LigLevel& l) { return l.ood_samples; }, 0);
    o->push_back(1);  // BLAKE3
}

std::vector<u128> eq_le(const ModN& F, const std::vector<u128>& pt) {
    std::vector<u128> t(1, 1);
    for (u128 c : pt) {
        const size_t half = t.size();
        t.resize(2 * half);
        for (size_t j = 0; j < half; j++) {
            const u128 one = F.mul(t[j], c);
            t[j + half] = one;
            t[j] = addmod(t[j], F.n - one, F.n);
        }
    }
    return t;
}

struct Transposer {
    bitz_ctx* ctx;
    const Batch* b;
    ChainDev* cd;
};
int transpose_cb(void* user, const void* u1, const void* u2, void* wf) {
    Transposer* T = (Transposer*)user;
    bitz_ctx* ctx = T->ctx;
    const Batch& b = *T->b;
    const Relation& r = *b.rel;
    const uint64_t cells = 1ull << (b.tf + b.sf);
    const uint64_t H = b.n >> b.th;
    const uint64_t live = 1 + b.n * (uint64_t)kFInstance;
    BITZ_REQUIRE(cells < (1ull << 40) && live <= cells && b.n <= (1ull << 20), BITZ_ERR_SHAPE,
                 "chain transpose: grid out of range");
    DevBuf P;
    if (!P.alloc(4 * H * (uint64_t)kChainFBar * 16)) return BITZ_ERR_ALLOC;
    const uint64_t psums = 4 * H * (uint64_t)kChainFBar;
    hipLaunchKernelGGL(k_chain_part_sums, dim3((unsigned)std::min<uint64_t>((psums + 255) / 256, 65536)), dim3(256), 0,
                       ctx->stream, (const uint4*)u2, H, T->cd->md[0], T->cd->md[1], T->cd->md[2], T->cd->md[3],
                       (uint4*)P.ptr);
    BITZ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_chain_transpose, dim3((unsigned)b.n), dim3(256), 0, ctx->stream, (const uint4*)u1,
                       (const uint4*)P.ptr, b.n, b.th, (uint4*)wf);
    BITZ_LAUNCH_CHECK();
    if (live < cells) {
        hipLaunchKernelGGL(k_chain_zero, dim3((unsigned)std::min<uint64_t>((cells - live + 255) / 256, 65536)),
                           dim3(256), 0, ctx->stream, (uint4*)wf, live, cells);
        BITZ_LAUNCH_CHECK();
    }
    // cell 0: every instance's local column-0 rows (sum_lo u1 times the
    // rows' u2 sums over the high instances), instance 0's first rows and
    // instance n - 1's last rows
    std::vector<hgf> hu1(1ull << b.th), hu2(1ull << b.sh);
    BITZ_HIP(hipMemcpyAsync(hu1.data(), u1, hu1.size() * 16, hipMemcpyDeviceToHost, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(hu2.data(), u2, hu2.size() * 16, hipMemcpyDeviceToHost, ctx->stream));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    hgf s1{0, 0};
    for (auto& v : hu1) s1 = hgf_add(s1, v);
    hgf w0{0, 0};
    for (uint32_t k = r.local.offsets[0]; k < r.local.offsets[1]; k++) {
        hgf acc{0, 0};
        for (uint64_t hi = 0; hi < H; hi++) acc = hgf_add(acc, hu2[r.local.rows[k] * H + hi]);
        w0 = hgf_add(w0, hgf_mul(s1, acc));
    }
    auto part0 = [&](const Csc& m, uint64_t inst) {
        hgf acc{0, 0};
        for (uint32_t k = m.offsets[0]; k < m.offsets[1]; k++) acc = hgf_add(acc, hu2[m.rows[k] * H + (inst >> b.th)]);
        return hgf_mul(hu1[inst & ((1ull << b.th) - 1)], acc);
    };
    w0 = hgf_add(w0, part0(r.first, 0));
    w0 = hgf_add(w0, part0(r.last, b.n - 1));
    BITZ_HIP(hipMemcpyAsync(wf, &w0, 16, hipMemcpyHostToDevice, ctx->stream));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    return BITZ_OK;
}

// Round 0 on the outer transcript (prover: prove it; verifier: check the
// proof's nonce and value), returning BITZ_OK or an error
using Round0Fn = int (*)(void* user, Transcript& tr);

struct Outer {
    u128 q = 0, mu = 0;
    std::vector<u128> roww, colw;  // the rank-one claim's weights (canonical)
    uint8_t tag[32];
};

// prove_linear / verify_linear's outer transcript for the chain
// (linear.rs:398-558, 586-705) from the statement frames to the fork tag
int chain_outer(const Batch& b, const uint32_t* blocks, const uint32_t* digest, const uint8_t root[32],
                const LigConfig& cfg, Transcript& tr, Round0Fn round0, void* user, Outer* o, PhaseClock* ph) {
    auto mark = [&](const char* name) {
        if (ph) ph->mark(name);
    };
    int rc;
    const Relation& rel = *b.rel;
    const uint64_t n = b.n;
    const int k = b.k;
    const std::array<uint32_t, 8>& iv = rel.initial_state;
    // chain_statement_binding (chain.rs:1364-1383)
    uint8_t stmt[32], binding[32];
    {
        std::vector<uint8_t> o;
        put_str(&o, kStatementDomain);
        for (uint32_t w : iv) put32(&o, w);
        put64(&o, n);
        for (uint64_t i = 0; i < n * 16; i++) put32(&o, blocks[i]);
        for (int i = 0; i < 8; i++) put32(&o, digest[i]);
        Blake3 hs;
        hs.update(o.data(), o.size());
        hs.finalize_xof(stmt, 32);
    }
    // chain_assignment_binding (chain.rs:1385-1418)
    {
        std::vector<uint8_t> o;
        put_str(&o, kAssignmentDomain);
        o.insert(o.end(), root, root + 32);
        o.insert(o.end(), rel.digest, rel.digest + 32);
        o.insert(o.end(), b.map_digest, b.map_digest + 32);
        o.insert(o.end(), stmt, stmt + 32);
        const uint64_t vals[11] = {(uint64_t)(b.tf + b.sf), kRate, kBatch, n, (uint64_t)k, (uint64_t)b.th,
                                   (uint64_t)b.sh, 1, (uint64_t)b.tf, (uint64_t)b.sf, 1};
        for (uint64_t v : vals) put64(&o, v);
        o.push_back(0);  // profile code of the commitment
        o.push_back(1);  // BLAKE3
        security_bytes(b, cfg.round0_grinding_bits, &o);
        ligerito_bytes(cfg, &o);
        blake3_hash_host(o.data(), o.size(), binding);
    }
    spartan_frame(tr, "protocol", kProtocolDomain, strlen(kProtocolDomain));
    spartan_frame(tr, "integer-relation", rel.digest, 32);
    spartan_frame(tr, "chained-boolean-map", b.map_digest, 32);
    {
        const char* tags[6] = {"instance-vars", "instance-count", "assignment-row-vars", "assignment-column-vars",
                               "source-row-vars", "source-column-vars"};
        const uint64_t vals[6] = {(uint64_t)k, n, (uint64_t)b.th, (uint64_t)b.sh, (uint64_t)b.tf, (uint64_t)b.sf};
        for (int i = 0; i < 6; i++) spartan_frame(tr, tags[i], &vals[i], 8);
    }
    spartan_frame(tr, "public-sha256-chain", stmt, 32);
    spartan_frame(tr, "assignment-oracle", binding, 32);
    tr.absorb_slice(kPolicyLabel);
    {
        uint8_t dg[32];
        ligerito_digest(cfg, dg);
        tr.absorb_slice(dg, 32);
    }
    mark("chain_round0");
    if ((rc = round0(user, tr))) return rc;
    mark("chain_prime");
    // the runtime prime and the relation frames
    {
        uint8_t v[16];
        spartan_frame(tr, "prime-domain", kPrimeDomain, strlen(kPrimeDomain));
        const uint64_t kk = (uint64_t)k;
        spartan_frame(tr, "log-instance-capacity", &kk, 8);
        u128_put(b.pmin, v);
        spartan_frame(tr, "prime-min", v, 16);
        u128_put(b.pmax, v);
        spartan_frame(tr, "prime-max", v, 16);
    }
    u128 q;
    {
        std::string err;
        if (!sample_prime_context(tr, b.pmin, b.pmax, 128, &q, &err)) {
            set_error("sha256 chain: " + err);
            return BITZ_ERR_VERIFY;
        }
        uint8_t v[16];
        u128_put(q, v);
        spartan_frame(tr, "prime-q", v, 16);
        spartan_frame(tr, "runtime-field-modulus", v, 16);
        spartan_frame(tr, "projected-linear-relation", rel.digest, 32);
    }
    mark("chain_points");
    ModN F;
    F.set(q);
    auto point = [&](const char* domain, int vars, std::vector<u128>* out) -> bool {
        spartan_frame(tr, "challenge-domain", domain, strlen(domain));
        out->resize(vars);
        for (auto& v : *out)
            if (!spartan_squeeze(tr, q, &v)) return false;
        return true;
    };
    std::vector<u128> xi, eta, a0, lam;
    if (!point(kLocalPointDomain, kLocalVars, &xi) || !point(kInstancePointDomain, k, &eta) ||
        !point(kConstantOneDomain, 1, &a0) || !point(kPublicIoDomain, kPublicBits + 1, &lam)) {
        set_error("sha256 chain: sampling exhausted");
        return BITZ_ERR_VERIFY;
    }
    const u128 alpha0 = a0[0], alpha_pub = lam.back();
    lam.pop_back();
    mark("chain_batching");
    // beta = C^T eq(xi) and the rank-one coefficients (chain.rs:1080-1130)
    std::vector<u128> coef(kChainHBar, 0);
    {
        // eq(xi) in Montgomery form; a column's signed small coefficients
        // (|v| < 2^37) are summed as exact 256-bit products, positive and
        // negative apart, then reduced once each: redc(sum a R v) = sum a v
        std::vector<u128> ex = eq_le(F, xi);
        for (auto& v : ex) v = F.to_mont(v);
        for (int c = 0; c < kChainHBar; c++) {
            u128 ph = 0, pl = 0, nh = 0, nl = 0;
            for (uint32_t e = rel.c_off[c]; e < rel.c_off[c + 1]; e++) {
                const int64_t v = rel.c_vals[e];
                u128 hi, lo;
                ModN::wide(ex[rel.c_rows[e]], (u128)(v >= 0 ? (uint64_t)v : (uint64_t)0 - (uint64_t)v), &hi, &lo);
                u128& H = v >= 0 ? ph : nh;
                u128& L = v >= 0 ? pl : nl;
                L += lo;
                H += hi + (L < lo ? 1 : 0);
            }
            const u128 pos = F.redc(ph, pl), neg = F.redc(nh, nl);
            coef[c] = addmod(pos, neg ? q - neg : 0, q);
        }
        coef[0] = addmod(coef[0], alpha0, q);
        for (int p = 0; p < kPublicBits; p++) {
            const int c = public_h_column(p);
            coef[c] = addmod(coef[c], F.mul(alpha_pub, lam[p]), q);
        }
    }
    // mu = alpha0 sum_i u_i + alpha_pub sum_i u_i sum_p lambda_p bit_p (chain.rs:1131-1160)
    mark("chain_mu");
    u128 mu;
    {
        const std::vector<u128> u = eq_le(F, eta);
        // byte tables of the slot weights, unreduced: an entry sums at most 8
        // weights (< 2^116) and an instance's value 96 entries (< 2^123), so
        // u_i * value < 2^236 < R q and one Montgomery product per instance
        // gives u_i value R^-1; the sum is brought back by R at the end
        std::vector<u128> tab((size_t)96 * 256);
        for (int t8 = 0; t8 < 96; t8++) {
            tab[t8 * 256] = 0;
            for (int v = 1; v < 256; v++) tab[t8 * 256 + v] = tab[t8 * 256 + (v & (v - 1))] + lam[8 * t8 + __builtin_ctz(v)];
        }
        u128 usum = 0, acc = 0;
        for (uint64_t i = 0; i < n; i++) {
            usum = addmod(usum, u[i], q);
            u128 val = 0;
            for (int ws = 0; ws < 24; ws++) {
                const uint32_t word = ws < 16 ? blocks[i * 16 + ws] : (i + 1 == n ? digest[ws - 16] : 0u);
                for (int by = 0; by < 4; by++) val += tab[(4 * ws + by) * 256 + ((word >> (8 * by)) & 255)];
            }
            acc = addmod(acc, F.mont(u[i], val), q);
        }
        acc = F.to_mont(acc);
        mu = addmod(F.mul(alpha0, usum), F.mul(alpha_pub, acc), q);
    }
    // chain_product_opening_claim (chain.rs:1201-1251)
    o->roww = eq_le(F, std::vector<u128>(eta.begin(), eta.begin() + b.th));
    const std::vector<u128>& roww = o->roww;
    const std::vector<u128> high(eq_le(F, std::vector<u128>(eta.begin() + b.th, eta.end())));
    const uint64_t H = high.size(), ncols = 1ull << b.sh;
    o->colw.assign(ncols, 0);
    std::vector<u128>& colw = o->colw;
    if (H == 1)  // eq over no coordinates: the coefficients themselves
        std::copy(coef.begin(), coef.end(), colw.begin());
    else
        for (uint64_t c = 0; c < std::min<uint64_t>(ncols, (uint64_t)kChainHBar * H); c++)
            colw[c] = F.mul(coef[c / H], high[c % H]);
    // the claim frame (linear.rs:324-368)
    mark("chain_claim_frame");
    {
        Blake3 hd;
        std::vector<uint8_t> o;
        put_str(&o, kClaimDomain);
        o.insert(o.end(), binding, binding + 32);
        put64(&o, xi.size());
        for (u128 v : xi) put128(&o, v);
        put64(&o, eta.size());
        for (u128 v : eta) put128(&o, v);
        put128(&o, alpha0);
        put64(&o, lam.size());
        for (u128 v : lam) put128(&o, v);
        put128(&o, alpha_pub);
        put64(&o, roww.size());
        for (u128 v : roww) put128(&o, v);
        hd.update(o.data(), o.size());
        o.clear();
        put64(&o, colw.size());
        hd.update(o.data(), o.size());
        hd.update(colw.data(), colw.size() * 16);
        o.clear();
        put128(&o, mu);
        hd.update(o.data(), o.size());
        uint8_t dg[32];
        hd.finalize_xof(dg, 32);
        spartan_frame(tr, kClaimTag, dg, 32);
    }
    o->q = q;
    o->mu = mu;
    tr.challenge(o->tag, 16);
    tr.challenge(o->tag + 16, 16);
    return BITZ_OK;
}

// the virtual statement's public messages (virt.rs:132-150): label, root,
// claim parameters on the derived grid + committed shape + map digest, the
// 40-byte Pcs frame; the claim follows
std::vector<std::vector<uint8_t>> virtual_lead(const Batch& b, const uint8_t root[32], u128 q) {
    std::vector<std::vector<uint8_t>> lead(4);
    put_str(&lead[0], kVirtualLabel);
    lead[1].assign(root, root + 32);
    std::vector<uint8_t>& o = lead[2];
    put64(&o, b.th);
    put64(&o, b.sh);
    put128(&o, q);
    const hgf gen = bitz_generator();
    put64(&o, gen.lo);
    put64(&o, gen.hi);
    put64(&o, b.tf);
    put64(&o, b.sf);
    o.insert(o.end(), b.map_digest, b.map_digest + 32);
    const uint64_t pcs40[5] = {(uint64_t)(b.tf + b.sf), kRate, kBatch, 0, 1};
    for (uint64_t v : pcs40) put64(&lead[3], v);
    return lead;
}

struct ProveR0 {
    bitz_ctx* ctx;
    const bitz_commit_handle* h;
    Round0Claim r0;
};
int prove_round0(void* user, Transcript& tr) {
    ProveR0* P = (ProveR0*)user;
    return round0_prove(P->ctx, P->h, tr, &P->r0);
}

int prove_chain(bitz_ctx* ctx, const uint32_t* blocks, int k, const uint32_t* initial_state, uint8_t root[32],
                uint32_t digest_out[8], uint8_t state_out[32], std::vector<uint8_t>* proof) {
    std::array<uint32_t, 8> iv;
    for (int i = 0; i < 8; i++) iv[i] = initial_state ? initial_state[i] : kIV[i];
    Batch b;
    BITZ_REQUIRE(make_batch(k, iv, &b), BITZ_ERR_SHAPE, "sha256 chain: needs 2^7..2^16 compressions");
    const Relation& rel = *b.rel;
    const uint64_t n = b.n;
    PhaseClock phase(ctx);
    phase.mark("chain_states");
    // the native chain (the statement's digest)
    std::vector<uint32_t> states((n + 1) * 8);
    memcpy(states.data(), iv.data(), 32);
    for (uint64_t i = 0; i < n; i++) compress(&states[i * 8], &blocks[i * 16], &states[(i + 1) * 8]);
    const uint32_t* digest = &states[n * 8];
    memcpy(digest_out, digest, 32);
    phase.mark("chain_witness");
    int rc;
    ChainDev cd;
    if ((rc = upload_relation(ctx, rel, &cd))) return rc;
    DevBuf dblocks, dstates, src, D;
    const uint64_t src_words = 1ull << (b.tf + b.sf - 6), d_words = 1ull << (b.th + b.sh - 6);
    if (!dblocks.alloc(n * 64) || !dstates.alloc((n + 1) * 32) || !src.alloc(src_words * 8) || !D.alloc(d_words * 8))
        return BITZ_ERR_ALLOC;
    BITZ_HIP(hipMemcpyAsync(dblocks.ptr, blocks, n * 64, hipMemcpyHostToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(dstates.ptr, states.data(), (n + 1) * 32, hipMemcpyHostToDevice, ctx->stream));
    BITZ_HIP(hipMemsetAsync(src.ptr, 0, src_words * 8, ctx->stream));
    BITZ_HIP(hipMemsetAsync(D.ptr, 0, d_words * 8, ctx->stream));
    hipLaunchKernelGGL(k_chain_witness, dim3((unsigned)(n / 64)), dim3(256), 0, ctx->stream,
                       (const uint32_t*)dblocks.ptr, (const uint32_t*)dstates.ptr, n, (const uint32_t*)cd.m_off.ptr,
                       (const uint32_t*)cd.m_cols.ptr, b.th, (uint64_t*)D.ptr, (uint64_t*)src.ptr);
    BITZ_LAUNCH_CHECK();
    BITZ_HIP(hipStreamSynchronize(ctx->stream));  // blocks and states are the caller's / local host memory
    phase.mark("chain_commit");
    std::unique_ptr<bitz_commit_handle, void (*)(bitz_commit_handle*)> h(nullptr, bitz_commit_free);
    {
        bitz_commit_handle* hp = nullptr;
        if ((rc = bitz_commit_device_rows(ctx, src.ptr, b.tf, b.sf, kBatch, kRate, root, &hp))) return rc;
        h.reset(hp);
    }
    src = DevBuf();
    LigConfig cfg;
    {
        std::string err = custom_johnson_config(b.tf + b.sf, kRate, kBatch, 100, &cfg);
        BITZ_REQUIRE(err.empty(), BITZ_ERR_SHAPE, "sha256 chain: Ligerito configuration: " + err);
    }
    phase.mark("chain_statement");
    Transcript tr;
    ProveR0 pr0{ctx, h.get(), {}};
    Outer O;
    if ((rc = chain_outer(b, blocks, digest, root, cfg, tr, prove_round0, &pr0, &O, &phase))) return rc;
    if (state_out) tr.digest(state_out);
    phase.mark("chain_open");
    const std::vector<std::vector<uint8_t>> lead = virtual_lead(b, root, O.q);
    DevBuf droww, dcolw;
    if (!droww.alloc(O.roww.size() * 16) || !dcolw.alloc(O.colw.size() * 16)) return BITZ_ERR_ALLOC;
    BITZ_HIP(hipMemcpyAsync(droww.ptr, O.roww.data(), O.roww.size() * 16, hipMemcpyHostToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(dcolw.ptr, O.colw.data(), O.colw.size() * 16, hipMemcpyHostToDevice, ctx->stream));
    Transposer T{ctx, &b, &cd};
    rc = open_virtual_claim(ctx, h.get(), kVirtualSession, O.tag, lead, D.ptr, b.th, b.sh, O.q, droww.ptr, dcolw.ptr,
                            O.mu, transpose_cb, &T, pr0.r0, proof);
    BITZ_HIP(hipStreamSynchronize(ctx->stream));  // the host weights are copied before they go
    return rc;
}

// the opening's bytes (BitZOpeningProof::from_bytes, bitz_opener.rs:325-372)
struct OpeningBytes {
    const uint8_t *narg = nullptr, *hints = nullptr, *y = nullptr;
    uint64_t narg_len = 0, hints_len = 0, nonce = 0;
    bool has_nonce = false;
};
bool parse_opening(const uint8_t* p, size_t len, OpeningBytes* o) {
    size_t pos = 0;
    auto rd64 = [&](uint64_t* v) {
        if (pos + 8 > len) return false;
        memcpy(v, p + pos, 8);
        pos += 8;
        return true;
    };
    if (!rd64(&o->narg_len) || o->narg_len > len - pos) return false;
    o->narg = p + pos;
    pos += o->narg_len;
    if (!rd64(&o->hints_len) || o->hints_len > len - pos) return false;
    o->hints = p + pos;
    pos += o->hints_len;
    if (pos + 17 > len || p[pos] != 1) return false;
    o->y = p + pos + 1;
    pos += 17;
    if (pos >= len || p[pos] > 1) return false;
    o->has_nonce = p[pos++] == 1;
    if (o->has_nonce && !rd64(&o->nonce)) return false;
    return pos == len;
}

struct VerifyR0 {
    const OpeningBytes* ob;
    int mp;
    uint32_t bits;
    std::vector<hgf> point;
    hgf y;
};
// verify_ood_round (ligerito_flock.rs:1551-1600): the header, the grinding
// boundary with the proof's nonce, zeta, the proof's y absorbed
int verify_round0(void* user, Transcript& tr) {
    VerifyR0* V = (VerifyR0*)user;
    ood_round_header(tr, V->mp, V->bits);
    if (V->bits) {
        if (!V->ob->has_nonce) {
            set_error("verify: missing Round-0 nonce");
            return BITZ_ERR_VERIFY;
        }
        uint8_t seed[32];
        grinding_seed(tr, "bitz/core/ood-round-grinding/v1", 0, V->bits, seed);
        const bool ok = round0_pow_valid(seed, V->ob->nonce, V->bits);
        absorb_grinding_nonce(tr, V->ob->nonce);
        if (!ok) {
            set_error("verify: Round-0 proof of work");
            return BITZ_ERR_VERIFY;
        }
    } else if (V->ob->has_nonce) {
        set_error("verify: unexpected Round-0 nonce");
        return BITZ_ERR_VERIFY;
    }
    uint8_t zb[16];
    tr.challenge(zb, 16);
    hgf cur;
    memcpy(&cur.lo, zb, 8);
    memcpy(&cur.hi, zb + 8, 8);
    V->point.resize(V->mp);
    for (int i = 0; i < V->mp; i++) {
        V->point[i] = cur;
        cur = hgf_mul(cur, cur);
    }
    uint8_t tagged[17];
    tagged[0] = 0x50;
    memcpy(tagged + 1, V->ob->y, 16);
    tr.absorb_slice(tagged, 17);
    memcpy(&V->y.lo, V->ob->y, 8);
    memcpy(&V->y.hi, V->ob->y + 8, 8);
    return BITZ_OK;
}

int verify_chain(bitz_ctx* ctx, const uint32_t* blocks, int k, const uint32_t* initial_state, const uint32_t digest[8],
                 const uint8_t root[32], const uint8_t* proof, size_t plen, uint8_t state_out[32]) {
    std::array<uint32_t, 8> iv;
    for (int i = 0; i < 8; i++) iv[i] = initial_state ? initial_state[i] : kIV[i];
    Batch b;
    BITZ_REQUIRE(make_batch(k, iv, &b), BITZ_ERR_SHAPE, "sha256 chain: needs 2^7..2^16 compressions");
    OpeningBytes ob;
    if (!parse_opening(proof, plen, &ob)) {
        set_error("verify: malformed opening proof");
        return BITZ_ERR_VERIFY;
    }
    LigConfig cfg;
    {
        std::string err = custom_johnson_config(b.tf + b.sf, kRate, kBatch, 100, &cfg);
        BITZ_REQUIRE(err.empty(), BITZ_ERR_SHAPE, "sha256 chain: Ligerito configuration: " + err);
    }
    int rc;
    Transcript tr;
    VerifyR0 vr0{&ob, b.tf + b.sf - 7, (uint32_t)cfg.round0_grinding_bits, {}, {0, 0}};
    Outer O;
    if ((rc = chain_outer(b, blocks, digest, root, cfg, tr, verify_round0, &vr0, &O, nullptr))) return rc;
    if (state_out) tr.digest(state_out);
    ChainDev cd;
    if ((rc = upload_relation(ctx, *b.rel, &cd))) return rc;
    const std::vector<std::vector<uint8_t>> lead = virtual_lead(b, root, O.q);
    DevBuf droww, dcolw;
    if (!droww.alloc(O.roww.size() * 16) || !dcolw.alloc(O.colw.size() * 16)) return BITZ_ERR_ALLOC;
    BITZ_HIP(hipMemcpyAsync(droww.ptr, O.roww.data(), O.roww.size() * 16, hipMemcpyHostToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(dcolw.ptr, O.colw.data(), O.colw.size() * 16, hipMemcpyHostToDevice, ctx->stream));
    Transposer T{ctx, &b, &cd};
    rc = verify_virtual_claim(ctx, root, b.tf, b.sf, kVirtualSession, O.tag, lead, b.th, b.sh, O.q, droww.ptr,
                              dcolw.ptr, O.mu, transpose_cb, &T, vr0.point, vr0.y, ob.narg, ob.narg_len, ob.hints,
                              ob.hints_len);
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    return rc;
}

}  // namespace
}  // namespace bitz

extern "C" int bitz_prove_sha256_chain(bitz_ctx* ctx, const uint32_t* blocks, uint32_t log_n,
                                       const uint32_t* initial_state, uint8_t root[32], uint32_t digest[8],
                                       uint8_t state[32], uint8_t* proof, uint64_t capacity, uint64_t* proof_len) {
    BITZ_REQUIRE(ctx && blocks && root && digest && proof_len, BITZ_ERR_ARG, "bitz_prove_sha256_chain: null argument");
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    std::vector<uint8_t> bytes;
    uint8_t st[32];
    int rc = bitz::prove_chain(ctx, blocks, (int)log_n, initial_state, root, digest, st, &bytes);
    if (rc) return rc;
    if (state) memcpy(state, st, 32);
    *proof_len = bytes.size();
    BITZ_REQUIRE(proof && capacity >= bytes.size(), BITZ_ERR_ARG, "bitz_prove_sha256_chain: proof buffer too small");
    memcpy(proof, bytes.data(), bytes.size());
    return BITZ_OK;
}

extern "C" int bitz_verify_sha256_chain(bitz_ctx* ctx, const uint32_t* blocks, uint32_t log_n,
                                        const uint32_t* initial_state, const uint32_t digest[8],
                                        const uint8_t root[32], const uint8_t* proof, uint64_t proof_len,
                                        uint8_t state[32]) {
    BITZ_REQUIRE(ctx && blocks && digest && root && (proof || proof_len == 0), BITZ_ERR_ARG,
                 "bitz_verify_sha256_chain: null argument");
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    return bitz::verify_chain(ctx, blocks, (int)log_n, initial_state, digest, root, proof, (size_t)proof_len, state);
}
