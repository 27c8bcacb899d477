// The u64 multiplication proof: x y = lo + 2^64 hi for a batch of u64
// pairs, committed as the 256-slot bit grid and proved by the Spartan PIOP
// over a runtime prime field F_q whose terminal claim is bitified and opened
// by BitZ.
//
// protocol::prove for MulLayout<u64> (src/piop/spartan/protocol/mod.rs:
// 1138-1170) -> bitz_opener::prove_direct (bitz_opener.rs:439-494):
//   the statement: the assignment binding as a Spartan frame
//     (u64_bitz.rs:135-171, binding.rs:185-208, mod.rs:866-887), the opener
//     policy digest (configuration.rs:224-227) and Round 0;
//   the session label and the ladder digest;
//   prove_piop (mod.rs:1410-1506): the runtime prime (sample_mod_q,
//     mod.rs:1852-1870), the Spartan statement (piop.rs:963-983, with the
//     constraint-matrix digest, matrix.rs:1478-1513), tau, the outer cubic
//     sumcheck over the gate variables (sumcheck/outer: ordinary.rs,
//     engine.rs, proof.rs:337-377), rho, the inner quadratic sumcheck over
//     the assignment (sumcheck/inner/mod.rs:80-150), bitification
//     (bitify.rs:119-352) and the bridge digest (bitify.rs:365-410);
//   the fork tag, then the BitZ opening of the bitified linear claim on the
//   inner transcript (session "bitz/bitz-opener/v1", the bridge digest first).
// Every grinding boundary of the lambda = 100 profile is at difficulty 0 for
// the supported shapes (t <= 24) and absorbs nothing.
//
// F_q arithmetic on the device is Montgomery with R = 2^128 on unsigned
// 128-bit integers (q < 2^113). Tables: A = x, B = y, C = lo + 2^64 hi, and
// the assignment blocks lo, hi, one entry per gate, in Montgomery form.
#include <string.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <thread>

#include "common.h"
#include "host_gf.h"
#include "host_proto.h"
#include "modq.h"
#include "open_impl.h"
#include "spartan_impl.h"

namespace bitz {
namespace {

constexpr char kBindingDomain[] = "bitz/spartan-u64-bitz/assignment/v1-runtime";
constexpr char kBlockOrder[] = "e0|x|y|zlo|zhi|zero|zero|zero";
constexpr char kStatementTag[] = "u64-mul-statement";
constexpr char kPrimeDomain[] = "bitz/spartan-u64-mul/runtime-prime/v1";
constexpr char kBridgeDomain[] = "bitz/spartan-u64-bitz/bitified-claim/v1";
constexpr char kFrameDomain[] = "bitz/spartan/transcript-frame/v2";
constexpr char kPiopDomain[] = "bitz/spartan/piop/v2";
constexpr char kOracleDomain[] = "bitz/spartan/assignment-oracle/v1";
constexpr char kMatrixDomain[] = "bitz/spartan/constraint-matrices/v2";
constexpr char kPolicyLabel[] = "bitz/ligerito-policy/early-ood/v1";
constexpr char kSession[] = "bitz/bitz-opener/v1";
constexpr int kRate = 1, kBatch = 4;  // LigeritoSelection::JOHNSON, custom:1:4 (configuration.rs:24-27)
constexpr int kLambda = 100;
constexpr int kT = 256;

// ------------------------------------------------------------- device F_q
__device__ __forceinline__ void fq_wide(u128 a, u128 b, u128* hi, u128* lo) {
    const uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64), b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
    const u128 p00 = (u128)a0 * b0, p01 = (u128)a0 * b1, p10 = (u128)a1 * b0, p11 = (u128)a1 * b1;
    const u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
    *lo = (uint64_t)p00 | (mid << 64);
    *hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
}
// a b R^-1 mod q for a b < R q
__device__ __forceinline__ u128 fq_mont(u128 a, u128 b, u128 q, u128 qinv) {
    u128 hi, lo, mh, ml;
    fq_wide(a, b, &hi, &lo);
    const u128 m = lo * qinv;
    fq_wide(m, q, &mh, &ml);
    const u128 c = (lo + ml) < lo ? 1 : 0;
    const u128 r = hi + mh + c;
    return r >= q ? r - q : r;
}
__device__ __forceinline__ u128 fq_add(u128 a, u128 b, u128 q) {
    const u128 s = a + b;  // q < 2^126: no overflow
    return s >= q ? s - q : s;
}
__device__ __forceinline__ u128 fq_sub(u128 a, u128 b, u128 q) { return a >= b ? a - b : a + q - b; }
__device__ __forceinline__ u128 mk(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

__device__ __forceinline__ uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t gthreads() { return (uint64_t)gridDim.x * blockDim.x; }

// K sums mod q over the block; thread 0 writes them to out[0..K)
template <int K>
__device__ __forceinline__ void block_sum_store(u128 (&v)[K], u128 q, u128* out) {
    __shared__ u128 sh[K][kT];
#pragma unroll
    for (int k = 0; k < K; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < K; k++) sh[k][threadIdx.x] = fq_add(sh[k][threadIdx.x], sh[k][threadIdx.x + w], q);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) out[k] = sh[k][0];
    }
}

struct FqArgs {
    uint64_t q[2], qinv[2], r2[2];
    __device__ u128 Q() const { return mk(q[0], q[1]); }
    __device__ u128 Qinv() const { return mk(qinv[0], qinv[1]); }
    __device__ u128 R2() const { return mk(r2[0], r2[1]); }
};

// project_u64_mul_witness (u64_mul.rs:186-246) into Montgomery tables:
// A = x, B = y, C = lo + 2^64 hi (the caller's limbs when given, else the
// product), L = lo, H = hi; zero past the n live gates
__global__ __launch_bounds__(kT) void k_sp_tables(FqArgs f, const uint64_t* __restrict__ x,
                                                  const uint64_t* __restrict__ y, const uint64_t* __restrict__ lo_in,
                                                  const uint64_t* __restrict__ hi_in, uint64_t n, uint64_t len,
                                                  u128* __restrict__ A, u128* __restrict__ B, u128* __restrict__ C,
                                                  u128* __restrict__ L, u128* __restrict__ H) {
    const u128 q = f.Q(), qi = f.Qinv(), r2 = f.R2();
    for (uint64_t i = gtid(); i < len; i += gthreads()) {
        u128 a = 0, b = 0, c = 0, l = 0, h = 0;
        if (i < n) {
            const uint64_t xv = x[i], yv = y[i];
            uint64_t lv, hv;
            if (lo_in) {
                lv = lo_in[i];
                hv = hi_in[i];
            } else {
                const u128 p = (u128)xv * yv;
                lv = (uint64_t)p;
                hv = (uint64_t)(p >> 64);
            }
            // x R = mont(x, R^2 mod q), for any x < R
            a = fq_mont(xv, r2, q, qi);
            b = fq_mont(yv, r2, q, qi);
            c = fq_mont(mk(lv, hv), r2, q, qi);
            l = fq_mont(lv, r2, q, qi);
            h = fq_mont(hv, r2, q, qi);
        }
        A[i] = a;
        B[i] = b;
        C[i] = c;
        L[i] = l;
        H[i] = h;
    }
}

// eq(point, .) over n coordinates, coordinate i on index bit i
// (eq_le_table_fq_fast_with, bitify.rs:234-266): f[bit][i] = (1 - p_i) R or
// p_i R; Montgomery output, or canonical when canonical != 0; zero past live
constexpr int kMaxEq = 32;
struct EqArgs {
    uint64_t f[2][kMaxEq][2];
};
__global__ __launch_bounds__(kT) void k_sp_eq(FqArgs fa, EqArgs e, int n, uint64_t one_lo, uint64_t one_hi,
                                              int canonical, uint64_t live, u128* __restrict__ out) {
    const u128 q = fa.Q(), qi = fa.Qinv();
    const uint64_t count = 1ull << n;
    for (uint64_t x = gtid(); x < count; x += gthreads()) {
        u128 acc = mk(one_lo, one_hi);
        for (int i = 0; i < n; i++) {
            const int bit = (int)((x >> i) & 1);
            acc = fq_mont(acc, mk(e.f[bit][i][0], e.f[bit][i][1]), q, qi);
        }
        if (canonical) acc = fq_mont(acc, 1, q, qi);
        out[x] = x < live ? acc : (u128)0;
    }
}

// One outer round over the pairs (2p, 2p+1) of A, B, C, E: the evaluations
// at X = 0, 1, 2, 3 of sum_p E(X) (A(X) B(X) - C(X)), E the eq(tau) table.
// fold != 0: the inputs are the previous level (twice as long), folded by r
// (Montgomery) into the outputs first, pair entry by entry.
__global__ __launch_bounds__(kT) void k_sp_outer(FqArgs fa, const u128* __restrict__ A, const u128* __restrict__ B,
                                                 const u128* __restrict__ C, const u128* __restrict__ E,
                                                 uint64_t pairs, int fold, uint64_t r_lo, uint64_t r_hi,
                                                 u128* __restrict__ A2, u128* __restrict__ B2, u128* __restrict__ C2,
                                                 u128* __restrict__ E2, u128* __restrict__ partials) {
    const u128 q = fa.Q(), qi = fa.Qinv(), r = mk(r_lo, r_hi);
    u128 acc[4] = {0, 0, 0, 0};
    for (uint64_t p = gtid(); p < pairs; p += gthreads()) {
        u128 a[2], b[2], c[2], e[2];
        for (int j = 0; j < 2; j++) {
            const uint64_t i = 2 * p + j;
            if (fold) {
                const u128 a0 = A[2 * i], b0 = B[2 * i], c0 = C[2 * i], e0 = E[2 * i];
                a[j] = fq_add(a0, fq_mont(r, fq_sub(A[2 * i + 1], a0, q), q, qi), q);
                b[j] = fq_add(b0, fq_mont(r, fq_sub(B[2 * i + 1], b0, q), q, qi), q);
                c[j] = fq_add(c0, fq_mont(r, fq_sub(C[2 * i + 1], c0, q), q, qi), q);
                e[j] = fq_add(e0, fq_mont(r, fq_sub(E[2 * i + 1], e0, q), q, qi), q);
                A2[i] = a[j];
                B2[i] = b[j];
                C2[i] = c[j];
                E2[i] = e[j];
            } else {
                a[j] = A[i];
                b[j] = B[i];
                c[j] = C[i];
                e[j] = E[i];
            }
        }
        const u128 da = fq_sub(a[1], a[0], q), db = fq_sub(b[1], b[0], q), dc = fq_sub(c[1], c[0], q),
                   de = fq_sub(e[1], e[0], q);
        u128 av = a[0], bv = b[0], cv = c[0], ev = e[0];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k) {
                av = fq_add(av, da, q);
                bv = fq_add(bv, db, q);
                cv = fq_add(cv, dc, q);
                ev = fq_add(ev, de, q);
            }
            acc[k] = fq_add(acc[k], fq_mont(ev, fq_sub(fq_mont(av, bv, q, qi), cv, q), q, qi), q);
        }
    }
    block_sum_store<4>(acc, q, partials + 4 * (uint64_t)blockIdx.x);
}

// U = sum_b coef_b Z_b over the four value blocks (x, y, lo, hi), the inner
// sumcheck's combined assignment (its weights are coef_b eq(r_x, gate))
struct Coefs {
    uint64_t c[4][2];
};
__global__ __launch_bounds__(kT) void k_sp_combine(FqArgs fa, Coefs k, const u128* __restrict__ A,
                                                   const u128* __restrict__ B, const u128* __restrict__ L,
                                                   const u128* __restrict__ H, uint64_t len, u128* __restrict__ U) {
    const u128 q = fa.Q(), qi = fa.Qinv();
    const u128 c0 = mk(k.c[0][0], k.c[0][1]), c1 = mk(k.c[1][0], k.c[1][1]), c2 = mk(k.c[2][0], k.c[2][1]),
               c3 = mk(k.c[3][0], k.c[3][1]);
    for (uint64_t i = gtid(); i < len; i += gthreads()) {
        u128 u = fq_mont(c0, A[i], q, qi);
        u = fq_add(u, fq_mont(c1, B[i], q, qi), q);
        u = fq_add(u, fq_mont(c2, L[i], q, qi), q);
        u = fq_add(u, fq_mont(c3, H[i], q, qi), q);
        U[i] = u;
    }
}

// One inner round over the gate variables: c0 = sum E0 U0 and
// c2 = sum (E1 - E0)(U1 - U0) over the pairs, with the same fused fold
__global__ __launch_bounds__(kT) void k_sp_inner(FqArgs fa, const u128* __restrict__ E, const u128* __restrict__ U,
                                                 uint64_t pairs, int fold, uint64_t r_lo, uint64_t r_hi,
                                                 u128* __restrict__ E2, u128* __restrict__ U2,
                                                 u128* __restrict__ partials) {
    const u128 q = fa.Q(), qi = fa.Qinv(), r = mk(r_lo, r_hi);
    u128 acc[2] = {0, 0};
    for (uint64_t p = gtid(); p < pairs; p += gthreads()) {
        u128 e[2], u[2];
        for (int j = 0; j < 2; j++) {
            const uint64_t i = 2 * p + j;
            if (fold) {
                const u128 e0 = E[2 * i], u0 = U[2 * i];
                e[j] = fq_add(e0, fq_mont(r, fq_sub(E[2 * i + 1], e0, q), q, qi), q);
                u[j] = fq_add(u0, fq_mont(r, fq_sub(U[2 * i + 1], u0, q), q, qi), q);
                E2[i] = e[j];
                U2[i] = u[j];
            } else {
                e[j] = E[i];
                u[j] = U[i];
            }
        }
        acc[0] = fq_add(acc[0], fq_mont(e[0], u[0], q, qi), q);
        acc[1] = fq_add(acc[1], fq_mont(fq_sub(e[1], e[0], q), fq_sub(u[1], u[0], q), q, qi), q);
    }
    block_sum_store<2>(acc, q, partials + 2 * (uint64_t)blockIdx.x);
}

// sum_i eqg[i] Z_b[i] for the four value blocks
__global__ __launch_bounds__(kT) void k_sp_dot4(FqArgs fa, const u128* __restrict__ eqg, const u128* __restrict__ A,
                                                const u128* __restrict__ B, const u128* __restrict__ L,
                                                const u128* __restrict__ H, uint64_t len, u128* __restrict__ partials) {
    const u128 q = fa.Q(), qi = fa.Qinv();
    u128 acc[4] = {0, 0, 0, 0};
    for (uint64_t i = gtid(); i < len; i += gthreads()) {
        const u128 w = eqg[i];
        acc[0] = fq_add(acc[0], fq_mont(w, A[i], q, qi), q);
        acc[1] = fq_add(acc[1], fq_mont(w, B[i], q, qi), q);
        acc[2] = fq_add(acc[2], fq_mont(w, L[i], q, qi), q);
        acc[3] = fq_add(acc[3], fq_mont(w, H[i], q, qi), q);
    }
    block_sum_store<4>(acc, q, partials + 4 * (uint64_t)blockIdx.x);
}

// structured_row_weights (bitify.rs:268-318): row (slot << h) | gh =
// words[slot] eq_hi[gh] (words canonical, eq_hi Montgomery: the product is
// canonical)
__global__ __launch_bounds__(kT) void k_sp_row_weights(FqArgs fa, const u128* __restrict__ words,
                                                       const u128* __restrict__ eqhi, int h, uint64_t rows,
                                                       u128* __restrict__ out) {
    const u128 q = fa.Q(), qi = fa.Qinv();
    for (uint64_t i = gtid(); i < rows; i += gthreads()) {
        const u128 w = words[i >> h];
        out[i] = w ? fq_mont(w, eqhi[i & ((1ull << h) - 1)], q, qi) : (u128)0;
    }
}

unsigned blocks_for(uint64_t n) {
    const uint64_t b = (n + kT - 1) / kT;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 2048));
}
// sumcheck rounds: at most 256 blocks, so the partials are at most 16 KiB
unsigned round_blocks(uint64_t pairs) {
    const uint64_t b = (pairs + kT - 1) / kT;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 256));
}
// rounds over at most this many pairs run on the host
constexpr uint64_t kHostPairs = 128;

// ------------------------------------------------------------- host F_q
struct Fq {
    ModN m;
    u128 q = 0, inv2 = 0, inv6 = 0;
    FqArgs args{};
    void set(u128 q_) {
        q = q_;
        m.set(q);
        inv2 = powmod(2, q - 2, q);
        inv6 = powmod(6, q - 2, q);
        args.q[0] = (uint64_t)q;
        args.q[1] = (uint64_t)(q >> 64);
        args.qinv[0] = (uint64_t)m.ninv;
        args.qinv[1] = (uint64_t)(m.ninv >> 64);
        args.r2[0] = (uint64_t)m.r2;
        args.r2[1] = (uint64_t)(m.r2 >> 64);
    }
    u128 mul(u128 a, u128 b) const { return m.mul(a, b); }
    u128 add(u128 a, u128 b) const { return addmod(a % q, b % q, q); }
    u128 sub(u128 a, u128 b) const { return addmod(a % q, q - b % q, q); }
    u128 to_mont(u128 a) const { return m.to_mont(a); }
    u128 from_mont(u128 a) const { return m.mont(a, 1); }
};

void put64(std::vector<uint8_t>* o, uint64_t v) {
    for (int i = 0; i < 8; i++) o->push_back((uint8_t)(v >> (8 * i)));
}
void put128(std::vector<uint8_t>* o, u128 v) {
    put64(o, (uint64_t)v);
    put64(o, (uint64_t)(v >> 64));
}
void put_str(std::vector<uint8_t>* o, const char* s) { o->insert(o->end(), s, s + strlen(s)); }

// absorb_spartan_message (src/piop/spartan/mod.rs:349-358)
void spartan_message(Transcript& tr, const char* tag, const void* payload, size_t len) {
    std::vector<uint8_t> f;
    put_str(&f, kFrameDomain);
    put64(&f, strlen(tag));
    put_str(&f, tag);
    put64(&f, len);
    f.insert(f.end(), (const uint8_t*)payload, (const uint8_t*)payload + len);
    tr.absorb_slice(f.data(), f.size());
}
void spartan_u128(Transcript& tr, const char* tag, u128 v) {
    uint8_t b[16];
    u128_put(v, b);
    spartan_message(tr, tag, b, 16);
}
// absorb_field_elements (mod.rs:330-346)
void absorb_elements(Transcript& tr, const u128* v, size_t n) {
    std::vector<uint8_t> p;
    put64(&p, n);
    for (size_t i = 0; i < n; i++) {
        put64(&p, 16);
        put128(&p, v[i]);
    }
    spartan_message(tr, "field-elements", p.data(), p.size());
}
// squeeze_field (mod.rs:371-384): sample_below(q) over 256 draws, absorbed
bool squeeze(Transcript& tr, u128 q, u128* out) {
    if (!sample_below(tr, q, 256, out)) return false;
    absorb_elements(tr, out, 1);
    return true;
}

// ------------------------------------------------------------- security profile
struct SecParams {
    u128 pmin, pmax;
    uint32_t initial, piop, terminal, forest;
};
// derive_params (profile.rs:506-646), single derived prime, at the u64
// relation's facts (u64_bitz.rs:56-66: defect < 2^130, lift arity and opening
// t = the row variables, a direct opening, degree 3)
SecParams security_params(int t, int tau_arity) {
    SecParams p;
    const int width = std::min(std::min(113, 128 - t), 126 - t);
    p.pmin = (u128)1 << (width - 1);
    const u128 bit_max = ((u128)1 << width) - 1;
    const u128 no_wrap_max = 1 + (~(u128)0) / (((u128)1 << t) + 1);
    p.pmax = std::min(bit_max, no_wrap_max);
    auto ceil_pos = [](double v) { return v <= 0.0 ? 0u : (uint32_t)std::ceil(v); };
    const double b = width, ln2 = std::log(2.0);
    const double density = 1.0 / (b * ln2) - 0.63 / ((b - 1.0) * ln2);
    const int divisors = 130 / (width - 1);
    const double tau_raw = (width - 1) - std::log2((double)std::max(tau_arity, 1));
    double draw = tau_raw;
    if (divisors) draw = std::min(draw, b + std::log2(density) - std::log2((double)divisors));
    p.initial = ceil_pos(kLambda - draw);
    p.piop = ceil_pos(kLambda - ((width - 1) - std::log2(3.0)));
    p.terminal = p.piop;
    p.forest = kLambda <= 126 ? 0 : kLambda - 126;
    return p;
}

struct MulShape {
    uint64_t n, cap;
    int g, s, t;
};

// bind_assignment (binding.rs:185-208) with the u64 relation's middle
// section (u64_bitz.rs:135-171)
void assignment_binding(const uint8_t root[32], const LigConfig& cfg, const SecParams& sec, const MulShape& sh,
                        uint8_t out[32]) {
    std::vector<uint8_t> b;
    put_str(&b, kBindingDomain);
    put_str(&b, kBlockOrder);
    b.insert(b.end(), root, root + 32);
    put64(&b, cfg.m);
    put64(&b, cfg.levels[0].log_inv_rate);
    put64(&b, cfg.initial_k);
    b.push_back(0);  // LigeritoProfile::Fast
    b.push_back(1);  // HashKind::Blake3
    put128(&b, sec.pmin);
    put128(&b, sec.pmax);
    for (uint64_t v : {(uint64_t)kLambda, (uint64_t)kLambda, (uint64_t)sec.initial, (uint64_t)sec.piop,
                       (uint64_t)sec.terminal, (uint64_t)sec.forest})
        put64(&b, v);
    put128(&b, (u128)1 << 64);
    for (uint64_t v : {sh.n, sh.cap, 5 * sh.cap, 8 * sh.cap, (uint64_t)sh.g, (uint64_t)5, (uint64_t)8, (uint64_t)64,
                       (uint64_t)0, (uint64_t)64, (uint64_t)128, (uint64_t)192, (uint64_t)256, (uint64_t)sh.t,
                       (uint64_t)sh.s, (uint64_t)1})
        put64(&b, v);
    blake3_hash_host(b.data(), b.size(), out);
}

// ------------------------------------------------------------- matrix digest
// constraint_matrix_digest (matrix.rs:1478-1513) of the u64 matrices
// (u64_mul.rs:155-176, selector_matrix mul.rs:476-503). Only the modulus
// bytes inside the first chunk depend on q (the reference splices them into
// a cached stream the same way, matrix.rs:567-620), so per shape the stream
// after the first chunk is hashed once: the subtree chaining values of the
// whole chunks after the first and the bytes of the last chunk are cached.
struct MatrixStream {
    std::vector<uint8_t> first;  // chunk 0, modulus bytes zero
    uint64_t chunks = 0;         // total chunks
    std::vector<int> run_k;      // the aligned runs tiling chunks [1, chunks - 1)
    std::vector<std::array<uint32_t, 8>> run_cv;
    std::vector<uint8_t> last;   // the last chunk's bytes (chunks >= 2)
};
constexpr size_t kModHole = sizeof(kMatrixDomain) - 1 + 8;

void matrix_stream_bytes(uint64_t n, uint64_t cap, std::vector<uint8_t>* o) {
    o->clear();
    o->reserve(64 + 152 * n);
    put_str(o, kMatrixDomain);
    put64(o, 16);
    put128(o, 0);  // the modulus hole
    put64(o, n);
    put64(o, 5 * cap);
    const u128 one = 1, base = (u128)1 << 64;
    o->push_back('A');
    for (uint64_t i = 0; i < n; i++) {
        put64(o, 1);
        put64(o, cap + i);
        put64(o, 16);
        put128(o, one);
    }
    o->push_back('B');
    for (uint64_t i = 0; i < n; i++) {
        put64(o, 1);
        put64(o, 2 * cap + i);
        put64(o, 16);
        put128(o, one);
    }
    o->push_back('C');
    for (uint64_t i = 0; i < n; i++) {
        put64(o, 2);
        put64(o, 3 * cap + i);
        put64(o, 16);
        put128(o, one);
        put64(o, 4 * cap + i);
        put64(o, 16);
        put128(o, base);
    }
}

std::shared_ptr<MatrixStream> matrix_stream(uint64_t n, uint64_t cap) {
    static std::mutex mu;
    static std::map<std::pair<uint64_t, uint64_t>, std::shared_ptr<MatrixStream>> cache;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find({n, cap});
        if (it != cache.end()) return it->second;
    }
    auto ms = std::make_shared<MatrixStream>();
    std::vector<uint8_t> bytes;
    matrix_stream_bytes(n, cap, &bytes);
    const uint64_t len = bytes.size();
    ms->chunks = (len + 1023) / 1024;
    ms->first.assign(bytes.begin(), bytes.begin() + std::min<uint64_t>(len, 1024));
    if (ms->chunks >= 2) {
        const uint64_t whole = ms->chunks - 2;  // chunks [1, chunks - 1)
        ms->last.assign(bytes.begin() + (ms->chunks - 1) * 1024, bytes.end());
        std::vector<std::array<uint32_t, 8>> cvs(whole);
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nt; w++)
            th.emplace_back([&, w] {
                for (uint64_t c = w; c < whole; c += nt)
                    blake3_chunk_cv(bytes.data() + (c + 1) * 1024, c + 1, cvs[c].data());
            });
        for (auto& x : th) x.join();
        ms->run_k = blake3_subtree_runs(1, whole);
        uint64_t at = 0;
        for (int k : ms->run_k) {
            std::vector<std::array<uint32_t, 8>> lvl(cvs.begin() + at, cvs.begin() + at + (1ull << k));
            while (lvl.size() > 1) {
                std::vector<std::array<uint32_t, 8>> up(lvl.size() / 2);
                for (size_t i = 0; i < up.size(); i++) blake3_parent_cv(lvl[2 * i].data(), lvl[2 * i + 1].data(), up[i].data());
                lvl.swap(up);
            }
            ms->run_cv.push_back(lvl[0]);
            at += 1ull << k;
        }
    }
    std::lock_guard<std::mutex> lk(mu);
    cache[{n, cap}] = ms;
    return ms;
}

void matrix_digest(uint64_t n, uint64_t cap, u128 q, uint8_t out[32]) {
    auto ms = matrix_stream(n, cap);
    std::vector<uint8_t> first = ms->first;
    u128_put(q, first.data() + kModHole);
    Blake3 h;
    h.update(first.data(), first.size());
    for (size_t i = 0; i < ms->run_k.size(); i++) h.absorb_subtree(ms->run_cv[i].data(), ms->run_k[i]);
    if (!ms->last.empty()) h.update(ms->last.data(), ms->last.size());
    h.finalize_xof(out, 32);
}

// ------------------------------------------------------------- bitification
struct Bitified {
    std::vector<u128> gate, factors;  // gate point; per-block row factors (scaled)
    u128 col_scale = 0, claimed = 0;
};
// bitify (bitify.rs:119-219), scale on the rows
bool bitify(const Fq& F, const std::vector<u128>& ry, int g, u128 scale, u128 value, Bitified* b) {
    b->gate.assign(ry.begin(), ry.begin() + g);
    const std::vector<u128> sel(ry.begin() + g, ry.end());
    auto block_factor = [&](int code) {
        u128 acc = 1;
        for (size_t bit = 0; bit < sel.size(); bit++)
            acc = F.mul(acc, ((code >> bit) & 1) ? sel[bit] : F.sub(1, sel[bit]));
        return acc;
    };
    u128 cst = block_factor(0);
    for (u128 c : b->gate) cst = F.mul(cst, F.sub(1, c));
    b->claimed = F.sub(value, F.mul(scale, cst));
    std::vector<u128> f;
    bool all_zero = true;
    for (int code = 1; code <= 4; code++) {
        f.push_back(block_factor(code));
        all_zero = all_zero && f.back() == 0;
    }
    if (all_zero) return false;  // the constant-only point (probability ~ 4/q)
    if (scale == 0 || scale == 1) {
        b->factors = f;
        b->col_scale = scale;
    } else {
        for (auto& x : f) x = F.mul(scale, x);
        b->factors = f;
        b->col_scale = 1;
    }
    return true;
}

// bridge_digest (bitify.rs:365-410) with the u64 constants (u64_bitz.rs:173-202)
void bridge_digest(const uint8_t binding[32], u128 q, const uint8_t mdigest[32], const MulShape& sh,
                   const std::vector<u128>& ry, u128 scale, u128 value, const Bitified& b, uint8_t out[32]) {
    std::vector<uint8_t> h;
    put_str(&h, kBridgeDomain);
    h.insert(h.end(), binding, binding + 32);
    put64(&h, 16);
    put128(&h, q);
    h.insert(h.end(), mdigest, mdigest + 32);
    put128(&h, q);
    put128(&h, (u128)1 << 64);
    for (uint64_t v : {sh.n, sh.cap, 5 * sh.cap, 8 * sh.cap, (uint64_t)sh.g, (uint64_t)5, (uint64_t)8, (uint64_t)sh.t,
                       (uint64_t)sh.s, (uint64_t)1, (uint64_t)0, (uint64_t)64, (uint64_t)128, (uint64_t)192,
                       (uint64_t)64, (uint64_t)256})
        put64(&h, v);
    for (uint8_t v : {1, 0, 1, 2, 3, 4, 5, 6, 7}) h.push_back(v);
    put64(&h, ry.size());
    for (u128 c : ry) put128(&h, c);
    put128(&h, scale);
    put128(&h, value);
    put64(&h, sh.s);
    for (int i = 0; i < sh.s; i++) put128(&h, b.gate[i]);
    put64(&h, sh.g - sh.s);
    for (int i = sh.s; i < sh.g; i++) put128(&h, b.gate[i]);
    h.push_back(0);
    for (u128 f : b.factors) put128(&h, f);
    put128(&h, b.col_scale);
    put128(&h, b.claimed);
    blake3_hash_host(h.data(), h.size(), out);
}

// ------------------------------------------------------------- the shared driver
int mul_shape(uint64_t n, int split, MulShape* sh) {
    uint32_t lay[4];
    if (int rc = bitz_mul_layout(64, n, split, lay)) return rc;
    sh->n = n;
    sh->g = (int)lay[0];
    sh->s = (int)lay[1];
    sh->t = (int)lay[2];
    sh->cap = 1ull << sh->g;
    BITZ_REQUIRE(n > sh->cap / 2, BITZ_ERR_SHAPE,
                 "u64 multiplication proof: the batch must fill more than half its capacity (n > 128)");
    BITZ_REQUIRE(sh->t + sh->s >= 20 && sh->t + sh->s <= 35 && sh->t <= 24, BITZ_ERR_SHAPE,
                 "u64 multiplication proof: supported grids are 2^20..2^35 bits with t <= 24");
    return BITZ_OK;
}

struct Eqs {
    // eq(point, .) into out, coordinate i on bit i (Montgomery, or canonical)
    static int table(bitz_ctx* ctx, const Fq& F, const std::vector<u128>& pt, bool canonical, uint64_t live,
                     void* out) {
        const int n = (int)pt.size();
        BITZ_REQUIRE(n <= kMaxEq, BITZ_ERR_SHAPE, "eq table: more than 32 coordinates");
        EqArgs e{};
        for (int i = 0; i < n; i++) {
            const u128 f1 = F.to_mont(pt[i]), f0 = F.to_mont(F.sub(1, pt[i]));
            e.f[0][i][0] = (uint64_t)f0;
            e.f[0][i][1] = (uint64_t)(f0 >> 64);
            e.f[1][i][0] = (uint64_t)f1;
            e.f[1][i][1] = (uint64_t)(f1 >> 64);
        }
        const u128 one = F.to_mont(1);
        const uint64_t count = 1ull << n;
        hipLaunchKernelGGL(k_sp_eq, dim3(blocks_for(count)), dim3(kT), 0, ctx->stream, F.args, e, n, (uint64_t)one,
                           (uint64_t)(one >> 64), canonical ? 1 : 0, live, (u128*)out);
        BITZ_LAUNCH_CHECK();
        return BITZ_OK;
    }
};

// sums of nb blocks' K partials (Montgomery) -> canonical values
int partial_sums(bitz_ctx* ctx, const Fq& F, const DevBuf& part, unsigned nb, int K, u128* out) {
    const size_t bytes = (size_t)nb * K * 16;
    if (int rc = ctx->scratch.ensure_stage(bytes)) return rc;
    const u128* h = (const u128*)ctx->scratch.stage;
    BITZ_HIP(hipMemcpyAsync(ctx->scratch.stage, part.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < K; k++) {
        u128 acc = 0;
        for (unsigned b = 0; b < nb; b++) acc = addmod(acc, h[(size_t)b * K + k], F.q);
        out[k] = F.from_mont(acc);
    }
    return BITZ_OK;
}

// device tables (Montgomery) of n entries each, down to the host
int download_tables(bitz_ctx* ctx, u128* const* src, int k, uint64_t n, std::vector<u128>* out) {
    const size_t bytes = (size_t)k * n * 16;
    if (int rc = ctx->scratch.ensure_stage(bytes)) return rc;
    for (int i = 0; i < k; i++)
        BITZ_HIP(hipMemcpyAsync((uint8_t*)ctx->scratch.stage + (size_t)i * n * 16, src[i], n * 16,
                                hipMemcpyDeviceToHost, ctx->stream));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    const u128* h = (const u128*)ctx->scratch.stage;
    for (int i = 0; i < k; i++) out[i].assign(h + (size_t)i * n, h + (size_t)(i + 1) * n);
    return BITZ_OK;
}

// the host's fold of Montgomery tables by r (Montgomery)
void host_fold(const Fq& F, std::vector<u128>& v, u128 rm) {
    const size_t n = v.size() / 2;
    for (size_t i = 0; i < n; i++) {
        const u128 v0 = v[2 * i], v1 = v[2 * i + 1];
        v[i] = addmod(v0, F.m.mont(rm, addmod(v1, F.q - v0, F.q)), F.q);
    }
    v.resize(n);
}

u128 horner(const Fq& F, const u128* c, int n, u128 x) {
    u128 acc = 0;
    for (int i = n - 1; i >= 0; i--) acc = F.add(F.mul(acc, x), c[i]);
    return acc;
}

// the value of the sum over gates g < live of eq(a, g) eq(b, g), g of n bits
u128 eq_prefix_sum(const Fq& F, const std::vector<u128>& a, const std::vector<u128>& b, uint64_t live) {
    const int n = (int)a.size();
    std::vector<u128> h0(n), h1(n), full(n + 1);
    full[0] = 1;
    for (int i = 0; i < n; i++) {
        h0[i] = F.mul(F.sub(1, a[i]), F.sub(1, b[i]));
        h1[i] = F.mul(a[i], b[i]);
        full[i + 1] = F.mul(full[i], F.add(h0[i], h1[i]));
    }
    if (live >= (1ull << n)) return full[n];
    u128 total = 0, prefix = 1;
    for (int i = n - 1; i >= 0; i--) {
        if ((live >> i) & 1) {
            total = F.add(total, F.mul(F.mul(prefix, h0[i]), full[i]));
            prefix = F.mul(prefix, h1[i]);
        } else {
            prefix = F.mul(prefix, h0[i]);
        }
    }
    return total;
}

// The statement, Round 0 (prover side through round0), the session and
// ladder digest, the prime and the Spartan statement; both sides.
struct Prefix {
    MulShape sh;
    LigConfig cfg;
    SecParams sec;
    uint8_t binding[32], mdigest[32];
    u128 q;
    Fq F;
};
int bind_statement(Transcript& tr, const uint8_t root[32], Prefix* P) {
    assignment_binding(root, P->cfg, P->sec, P->sh, P->binding);
    spartan_message(tr, kStatementTag, P->binding, 32);
    tr.absorb_slice(std::string(kPolicyLabel));
    uint8_t d[32];
    ligerito_digest(P->cfg, d);
    tr.absorb_slice(d, 32);
    return BITZ_OK;
}
int bind_prime_and_statement(Transcript& tr, Prefix* P) {
    tr.absorb_slice(std::string(kSession));
    uint8_t d[32];
    ligerito_digest(P->cfg, d);
    tr.absorb_slice(d, 32);
    spartan_message(tr, "prime-domain", kPrimeDomain, strlen(kPrimeDomain));
    spartan_u128(tr, "prime-min", P->sec.pmin);
    spartan_u128(tr, "prime-max", P->sec.pmax);
    std::string err;
    if (!sample_prime_context(tr, P->sec.pmin, P->sec.pmax, 128, &P->q, &err)) {
        set_error("u64 multiplication proof: " + err);
        return BITZ_ERR_VERIFY;
    }
    spartan_u128(tr, "prime-q", P->q);
    P->F.set(P->q);
    matrix_digest(P->sh.n, P->sh.cap, P->q, P->mdigest);
    spartan_message(tr, "protocol", kPiopDomain, strlen(kPiopDomain));
    spartan_u128(tr, "field-modulus", P->q);
    spartan_message(tr, "matrix-statement", P->mdigest, 32);
    spartan_message(tr, kOracleDomain, P->binding, 32);
    return BITZ_OK;
}

int make_prefix(uint64_t n, int split, Prefix* P) {
    if (int rc = mul_shape(n, split, &P->sh)) return rc;
    std::string err = custom_johnson_config(P->sh.t + P->sh.s, kRate, kBatch, kLambda, &P->cfg);
    BITZ_REQUIRE(err.empty(), BITZ_ERR_SHAPE, "Ligerito configuration: " + err);
    P->sec = security_params(P->sh.t, P->sh.g);
    BITZ_REQUIRE(P->sec.initial == 0 && P->sec.piop == 0 && P->sec.terminal == 0, BITZ_ERR_SHAPE,
                 "u64 multiplication proof: shapes with grinding boundaries are not supported");
    return BITZ_OK;
}

// row weights (canonical, device) and column weights of the bitified claim
int claim_weights(bitz_ctx* ctx, const Prefix& P, const Bitified& b, DevBuf& roww, DevBuf& colw) {
    const int h = P.sh.g - P.sh.s;
    const uint64_t rows = 1ull << P.sh.t, cols = 1ull << P.sh.s;
    std::vector<u128> words(256, 0);
    for (int blk = 0; blk < 4; blk++) {
        u128 sc = b.factors[blk];
        for (int k = 0; k < 64; k++) {
            words[64 * blk + k] = sc;
            sc = P.F.add(sc, sc);
        }
    }
    DevBuf wd, eqhi;
    if (!roww.alloc(rows * 16) || !colw.alloc(cols * 16) || !wd.alloc(256 * 16) || !eqhi.alloc(16ull << h))
        return BITZ_ERR_ALLOC;
    if (int rc = upload_async(ctx, words.data(), 256 * 16, wd.ptr)) return rc;
    std::vector<u128> hi(b.gate.begin() + P.sh.s, b.gate.end()), lo(b.gate.begin(), b.gate.begin() + P.sh.s);
    if (int rc = Eqs::table(ctx, P.F, hi, false, ~0ull, eqhi.ptr)) return rc;
    hipLaunchKernelGGL(k_sp_row_weights, dim3(blocks_for(rows)), dim3(kT), 0, ctx->stream, P.F.args,
                       (const u128*)wd.ptr, (const u128*)eqhi.ptr, h, rows, (u128*)roww.ptr);
    BITZ_LAUNCH_CHECK();
    if (b.col_scale == 0) {
        BITZ_HIP(hipMemsetAsync(colw.ptr, 0, cols * 16, ctx->stream));
    } else {
        if (int rc = Eqs::table(ctx, P.F, lo, true, ~0ull, colw.ptr)) return rc;
    }
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    return BITZ_OK;
}

// the public weight at r_y: sum_b coef_b eq(sel, b), times
// sum_{gate < n} eq(r_x, gate) eq(r_gate, gate)
u128 public_weight(const Fq& F, const std::vector<u128>& rx, const std::vector<u128>& ry, int g, u128 rho,
                   uint64_t n) {
    const u128 rho2 = F.mul(rho, rho);
    const u128 coef[8] = {0, 1, rho, rho2, F.mul(rho2, F.m.mul((u128)1 << 64, 1)), 0, 0, 0};
    u128 ws = 0;
    for (int code = 0; code < 8; code++) {
        u128 f = coef[code];
        for (int bit = 0; bit < 3; bit++) f = F.mul(f, ((code >> bit) & 1) ? ry[g + bit] : F.sub(1, ry[g + bit]));
        ws = F.add(ws, f);
    }
    std::vector<u128> gate(ry.begin(), ry.begin() + g);
    return F.mul(ws, eq_prefix_sum(F, rx, gate, n));
}

// the fork tag: two 16-byte challenges of the outer transcript
void fork_tag(Transcript& tr, uint8_t tag[32]) {
    tr.challenge(tag, 16);
    tr.challenge(tag + 16, 16);
}

bool squeeze_or_fail(Transcript& tr, u128 q, u128* out) {
    if (squeeze(tr, q, out)) return true;
    set_error("u64 multiplication proof: sampling exhausted");
    return false;
}

void put_elems(std::vector<uint8_t>* payload, const u128* v, int n) {
    for (int k = 0; k < n; k++) {
        uint8_t b[16];
        u128_put(v[k], b);
        payload->insert(payload->end(), b, b + 16);
    }
}

// the cubic through (k, e_k), k = 0..3
void cubic_coeffs(const Fq& F, const u128 e[4], u128 c[4]) {
    c[0] = e[0];
    c[3] = F.mul(F.sub(F.add(F.sub(e[3], F.mul(3, e[2])), F.mul(3, e[1])), e[0]), F.inv6);
    c[2] = F.mul(F.add(F.sub(F.sub(F.mul(4, e[2]), e[3]), F.mul(5, e[1])), F.mul(2, e[0])), F.inv2);
    c[1] = F.mul(F.sub(F.add(F.sub(F.mul(2, e[3]), F.mul(9, e[2])), F.mul(18, e[1])), F.mul(11, e[0])), F.inv6);
}

// The three selector rounds of the inner sumcheck on the eight blocks:
// W_b = coef_b E(r_gate) (zero for blocks 0, 5, 6, 7), Z_b(r_gate).
template <class Msg>
bool selector_rounds(const Fq& F, Transcript& tr, u128 w[8], u128 z[8], u128* claim, std::vector<u128>* ry,
                     Msg&& msg) {
    int len = 8;
    for (int round = 0; round < 3; round++) {
        u128 c0 = 0, c2 = 0;
        for (int i = 0; i < len; i += 2) {
            c0 = F.add(c0, F.mul(w[i], z[i]));
            c2 = F.add(c2, F.mul(F.sub(w[i + 1], w[i]), F.sub(z[i + 1], z[i])));
        }
        u128 c[3] = {c0, F.sub(F.sub(*claim, F.add(c0, c0)), c2), c2};
        if (!msg(c)) return false;
        u128 r;
        if (!squeeze_or_fail(tr, F.q, &r)) return false;
        *claim = horner(F, c, 3, r);
        ry->push_back(r);
        for (int i = 0; i < len / 2; i++) {
            w[i] = F.add(w[2 * i], F.mul(r, F.sub(w[2 * i + 1], w[2 * i])));
            z[i] = F.add(z[2 * i], F.mul(r, F.sub(z[2 * i + 1], z[2 * i])));
        }
        len /= 2;
    }
    return true;
}

int prove_mul_u64(bitz_ctx* ctx, const uint64_t* x, const uint64_t* y, const uint64_t* lo, const uint64_t* hi,
                  uint64_t n, int split, uint8_t root[32], uint8_t state[32], std::vector<uint8_t>* proof) {
    int rc;
    Prefix P;
    if ((rc = make_prefix(n, split, &P))) return rc;
    const MulShape& sh = P.sh;
    const int g = sh.g;
    const uint64_t len = sh.cap;
    PhaseClock phase(ctx);
    phase.mark("mul_commit");
    // the committed grid (MulWitness::bitz_bit_rows) and its commitment
    std::unique_ptr<bitz_commit_handle, void (*)(bitz_commit_handle*)> h(nullptr, bitz_commit_free);
    {
        DevBuf rows;
        const uint64_t words = 1ull << (sh.t + sh.s - 6);
        if (!rows.alloc(words * 8)) return BITZ_ERR_ALLOC;
        if ((rc = bitz_mul_bit_rows_dev(ctx, 64, n, split, x, y, lo, hi, rows.ptr, words))) return rc;
        bitz_commit_handle* hp = nullptr;
        if ((rc = bitz_commit_device_rows(ctx, rows.ptr, sh.t, sh.s, kBatch, kRate, root, &hp))) return rc;
        h.reset(hp);
    }
    phase.mark("mul_statement");
    Transcript tr;
    bind_statement(tr, root, &P);
    Round0Claim r0;
    if ((rc = round0_prove(ctx, h.get(), tr, &r0))) return rc;
    if ((rc = bind_prime_and_statement(tr, &P))) return rc;
    const Fq& F = P.F;
    phase.mark("spartan_outer");
    std::vector<u128> tau(g);
    for (auto& v : tau)
        if (!squeeze_or_fail(tr, P.q, &v)) return BITZ_ERR_DEVICE;
    // base tables (len entries): A, B, C, E (outer), L, H; two half-length
    // sets ping-pong the folds
    DevBuf T0[6], T1[4], T2[4], part;
    for (auto& b : T0)
        if (!b.alloc(len * 16)) return BITZ_ERR_ALLOC;
    for (int k = 0; k < 4; k++)
        if (!T1[k].alloc(len * 8) || !T2[k].alloc(len * 8)) return BITZ_ERR_ALLOC;
    if (!part.alloc(2048 * 4 * 16)) return BITZ_ERR_ALLOC;
    u128 *A = (u128*)T0[0].ptr, *B = (u128*)T0[1].ptr, *C = (u128*)T0[2].ptr, *E = (u128*)T0[3].ptr,
         *L = (u128*)T0[4].ptr, *H = (u128*)T0[5].ptr;
    hipLaunchKernelGGL(k_sp_tables, dim3(blocks_for(len)), dim3(kT), 0, ctx->stream, F.args, x, y, lo, hi, n, len, A,
                       B, C, L, H);
    BITZ_LAUNCH_CHECK();
    if ((rc = Eqs::table(ctx, F, tau, false, ~0ull, E))) return rc;
    // round j reads src (the base set for j <= 1) and, for j >= 1, folds it
    // by r_(j-1) into dst (T1, T2 alternating) before its sums
    auto set_ptrs = [](DevBuf* s, u128** o, int k) {
        for (int i = 0; i < k; i++) o[i] = (u128*)s[i].ptr;
    };
    std::vector<uint8_t> payload;
    std::vector<u128> rx;
    u128 claim = 0, rprev = 0;
    u128 *src[4] = {A, B, C, E}, *dst[4];
    int j = 0;
    for (; j < g; j++) {
        const uint64_t pairs = len >> (j + 1);
        if (pairs <= kHostPairs) break;
        const unsigned nb = round_blocks(pairs);
        const u128 rm = F.to_mont(rprev);
        set_ptrs((j & 1) ? T1 : T2, dst, 4);
        hipLaunchKernelGGL(k_sp_outer, dim3(nb), dim3(kT), 0, ctx->stream, F.args, src[0], src[1], src[2], src[3],
                           pairs, j > 0 ? 1 : 0, (uint64_t)rm, (uint64_t)(rm >> 64), dst[0], dst[1], dst[2], dst[3],
                           (u128*)part.ptr);
        BITZ_LAUNCH_CHECK();
        if (j > 0)
            for (int k = 0; k < 4; k++) src[k] = dst[k];
        u128 e[4], c[4];
        if ((rc = partial_sums(ctx, F, part, nb, 4, e))) return rc;
        cubic_coeffs(F, e, c);
        absorb_elements(tr, c, 4);
        put_elems(&payload, c, 4);
        u128 r;
        if (!squeeze_or_fail(tr, P.q, &r)) return BITZ_ERR_DEVICE;
        claim = horner(F, c, 4, r);
        rx.push_back(r);
        rprev = r;
    }
    // the last rounds on the host: the tables (not yet folded by r_(j-1))
    // come down once
    std::vector<u128> ht[4];
    if ((rc = download_tables(ctx, src, 4, len >> (j > 0 ? j - 1 : 0), ht))) return rc;
    if (j > 0)
        for (auto& v : ht) host_fold(F, v, F.to_mont(rprev));
    for (; j < g; j++) {
        u128 e[4] = {0, 0, 0, 0}, c[4];
        for (size_t i = 0; i + 1 < ht[0].size(); i += 2) {
            u128 d[4], v[4];
            for (int k = 0; k < 4; k++) {
                v[k] = ht[k][i];
                d[k] = addmod(ht[k][i + 1], F.q - ht[k][i], F.q);
            }
            for (int x = 0; x < 4; x++) {
                if (x)
                    for (int k = 0; k < 4; k++) v[k] = addmod(v[k], d[k], F.q);
                const u128 ab = F.m.mont(v[0], v[1]);
                e[x] = addmod(e[x], F.m.mont(v[3], addmod(ab, F.q - v[2], F.q)), F.q);
            }
        }
        for (auto& x : e) x = F.from_mont(x);
        cubic_coeffs(F, e, c);
        absorb_elements(tr, c, 4);
        put_elems(&payload, c, 4);
        u128 r;
        if (!squeeze_or_fail(tr, P.q, &r)) return BITZ_ERR_DEVICE;
        claim = horner(F, c, 4, r);
        rx.push_back(r);
        rprev = r;
        for (auto& v : ht) host_fold(F, v, F.to_mont(r));
    }
    u128 fin[4];
    for (int k = 0; k < 4; k++) fin[k] = F.from_mont(ht[k][0]);
    // an honest witness satisfies the terminal check; a dishonest one (caller
    // limbs) is proved anyway, for the verifier to reject
    if (lo == nullptr && claim != F.mul(fin[3], F.sub(F.mul(fin[0], fin[1]), fin[2]))) {
        set_error("u64 multiplication proof: outer sumcheck terminal mismatch");
        return BITZ_ERR_DEVICE;
    }
    absorb_elements(tr, fin, 3);
    put_elems(&payload, fin, 3);
    u128 rho;
    if (!squeeze_or_fail(tr, P.q, &rho)) return BITZ_ERR_DEVICE;
    phase.mark("spartan_inner");
    // inner sumcheck: the gate rounds on E = eq(r_x, .) (live gates) and
    // U = sum_b coef_b Z_b (coef = 1, rho, rho^2, rho^2 2^64)
    const u128 rho2 = F.mul(rho, rho);
    const u128 coef[4] = {1, rho, rho2, F.mul(rho2, F.m.mul((u128)1 << 64, 1))};
    Coefs kc{};
    for (int b = 0; b < 4; b++) {
        const u128 mc = F.to_mont(coef[b]);
        kc.c[b][0] = (uint64_t)mc;
        kc.c[b][1] = (uint64_t)(mc >> 64);
    }
    u128* U = C;  // C is no longer needed
    if ((rc = Eqs::table(ctx, F, rx, false, n, E))) return rc;
    hipLaunchKernelGGL(k_sp_combine, dim3(blocks_for(len)), dim3(kT), 0, ctx->stream, F.args, kc, A, B, L, H, len, U);
    BITZ_LAUNCH_CHECK();
    u128 iclaim = F.add(F.add(fin[0], F.mul(rho, fin[1])), F.mul(rho2, fin[2]));
    std::vector<u128> ry;
    u128 *isrc[2] = {E, U}, *idst[2];
    j = 0;
    for (; j < g; j++) {
        const uint64_t pairs = len >> (j + 1);
        if (pairs <= kHostPairs) break;
        const unsigned nb = round_blocks(pairs);
        const u128 rm = F.to_mont(rprev);
        set_ptrs((j & 1) ? T1 : T2, idst, 2);
        hipLaunchKernelGGL(k_sp_inner, dim3(nb), dim3(kT), 0, ctx->stream, F.args, isrc[0], isrc[1], pairs,
                           j > 0 ? 1 : 0, (uint64_t)rm, (uint64_t)(rm >> 64), idst[0], idst[1], (u128*)part.ptr);
        BITZ_LAUNCH_CHECK();
        if (j > 0)
            for (int k = 0; k < 2; k++) isrc[k] = idst[k];
        u128 cs[2];
        if ((rc = partial_sums(ctx, F, part, nb, 2, cs))) return rc;
        u128 c[3] = {cs[0], F.sub(F.sub(iclaim, F.add(cs[0], cs[0])), cs[1]), cs[1]};
        absorb_elements(tr, c, 3);
        put_elems(&payload, c, 3);
        u128 r;
        if (!squeeze_or_fail(tr, P.q, &r)) return BITZ_ERR_DEVICE;
        iclaim = horner(F, c, 3, r);
        ry.push_back(r);
        rprev = r;
    }
    std::vector<u128> hi2[2];
    if ((rc = download_tables(ctx, isrc, 2, len >> (j > 0 ? j - 1 : 0), hi2))) return rc;
    if (j > 0)
        for (auto& v : hi2) host_fold(F, v, F.to_mont(rprev));
    for (; j < g; j++) {
        u128 c0 = 0, c2 = 0;
        for (size_t i = 0; i + 1 < hi2[0].size(); i += 2) {
            const u128 e0 = hi2[0][i], u0 = hi2[1][i];
            c0 = addmod(c0, F.m.mont(e0, u0), F.q);
            c2 = addmod(c2, F.m.mont(addmod(hi2[0][i + 1], F.q - e0, F.q), addmod(hi2[1][i + 1], F.q - u0, F.q)), F.q);
        }
        c0 = F.from_mont(c0);
        c2 = F.from_mont(c2);
        u128 c[3] = {c0, F.sub(F.sub(iclaim, F.add(c0, c0)), c2), c2};
        absorb_elements(tr, c, 3);
        put_elems(&payload, c, 3);
        u128 r;
        if (!squeeze_or_fail(tr, P.q, &r)) return BITZ_ERR_DEVICE;
        iclaim = horner(F, c, 3, r);
        ry.push_back(r);
        rprev = r;
        for (auto& v : hi2) host_fold(F, v, F.to_mont(r));
    }
    const u128 e_rg = F.from_mont(hi2[0][0]);
    // Z_b(r_gate) for the value blocks; the constant block is 1 at gate 0
    u128 z[8] = {1, 0, 0, 0, 0, 0, 0, 0}, w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (u128 c : ry) z[0] = F.mul(z[0], F.sub(1, c));
    if ((rc = Eqs::table(ctx, F, ry, false, ~0ull, E))) return rc;
    {
        const unsigned nb = blocks_for(len);
        hipLaunchKernelGGL(k_sp_dot4, dim3(nb), dim3(kT), 0, ctx->stream, F.args, E, A, B, L, H, len,
                           (u128*)part.ptr);
        BITZ_LAUNCH_CHECK();
        if ((rc = partial_sums(ctx, F, part, nb, 4, z + 1))) return rc;
    }
    for (int b = 0; b < 4; b++) w[b + 1] = F.mul(coef[b], e_rg);
    if (!selector_rounds(F, tr, w, z, &iclaim, &ry, [&](const u128* c) {
            absorb_elements(tr, c, 3);
            put_elems(&payload, c, 3);
            return true;
        }))
        return BITZ_ERR_DEVICE;
    if (iclaim != F.mul(w[0], z[0])) {
        set_error("u64 multiplication proof: inner sumcheck terminal mismatch");
        return BITZ_ERR_DEVICE;
    }
    for (auto& b : T1) b = DevBuf();
    for (auto& b : T2) b = DevBuf();
    for (auto& b : T0) b = DevBuf();
    // bitification, the bridge digest and the fork tag
    phase.mark("mul_bitify");
    const u128 scale = w[0];
    Bitified bf;
    if (!bitify(F, ry, g, scale, iclaim, &bf)) {
        set_error("u64 multiplication proof: constant-only claim point");
        return BITZ_ERR_DEVICE;
    }
    uint8_t bridge[32], tag[32];
    bridge_digest(P.binding, P.q, P.mdigest, sh, ry, scale, iclaim, bf, bridge);
    fork_tag(tr, tag);
    tr.digest(state);
    DevBuf roww, colw;
    if ((rc = claim_weights(ctx, P, bf, roww, colw))) return rc;
    std::vector<uint8_t> opening;
    if ((rc = open_linear_claim(ctx, h.get(), kSession, tag, bridge, 32, P.q, roww.ptr, colw.ptr, bf.claimed, r0,
                                &opening)))
        return rc;
    proof->swap(payload);
    proof->insert(proof->end(), opening.begin(), opening.end());
    return BITZ_OK;
}

int verify_mul_u64(bitz_ctx* ctx, const uint8_t root[32], uint64_t n, int split, const uint8_t* proof, size_t plen,
                   uint8_t state[32]) {
    int rc;
    Prefix P;
    if ((rc = make_prefix(n, split, &P))) return rc;
    const MulShape& sh = P.sh;
    const int g = sh.g;
    auto fail = [](const char* why) {
        set_error(std::string("verify: ") + why);
        return BITZ_ERR_VERIFY;
    };
    const size_t spartan_len = 16 * ((size_t)4 * g + 3 + 3 * ((size_t)g + 3));
    if (plen < spartan_len) return fail("truncated proof");
    const uint8_t* sp = proof;
    // the opening tail: narg, hints, Round 0 (BitZOpeningProof::from_bytes)
    const uint8_t* op = proof + spartan_len;
    const size_t olen = plen - spartan_len;
    size_t pos = 0;
    auto rd64 = [&](uint64_t* v) {
        if (pos + 8 > olen) return false;
        memcpy(v, op + pos, 8);
        pos += 8;
        return true;
    };
    uint64_t narg_len, hints_len;
    if (!rd64(&narg_len) || narg_len > olen - pos) return fail("truncated proof");
    const uint8_t* narg = op + pos;
    pos += narg_len;
    if (!rd64(&hints_len) || hints_len > olen - pos) return fail("truncated proof");
    const uint8_t* hints = op + pos;
    pos += hints_len;
    if (pos + 17 > olen || op[pos] != 1) return fail("missing Round-0 claim");
    const uint8_t* yb = op + pos + 1;
    pos += 17;
    if (pos >= olen || op[pos] > 1) return fail("truncated proof");
    const bool has_nonce = op[pos++] == 1;
    uint64_t nonce = 0;
    if (has_nonce && !rd64(&nonce)) return fail("truncated proof");
    if (pos != olen) return fail("trailing proof bytes");
    // statement and Round 0
    Transcript tr;
    bind_statement(tr, root, &P);
    const int mp = sh.t + sh.s - 7;
    const uint32_t bits = (uint32_t)P.cfg.round0_grinding_bits;
    ood_round_header(tr, mp, bits);
    if (bits) {
        if (!has_nonce) return fail("missing Round-0 nonce");
        uint8_t seed[32];
        grinding_seed(tr, "bitz/core/ood-round-grinding/v1", 0, bits, seed);
        const bool ok = round0_pow_valid(seed, nonce, bits);
        absorb_grinding_nonce(tr, nonce);
        if (!ok) return fail("Round-0 proof of work");
    } else if (has_nonce) {
        return fail("unexpected Round-0 nonce");
    }
    uint8_t zb[16];
    tr.challenge(zb, 16);
    std::vector<hgf> opt(mp);
    {
        hgf cur;
        memcpy(&cur.lo, zb, 8);
        memcpy(&cur.hi, zb + 8, 8);
        for (int i = 0; i < mp; i++) {
            opt[i] = cur;
            cur = hgf_mul(cur, cur);
        }
    }
    {
        uint8_t tagged[17];
        tagged[0] = 0x50;
        memcpy(tagged + 1, yb, 16);
        tr.absorb_slice(tagged, 17);
    }
    hgf ood_y;
    memcpy(&ood_y.lo, yb, 8);
    memcpy(&ood_y.hi, yb + 8, 8);
    if ((rc = bind_prime_and_statement(tr, &P))) return rc == BITZ_ERR_VERIFY ? fail("prime search") : rc;
    const Fq& F = P.F;
    size_t at = 0;
    auto elem = [&](u128* v) {
        *v = u128_from(sp + at);
        at += 16;
        return *v < F.q;
    };
    std::vector<u128> tau(g), rx, ry;
    for (auto& v : tau)
        if (!squeeze_or_fail(tr, P.q, &v)) return BITZ_ERR_VERIFY;
    u128 claim = 0;
    for (int j = 0; j < g; j++) {
        u128 c[4];
        for (auto& v : c)
            if (!elem(&v)) return fail("non-canonical field element");
        if (F.add(c[0], F.add(F.add(c[0], c[1]), F.add(c[2], c[3]))) != claim) return fail("outer sumcheck round");
        absorb_elements(tr, c, 4);
        u128 r;
        if (!squeeze_or_fail(tr, P.q, &r)) return BITZ_ERR_VERIFY;
        claim = horner(F, c, 4, r);
        rx.push_back(r);
    }
    u128 abc[3];
    for (auto& v : abc)
        if (!elem(&v)) return fail("non-canonical field element");
    u128 eqtr = 1;
    for (int i = 0; i < g; i++)
        eqtr = F.mul(eqtr, F.add(F.mul(F.sub(1, tau[i]), F.sub(1, rx[i])), F.mul(tau[i], rx[i])));
    if (claim != F.mul(eqtr, F.sub(F.mul(abc[0], abc[1]), abc[2]))) return fail("outer sumcheck terminal");
    absorb_elements(tr, abc, 3);
    u128 rho;
    if (!squeeze_or_fail(tr, P.q, &rho)) return BITZ_ERR_VERIFY;
    const u128 rho2 = F.mul(rho, rho);
    claim = F.add(F.add(abc[0], F.mul(rho, abc[1])), F.mul(rho2, abc[2]));
    for (int j = 0; j < g + 3; j++) {
        u128 c[3];
        for (auto& v : c)
            if (!elem(&v)) return fail("non-canonical field element");
        if (F.add(F.add(c[0], c[0]), F.add(c[1], c[2])) != claim) return fail("inner sumcheck round");
        absorb_elements(tr, c, 3);
        u128 r;
        if (!squeeze_or_fail(tr, P.q, &r)) return BITZ_ERR_VERIFY;
        claim = horner(F, c, 3, r);
        ry.push_back(r);
    }
    const u128 scale = public_weight(F, rx, ry, g, rho, n);
    Bitified bf;
    if (!bitify(F, ry, g, scale, claim, &bf)) return fail("constant-only claim point");
    uint8_t bridge[32], tag[32];
    bridge_digest(P.binding, P.q, P.mdigest, sh, ry, scale, claim, bf, bridge);
    fork_tag(tr, tag);
    if (state) tr.digest(state);
    DevBuf roww, colw;
    if ((rc = claim_weights(ctx, P, bf, roww, colw))) return rc;
    return verify_linear_claim(ctx, root, sh.t, sh.s, kRate, kBatch, kSession, tag, bridge, 32, P.q, roww.ptr,
                               colw.ptr, bf.claimed, opt, ood_y, narg, narg_len, hints, hints_len);
}

}  // namespace

void spartan_frame(Transcript& tr, const char* tag, const void* payload, size_t len) {
    spartan_message(tr, tag, payload, len);
}
bool spartan_squeeze(Transcript& tr, u128 q, u128* out) { return squeeze(tr, q, out); }

}  // namespace bitz

extern "C" int bitz_prove_mul_u64(bitz_ctx* ctx, const void* x, const void* y, const void* lo, const void* hi,
                                  uint64_t n, int32_t split_shift, uint8_t root[32], uint8_t state[32],
                                  uint8_t* proof, uint64_t capacity, uint64_t* proof_len) {
    BITZ_REQUIRE(ctx && x && y && root && proof_len, BITZ_ERR_ARG, "bitz_prove_mul_u64: null argument");
    BITZ_REQUIRE((lo == nullptr) == (hi == nullptr), BITZ_ERR_ARG,
                 "bitz_prove_mul_u64: supply both product limbs or neither");
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    std::vector<uint8_t> bytes;
    uint8_t st[32];
    int rc = bitz::prove_mul_u64(ctx, (const uint64_t*)x, (const uint64_t*)y, (const uint64_t*)lo,
                                 (const uint64_t*)hi, n, split_shift, root, st, &bytes);
    if (rc) return rc;
    if (state) memcpy(state, st, 32);
    *proof_len = bytes.size();
    BITZ_REQUIRE(proof && capacity >= bytes.size(), BITZ_ERR_ARG, "bitz_prove_mul_u64: proof buffer too small");
    memcpy(proof, bytes.data(), bytes.size());
    return BITZ_OK;
}

extern "C" int bitz_verify_mul_u64(bitz_ctx* ctx, const uint8_t root[32], uint64_t n, int32_t split_shift,
                                   const uint8_t* proof, uint64_t proof_len, uint8_t state[32]) {
    BITZ_REQUIRE(ctx && root && (proof || proof_len == 0), BITZ_ERR_ARG, "bitz_verify_mul_u64: null argument");
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    return bitz::verify_mul_u64(ctx, root, n, split_shift, proof, (size_t)proof_len, state);
}
