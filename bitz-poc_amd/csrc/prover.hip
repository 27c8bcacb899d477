// The BitZ standalone opening, orchestrated on the host, every bulk pass on
// the GPU. Follows src/piop/spartan/protocol/bitz_opener.rs:672-707
// (prove_standalone) -> src/bitz/mod.rs:142-199 (BitZProver::prove) ->
// src/bitz/pcs.rs:311-507 (prove_lin, prove_mle) ->
// vendor/flock-mod/crates/flock-core/src/pcs/ligerito.rs:3398-3964
// (recursive_prover_with_basis).
//
// Transcripts: the outer statement/claim transcript is the reference's
// Blake3Transcript (src/transcript/mod.rs). The inner BitZ transcript keeps
// the reference's message order, labels and encodings but runs them through
// the same BLAKE3 framing (the reference's spongefish 0.7.4 sponge is not
// available to restate).
#include <algorithm>
#include <array>
#include <set>
#include <thread>

#include <optional>
#include "host_gf.h"
#include "host_proto.h"
#include "modq.h"
#include "open_impl.h"
#include "residual.h"
#include "spartan_impl.h"

namespace bitz {
namespace {

// ------------------------------------------------------------- host values
using G = hgf;
inline G gzero() { return G{0, 0}; }
inline G gone() { return G{1, 0}; }
inline bool geq(G a, G b) { return a.lo == b.lo && a.hi == b.hi; }
inline G gfrom(const uint8_t* p) {
    G g;
    memcpy(&g.lo, p, 8);
    memcpy(&g.hi, p + 8, 8);
    return g;
}
inline void gput(G g, uint8_t* p) {
    memcpy(p, &g.lo, 8);
    memcpy(p + 8, &g.hi, 8);
}
// eq_factor (src/bitz/mod.rs:328-331): r z + (1 - r)(1 - z)
inline G eq_factor(G r, G z) { return hgf_add(hgf_mul(r, z), hgf_mul(hgf_add(r, gone()), hgf_add(z, gone()))); }

// A GKR layer's rounds with at most 2^host_tail_log slots (ctx->tune, 10 by
// default) run on the host (host_tail): one such round is a few microseconds
// of carry-less multiplies there, against a kernel launch, a reduction and a
// device-to-host round trip (tens of microseconds) on the device.

// out[i] = v[i]^-1 with one inversion (prefix products; Montgomery's
// trick); a zero entry maps to zero, as hgf_inv does
void batch_inverse(const std::vector<G>& v, std::vector<G>* out) {
    const size_t n = v.size();
    out->assign(n, gzero());
    std::vector<G> pre(n);
    G acc = gone();
    for (size_t i = 0; i < n; i++) {
        pre[i] = acc;
        if (!geq(v[i], gzero())) acc = hgf_mul(acc, v[i]);
    }
    G cur = hgf_inv(acc);
    for (size_t i = n; i-- > 0;) {
        if (geq(v[i], gzero())) continue;
        (*out)[i] = hgf_mul(pre[i], cur);
        cur = hgf_mul(cur, v[i]);
    }
}

std::vector<G> host_eq(const std::vector<G>& r) {
    std::vector<G> t(1ull << r.size());
    t[0] = gone();
    for (size_t i = 0; i < r.size(); i++) {
        size_t len = 1ull << i;
        for (size_t x = 0; x < len; x++) {
            G hi = hgf_mul(t[x], r[i]);
            t[x + len] = hi;
            t[x] = hgf_add(t[x], hi);
        }
    }
    return t;
}


// ---- eq tables mod q on the device ----
// eq_table_mod_q (src/ligerito_flock.rs:1330-1347), entry by entry: x's
// Montgomery product of the per-coordinate factors (1 - r_i or r_i,
// coordinate 0 on the top index bit), passed in Montgomery form as kernel
// arguments, then one product by 1 to leave the Montgomery domain.
// Montgomery arithmetic with R = 2^128 as ModN::mont (a b R^-1 mod q for
// a b < R q), in 128-bit integer arithmetic.
constexpr int kMaxModQN = 32;
struct ModQArgs {
    uint64_t f[2][kMaxModQN][2];  // [bit][coordinate] = {lo, hi} of the factor x R mod q
};

__device__ __forceinline__ void wide_dev(u128 a, u128 b, u128* hi, u128* lo) {
    const uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64), b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
    const u128 p00 = (u128)a0 * b0, p01 = (u128)a0 * b1, p10 = (u128)a1 * b0, p11 = (u128)a1 * b1;
    const u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
    *lo = (uint64_t)p00 | (mid << 64);
    *hi = p11 + (p01 >> 64) + (p10 >> 64) + (mid >> 64);
}

__device__ __forceinline__ u128 mont_dev(u128 a, u128 b, u128 q, u128 qinv) {
    u128 hi, lo, mh, ml;
    wide_dev(a, b, &hi, &lo);
    const u128 m = lo * qinv;
    wide_dev(m, q, &mh, &ml);
    const u128 c = (lo + ml) < lo ? 1 : 0;
    const u128 r = hi + mh + c;
    return r >= q ? r - q : r;
}

__global__ void k_eq_mod_q(ModQArgs a, int n, uint64_t q_lo, uint64_t q_hi, uint64_t qi_lo, uint64_t qi_hi,
                           uint64_t one_lo, uint64_t one_hi, uint64_t* __restrict__ out) {
    const u128 q = ((u128)q_hi << 64) | q_lo, qinv = ((u128)qi_hi << 64) | qi_lo;
    const uint64_t count = 1ull << n;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < count;
         x += (uint64_t)gridDim.x * blockDim.x) {
        u128 acc = ((u128)one_hi << 64) | one_lo;
        for (int i = 0; i < n; i++) {
            const int bit = (int)((x >> (n - 1 - i)) & 1);
            const u128 f = ((u128)a.f[bit][i][1] << 64) | a.f[bit][i][0];
            acc = mont_dev(acc, f, q, qinv);
        }
        acc = mont_dev(acc, 1, q, qinv);
        out[2 * x] = (uint64_t)acc;
        out[2 * x + 1] = (uint64_t)(acc >> 64);
    }
}

// eq_table_mod_q(r, q) into out (2^n entries of 16 bytes, u128 little-endian)
int eq_table_mod_q_dev(bitz_ctx* ctx, const std::vector<u128>& r, u128 q, void* out) {
    const int n = (int)r.size();
    BITZ_REQUIRE(n <= kMaxModQN, BITZ_ERR_SHAPE, "eq_table_mod_q_dev: more than 32 coordinates");
    ModN m;
    m.set(q);
    ModQArgs a;
    for (int i = 0; i < n; i++) {
        const u128 f1 = m.to_mont(r[i]), f0 = m.to_mont((1 + q - r[i] % q) % q);
        a.f[0][i][0] = (uint64_t)f0;
        a.f[0][i][1] = (uint64_t)(f0 >> 64);
        a.f[1][i][0] = (uint64_t)f1;
        a.f[1][i][1] = (uint64_t)(f1 >> 64);
    }
    const u128 one = m.to_mont(1);
    const uint64_t count = 1ull << n;
    const unsigned blocks = (unsigned)std::min<uint64_t>(4096, (count + 255) / 256);
    TimedLaunch tl(ctx, "eq_mod_q", 16.0 * count, 0.0);
    hipLaunchKernelGGL(k_eq_mod_q, dim3(blocks), dim3(256), 0, ctx->stream, a, n, (uint64_t)q, (uint64_t)(q >> 64),
                       (uint64_t)m.ninv, (uint64_t)(m.ninv >> 64), (uint64_t)one, (uint64_t)(one >> 64),
                       (uint64_t*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

// smallest_generator (src/pcs.rs:85-102): smallest n >= 2 generating GF(2^128)^x
G gf_pow_u128(G b, u128 e) {
    G acc = gone();
    while (e) {
        if (e & 1) acc = hgf_mul(acc, b);
        b = hgf_mul(b, b);
        e >>= 1;
    }
    return acc;
}
G smallest_generator() {
    static const u128 kFactors[9] = {3, 5, 17, 257, 641, 65537, 274177, 6700417, (u128)67280421310721ull};
    u128 order = ~(u128)0;
    for (uint64_t n = 2;; n++) {
        G a{n, 0};
        bool ok = true;
        for (u128 p : kFactors)
            if (geq(gf_pow_u128(a, order / p), gone())) {
                ok = false;
                break;
            }
        if (ok) return a;
    }
}

// The statement len || a || len || b || target (LinearClaim encoding,
// src/bitz/params.rs:273-285) absorbed as one public message, with a and b
// (na, nb field elements) in device memory: its whole chunks are hashed on
// the device as aligned subtree runs, the rest on the host. T is the prover's
// or the verifier's inner transcript.
template <class T>
int public_weights_dev(bitz_ctx* ctx, T& it, const void* a, uint64_t na, const void* b, uint64_t nb, G target) {
    const size_t len = 16 * (na + nb) + 32;
    int rc = ctx->scratch.ensure(64);  // the pinned header staging
    if (rc) return rc;
    const Transcript::SlicePlan plan = it.public_plan(len);
    const size_t pad = (4 - plan.lead % 4) % 4;  // chunk data 4-byte aligned
    DevBuf msg, cvs, roots;
    const std::vector<int> ks = blake3_subtree_runs(plan.c0, plan.chunks);
    if (!msg.alloc(pad + len + 64) || !cvs.alloc(32 * std::max<uint64_t>(plan.chunks, 1)) ||
        !roots.alloc(32 * std::max<size_t>(ks.size(), 1)))
        return BITZ_ERR_ALLOC;
    uint8_t* m = (uint8_t*)msg.ptr + pad;
    uint8_t* hdr = (uint8_t*)ctx->scratch.host;  // pinned: the copies below stay asynchronous
    uint64_t l = na;
    memcpy(hdr, &l, 8);
    l = nb;
    memcpy(hdr + 8, &l, 8);
    gput(target, hdr + 16);
    BITZ_HIP(hipMemcpyAsync(m, hdr, 8, hipMemcpyHostToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(m + 8, a, 16 * na, hipMemcpyDeviceToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(m + 8 + 16 * na, hdr + 8, 8, hipMemcpyHostToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(m + 16 + 16 * na, b, 16 * nb, hipMemcpyDeviceToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(m + 16 + 16 * (na + nb), hdr + 16, 16, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = blake3_runs(ctx, m + plan.lead, plan.chunks, plan.c0, ks, cvs.ptr, roots.ptr))) return rc;
    const size_t tail = len - plan.lead - 1024 * plan.chunks;
    if ((rc = ctx->scratch.ensure_stage(plan.lead + tail + 32 * ks.size()))) return rc;
    uint8_t* st = (uint8_t*)ctx->scratch.stage;
    BITZ_HIP(hipMemcpyAsync(st, m, plan.lead, hipMemcpyDeviceToHost, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(st + plan.lead, m + plan.lead + 1024 * plan.chunks, tail, hipMemcpyDeviceToHost,
                            ctx->stream));
    if (!ks.empty())
        BITZ_HIP(hipMemcpyAsync(st + plan.lead + tail, roots.ptr, 32 * ks.size(), hipMemcpyDeviceToHost,
                                ctx->stream));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<uint32_t> run_cvs(8 * ks.size());
    if (!ks.empty()) memcpy(run_cvs.data(), st + plan.lead + tail, 32 * ks.size());
    it.public_message_runs(len, plan, st, run_cvs.data(), st + plan.lead);
    return BITZ_OK;
}

// s_k(v_k) for k = 0..log_n and their inverses (eval_sk_at_vks,
// ligerito.rs:1720-1745): s_{i+1}(x) = s_i(x)^2 + s_i(v_i) s_i(x), v_i = X^i
void sk_at_vks(int log_n, std::vector<G>* sks, std::vector<G>* inv) {
    sks->assign(log_n + 1, gzero());
    (*sks)[0] = gone();
    std::vector<G> layer(log_n);
    for (int i = 1; i <= log_n; i++) layer[i - 1] = G{1ull << i, 0};
    int c = log_n;
    for (int i = 0; i < log_n; i++) {
        for (int j = 0; j < c; j++) {
            G sv = layer[j];
            G v = hgf_add(hgf_mul(sv, sv), hgf_mul((*sks)[i], sv));
            if (j == 0)
                (*sks)[i + 1] = v;
            else
                layer[j - 1] = v;
        }
        c--;
    }
    batch_inverse(*sks, inv);
}

// -------------------------------------------------------------- proof writer
struct Writer {
    std::vector<uint8_t> b;
    void u64(uint64_t v) { b.insert(b.end(), (uint8_t*)&v, (uint8_t*)&v + 8); }
    void bytes(const void* p, size_t n) { b.insert(b.end(), (const uint8_t*)p, (const uint8_t*)p + n); }
    void gf(G g) {
        u64(g.lo);
        u64(g.hi);
    }
};

struct LevelOpen {
    std::vector<G> rows;  // nq * lanes
    uint32_t lanes = 0;
    std::vector<uint8_t> proof;  // k * 32
};

void write_open(Writer& w, const LevelOpen& o) {
    size_t nq = o.lanes ? o.rows.size() / o.lanes : 0;
    w.u64(nq);
    for (size_t i = 0; i < nq; i++) {
        w.u64(o.lanes);
        for (uint32_t l = 0; l < o.lanes; l++) w.gf(o.rows[i * o.lanes + l]);
    }
    w.u64(o.proof.size() / 32);
    w.bytes(o.proof.data(), o.proof.size());
}

// ------------------------------------------------------------- the prover
struct Prover {
    bitz_ctx* ctx;
    const bitz_commit_handle* h;
    int t, s, m, mp;
    LigConfig cfg;
    G g;
    // device buffers
    DevBuf packed, rowimg, colimg, folds, roww, levels, eqc, eqy, eqcinv, tmp0, tmp1, tmp2, tmp3, qdev, pt, fsbuf;
    // the committed bits as 16-bit row words (gkr_nibble_rows), read by the
    // low GKR layers' table rounds
    DevBuf nibbles;
    // Split column weights of a weighted layer: eq_c[c] = eqlo[c & (2^wlo - 1)]
    // * eqhi[c >> wlo] (wlo = min(s, 8)). The right half carries eqhi only:
    // its multiplies are window-table products over 256-column runs, and
    // the row-only round sums weigh each thread's slots by one eqlo entry.
    DevBuf eqlo, eqhi;
    bool wsplit = false;
    int wlo = 0;
    std::vector<DevBuf*> lig_bufs;

    // through the pinned upload ring: no stream synchronisation, and the
    // host vector may be freed as soon as this returns
    int upload_gf(const std::vector<G>& v, DevBuf& buf) {
        if (!buf.alloc(std::max<size_t>(v.size(), 1) * 16)) return BITZ_ERR_ALLOC;
        return upload_async(ctx, v.data(), v.size() * 16, buf.ptr);
    }
    int download(const void* dev, void* host, size_t bytes) {
        BITZ_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
        BITZ_HIP(hipStreamSynchronize(ctx->stream));
        return BITZ_OK;
    }
    int result(int K, G* out) {
        uint8_t buf[64];
        int rc;
        {
            SubPhase sp(ctx, "sub_round_wait");
            rc = read_result(ctx, K, buf);
        }
        if (rc) return rc;
        for (int k = 0; k < K; k++) out[k] = gfrom(buf + 16 * k);
        return BITZ_OK;
    }
    // eq table over host coordinates into `buf`
    // eqhi for a pattern fold of a weighted layer (none when s <= 8: the
    // high bits' factor is then 1)
    const void* hi_weights(bool weigh) const { return weigh && wsplit && s > wlo ? eqhi.ptr : nullptr; }

    // Round 0 of a weighted layer: the right half (2h entries) takes the
    // column weights (eqhi only when split), then the row-only sums.
    int premul_round(const uint4* l, uint4* r, uint64_t hh, int send_one) {
        if (!wsplit) return gkr_premul_round_sums(ctx, l, r, hh, eqc.ptr, s, eqy.ptr, send_one, nullptr);
        if (s > wlo)
            if (int rc = scale_cols_hi(ctx, r, 2 * hh, s, eqhi.ptr)) return rc;
        return gkr_round_sums_rowonly(ctx, l, r, hh, s, eqy.ptr, send_one, nullptr, eqlo.ptr, wlo);
    }

    // Product tables queued for one launch: those of at most kTensorJobN
    // coordinates wait for flush() (tensor_args_multi, up to kTensorJobs per
    // launch); larger ones are launched at once (eq_dev / tensor_args).
    struct TableJobs {
        std::vector<std::vector<G>> a, b;
        std::vector<DevBuf*> bufs;
        int add(Prover* p, std::vector<G> av, std::vector<G> bv, DevBuf& buf) {
            const int n = (int)av.size();
            if (!buf.alloc(16ull << n)) return BITZ_ERR_ALLOC;
            if (n > kTensorJobN) return tensor_args(p->ctx, av.data(), bv.data(), n, buf.ptr);
            a.push_back(std::move(av));
            b.push_back(std::move(bv));
            bufs.push_back(&buf);
            return BITZ_OK;
        }
        int eq(Prover* p, const std::vector<G>& r, DevBuf& buf) {
            if ((int)r.size() > kTensorJobN) return p->eq_dev(r, buf);
            std::vector<G> av(r.size());
            for (size_t i = 0; i < r.size(); i++) av[i] = hgf_add(r[i], gone());
            return add(p, std::move(av), r, buf);
        }
        int flush(Prover* p) {
            for (size_t j0 = 0; j0 < bufs.size(); j0 += kTensorJobs) {
                const int nj = (int)std::min<size_t>(kTensorJobs, bufs.size() - j0);
                const hgf* pa[kTensorJobs];
                const hgf* pb[kTensorJobs];
                int n[kTensorJobs];
                void* out[kTensorJobs];
                for (int j = 0; j < nj; j++) {
                    pa[j] = a[j0 + j].data();
                    pb[j] = b[j0 + j].data();
                    n[j] = (int)a[j0 + j].size();
                    out[j] = bufs[j0 + j]->ptr;
                }
                if (int rc = tensor_args_multi(p->ctx, nj, pa, pb, n, out)) return rc;
            }
            a.clear();
            b.clear();
            bufs.clear();
            return BITZ_OK;
        }
    };

    int eq_dev(const std::vector<G>& r, DevBuf& buf) {
        if ((int)r.size() <= kMaxTensorN) {
            if (!buf.alloc((16ull << r.size()))) return BITZ_ERR_ALLOC;
            return eq_args(ctx, r.data(), (int)r.size(), buf.ptr);
        }
        if (!pt.alloc(std::max<size_t>(r.size(), 1) * 16)) return BITZ_ERR_ALLOC;
        if (int rc = upload_async(ctx, r.data(), r.size() * 16, pt.ptr)) return rc;
        if (!buf.alloc((16ull << r.size()))) return BITZ_ERR_ALLOC;
        return eq_table(ctx, pt.ptr, (int)r.size(), buf.ptr);
    }

    // Public message len(a) u64 || a || len(b) u64 || b || target (a, b device
    // arrays of field elements): bytes assembled in device memory, whole
    // chunks hashed there (blake3_runs), lead and tail bytes on the host.
    int public_weights(InnerTranscript& it, const void* a, uint64_t na, const void* b, uint64_t nb, G target) {
        return public_weights_dev(ctx, it, a, na, b, nb, target);
    }

    // Round 0 value y = MLE[p_msg](point) (ood_eval, src/ligerito_flock.rs:1459-1486)
    int mle_packed(const std::vector<G>& point, G* y) {
        int rc = eq_dev(point, tmp0);
        if (rc) return rc;
        rc = dot(ctx, h->packed.ptr, tmp0.ptr, 1ull << mp, nullptr);
        if (rc) return rc;
        return result(1, y);
    }

    // ---- batched grand-product GKR (Forest::prove == gpgkr_prove) ----
    // Layer kinds: halves materialised, or an implicit low level whose first
    // rounds run off the bits and the row words, and whose table round's fold
    // writes halves of 2^(t+s-5) entries: the leaves (rounds 0-1 from bit
    // patterns, round 2 from row words, round 3 from the row-word fold
    // tables), level 1 (round 0 from bit patterns, round 1 from row words,
    // round 2 from tables), level 2 (round 0 from row words, round 1 from
    // tables), level 3 (round 0 from its leaf product tables).
    enum LayerKind { kDense = 0, kLevel1 = 1, kLeaf = 2, kLevel2 = 3, kLevel3 = 4 };
    // the level a low layer serves, its row-word round (none for level 3) and its table round
    static int nibble_level(LayerKind kind) {
        return kind == kLeaf ? 0 : kind == kLevel1 ? 1 : kind == kLevel2 ? 2 : 3;
    }
    static int nibble_round(LayerKind kind) { return 2 - nibble_level(kind); }
    static int table_round(LayerKind kind) { return 3 - nibble_level(kind); }
    int gkr_layer(InnerTranscript& it, uint4* l, uint4* r, LayerKind kind, std::vector<G>& internal, G* claim) {
        std::optional<SubPhase> setup_phase(std::in_place, ctx, "sub_gkr_setup");
        const bool leaf = kind == kLeaf;
        const int L = (int)internal.size();
        std::vector<G> ext(internal.rbegin(), internal.rend());
        std::vector<G> next;
        G factor = gone();
        int rc;
        // Round `round` weighs its slots by eq over ext[0..rem) (rem = L-1-round)
        // as eq_y[x >> s_c] * eq_c[x & (2^s_c - 1)]. The tables are built once,
        // over ext[0..L-1); read at indices below 2^rem they equal the round's
        // weights times the constant prod_{i=1..round} (1 + internal[i]) (the
        // dropped coordinates sit at x = 0), which corr[round] divides out.
        // Should a coordinate equal one the constant vanishes, and that layer
        // rebuilds its tables every round instead.
        std::vector<G> corr(L, gone());
        bool per_layer = true;
        {
            std::vector<G> pp(L, gone());
            for (int i = 1; i < L; i++) pp[i] = hgf_mul(pp[i - 1], hgf_add(internal[i], gone()));
            if (geq(pp[L - 1], gzero())) {
                per_layer = false;
            } else {
                G cur = hgf_inv(pp[L - 1]);
                for (int i = L - 1; i >= 1; i--) {
                    corr[i] = cur;
                    cur = hgf_mul(cur, hgf_add(internal[i], gone()));
                }
            }
        }
        // the layer's tables of at most kTensorJobN coordinates go out in
        // one launch (tensor_args_multi) after the set-up below
        TableJobs tabs;
        if (per_layer) {
            int s_c0 = std::min(s, L - 1);
            std::vector<G> c(ext.begin(), ext.begin() + s_c0);
            std::vector<G> yv(ext.begin() + s_c0, ext.begin() + (L - 1));
            if ((rc = tabs.eq(this, c, eqc))) return rc;
            if ((rc = tabs.eq(this, yv, eqy))) return rc;
        }
        // Column weights carried by the right half (weighable / batch_inverse,
        // src/bitz/gkr.rs:48-80; kernels.rs:114-131): eq_c depends only on the
        // low s index bits, so r multiplied by it once (round 0, or the leaf
        // layer's second fold) stays so through every round that binds a row
        // variable, whose slots then skip the two eq_c multiplies. Before the
        // first column round the factor is divided out again with the table
        // prod_i (bit i ? ext_i : 1 + ext_i)^-1 (no coordinate is 0 or 1).
        bool weigh = per_layer && s > 0 && L - 1 >= s;
        for (int i = 0; weigh && i < s; i++)
            if (geq(ext[i], gzero()) || geq(ext[i], gone())) weigh = false;
        wsplit = weigh && s <= 28;
        wlo = std::min(s, 8);
        if (wsplit) {
            std::vector<G> clo(ext.begin(), ext.begin() + wlo), chi(ext.begin() + wlo, ext.begin() + s);
            if ((rc = tabs.eq(this, clo, eqlo))) return rc;
            if (s > wlo && (rc = tabs.eq(this, chi, eqhi))) return rc;
        }
        if (weigh) {
            std::vector<G> v(2 * s), inv(2 * s);
            for (int i = 0; i < s; i++) {
                v[i] = hgf_add(ext[i], gone());
                v[s + i] = ext[i];
            }
            G acc = gone();
            for (int i = 0; i < 2 * s; i++) {
                inv[i] = acc;  // prefix product
                acc = hgf_mul(acc, v[i]);
            }
            G cur = hgf_inv(acc);
            for (int i = 2 * s - 1; i >= 0; i--) {
                inv[i] = hgf_mul(inv[i], cur);
                cur = hgf_mul(cur, v[i]);
            }
            // split weights: r never carries the low bits' factors
            for (int i = 0; wsplit && i < wlo; i++) inv[i] = inv[s + i] = gone();
            if ((rc = tabs.add(this, std::vector<G>(inv.begin(), inv.begin() + s),
                               std::vector<G>(inv.begin() + s, inv.end()), eqcinv)))
                return rc;
        }
        if ((rc = tabs.flush(this))) return rc;
        if (per_layer && ctx->device_fs)
            return gkr_layer_device(it, l, r, kind, internal, corr, weigh, claim);
        bool weighted = false;  // r currently carries eq_c
        // a round's fold by its challenge is enqueued with the next round's
        // sums: fused into one pass where the right half carries eq_c (two
        // products per slot), a fold pass then a sums pass otherwise (with the
        // eq_c multiplies as well, the fused kernel loses to the two passes)
        bool pending_fold = false;
        uint8_t pending_rho[16], rho0[16];
        uint8_t last[32];
        bool tail_done = false;
        BITZ_REQUIRE(kind == kDense || (L >= 4 && t >= 7 && (!weigh || wsplit)), BITZ_ERR_SHAPE,
                     "GKR: layer too small for bit patterns");
        setup_phase.reset();
        std::vector<hgf> chal;  // the layer's challenges so far (for its row-word and table rounds)
        const int nround = kind == kDense ? -1 : nibble_round(kind);
        const int jround = kind == kDense ? -1 : table_round(kind);
        for (int round = 0; round < L; round++) {
            G z = internal[round];
            int rem = L - 1 - round;
            uint64_t hh = 1ull << rem;
            int s_c = std::min(s, rem);
            const bool bit_round = kind != kDense && round <= jround;
            if (!bit_round && rem <= ctx->tune.host_tail_log && (!weighted || rem < s)) {
                // the remaining rounds on the host, from dense halves
                if (pending_fold && (rc = fold_pair(ctx, l, r, 2 * hh, pending_rho))) return rc;
                if (weighted && (rc = mul_inplace(ctx, r, eqcinv.ptr, 1ull << s))) return rc;
                if ((rc = host_tail(it, l, r, round, ext, internal, &factor, &next, last))) return rc;
                pending_fold = false;
                tail_done = true;
                break;
            }
            if (!per_layer) {
                std::vector<G> c(ext.begin(), ext.begin() + s_c);
                std::vector<G> yv(ext.begin() + s_c, ext.begin() + rem);
                if ((rc = eq_dev(c, eqc))) return rc;
                if ((rc = eq_dev(yv, eqy))) return rc;
            }
            int send_one = geq(z, gzero()) ? 1 : 0;
            if (round == nround) {
                rc = gkr_nibble_round_sums(ctx, nibble_level(kind), nibbles.ptr, rowimg.ptr, t, s, chal.data(),
                                           nullptr, eqc.ptr, eqy.ptr, send_one, nullptr);
            } else if (round == jround) {
                rc = gkr_jit_tables(ctx, nibble_level(kind), rowimg.ptr, t, chal.data(), nullptr);
                if (!rc)
                    rc = gkr_jit_round_sums(ctx, nibble_level(kind), nibbles.ptr, t, s, eqc.ptr, eqy.ptr, send_one,
                                            nullptr);
            } else if (leaf && round == 0)
                rc = gkr_leaf_round_sums(ctx, (const uint64_t*)packed.ptr, rowimg.ptr, t, s, eqc.ptr, s_c, eqy.ptr,
                                         send_one, nullptr, nibbles.ptr);
            else if (leaf && round == 1)
                rc = gkr_leaf_round1_sums(ctx, (const uint64_t*)packed.ptr, rowimg.ptr, t, s, rho0, eqc.ptr, eqy.ptr,
                                          send_one, nullptr, nullptr, nibbles.ptr);
            else if (kind == kLevel1 && round == 0)
                rc = gkr_level1_round0_sums(ctx, (const uint64_t*)packed.ptr, rowimg.ptr, t, s, eqc.ptr, eqy.ptr,
                                            send_one, nullptr, nibbles.ptr);
            else if (pending_fold && weighted && rem >= s) {
                rc = gkr_fold_round_sums(ctx, l, r, hh, pending_rho, nullptr, s, eqy.ptr, send_one, nullptr, nullptr,
                                         wsplit ? eqlo.ptr : nullptr, wlo);
            } else {
                if (pending_fold && (rc = fold_pair(ctx, l, r, 2 * hh, pending_rho))) return rc;
                if (weighted && rem < s) {
                    // all row variables bound: r holds 2^s entries eq_c[c] * r[c]
                    if ((rc = mul_inplace(ctx, r, eqcinv.ptr, 1ull << s))) return rc;
                    weighted = false;
                }
                if (weighted) {
                    rc = gkr_round_sums_rowonly(ctx, l, r, hh, s, eqy.ptr, send_one, nullptr,
                                                wsplit ? eqlo.ptr : nullptr, wlo);
                } else if (weigh && round == 0) {
                    rc = premul_round(l, r, hh, send_one);
                    weighted = true;
                } else {
                    rc = gkr_round_sums(ctx, l, r, hh, eqc.ptr, s_c, eqy.ptr, send_one, nullptr);
                }
            }
            if (rc) return rc;
            G sums[2];
            if ((rc = result(2, sums))) return rc;
            uint8_t msg[32];
            G f = per_layer ? hgf_mul(factor, corr[round]) : factor;
            gput(hgf_mul(f, sums[0]), msg);
            gput(hgf_mul(f, sums[1]), msg + 16);
            it.prover_message(msg, 32);
            uint8_t rb[16];
            it.verifier_gf(rb);
            G rr = gfrom(rb);
            next.push_back(rr);
            if (bit_round) chal.push_back(hgf{rr.lo, rr.hi});
            if (round == jround) {
                // the level stays implicit up to here: the table fold writes the halves
                if ((rc = gkr_jit_fold(ctx, nibble_level(kind), nibbles.ptr, t, s, chal.data(), nullptr,
                                       hi_weights(weigh), l, r)))
                    return rc;
                weighted = weigh;
                pending_fold = false;
            } else if (leaf && round == 0) {
                // leaves stay implicit: round 1 reads the bits again
                memcpy(rho0, rb, 16);
                pending_fold = false;
            } else if (bit_round) {
                pending_fold = false;
            } else {
                pending_fold = true;
                memcpy(pending_rho, rb, 16);
            }
            factor = hgf_mul(factor, eq_factor(rr, z));
        }
        if (pending_fold && (rc = fold_pair(ctx, l, r, 1, pending_rho))) return rc;
        if (!tail_done) {
            if ((rc = download(l, last, 16))) return rc;
            if ((rc = download(r, last + 16, 16))) return rc;
        }
        it.prover_message(last, 32);
        uint8_t rb[16];
        it.verifier_gf(rb);
        G rc_ = gfrom(rb);
        G el = gfrom(last), er = gfrom(last + 16);
        *claim = hgf_add(el, hgf_mul(rc_, hgf_add(er, el)));
        internal.clear();
        internal.push_back(rc_);
        internal.insert(internal.end(), next.begin(), next.end());
        return BITZ_OK;
    }

    // Rounds round0..L-1 of a GKR layer on the host once each half holds at
    // most 2^(host_tail_log+1) entries: the halves come down once, then every
    // round is the same sums / message / challenge / fold sequence as the
    // device rounds (gkr_layer), with the round's own eq table over
    // ext[0..rem) as slot weights (so no corr[] constant; Sums::slot,
    // src/bitz/kernels.rs:299-310). `last` receives the folded l[0], r[0].
    int host_tail(InnerTranscript& it, const uint4* l, const uint4* r, int round0, const std::vector<G>& ext,
                  const std::vector<G>& internal, G* factor, std::vector<G>* next, uint8_t last[32]) {
        SubPhase sp(ctx, "sub_gkr_tail");
        using hostgf::E;
        const int L = (int)internal.size();
        const int rem0 = L - 1 - round0;
        const uint64_t n = 2ull << rem0;  // entries of each half at round0
        void* both;
        {
            SubPhase sw(ctx, "sub_gkr_tail_wait");
            if (int rc = download_pair_mapped(ctx, l, r, n, &both)) return rc;
        }
        E* hl = (E*)both;
        E* hr = hl + n;
        std::vector<E> coords(rem0 > 0 ? rem0 : 1), w(1ull << rem0);
        for (int i = 0; i < rem0; i++) coords[i] = E{ext[i].lo, ext[i].hi};
        hostgf::eq_table(coords.data(), rem0, w.data());
        for (int round = round0; round < L; round++) {
            const int rem = L - 1 - round;
            const uint64_t hh = 1ull << rem;
            const G z = internal[round];
            if (round > round0) hostgf::drop_top(w.data(), hh);  // eq over ext[0..rem)
            E sums[2];
            hostgf::gkr_round_sums(hl, hr, hh, w.data(), geq(z, gzero()) ? 1 : 0, sums);
            uint8_t msg[32];
            gput(hgf_mul(*factor, G{sums[0].lo, sums[0].hi}), msg);
            gput(hgf_mul(*factor, G{sums[1].lo, sums[1].hi}), msg + 16);
            it.prover_message(msg, 32);
            uint8_t rb[16];
            it.verifier_gf(rb);
            const G rr = gfrom(rb);
            next->push_back(rr);
            hostgf::fold(hl, hh, E{rr.lo, rr.hi});
            hostgf::fold(hr, hh, E{rr.lo, rr.hi});
            *factor = hgf_mul(*factor, eq_factor(rr, z));
        }
        memcpy(last, hl, 16);
        memcpy(last + 16, hr, 16);
        return BITZ_OK;
    }

    // returns the leaf point (LE, t+s coordinates) and claim
    int gkr(InnerTranscript& it, const std::vector<G>& zeta, std::vector<G>* point, G* claim) {
        // Storage: levels 4..t-1 back to back (level j: 2^(t+s-j) entries).
        // Levels 0-3 are never materialised; level 4 comes from level 3's
        // tables of the row words, the upper levels from products. Once layer
        // 4 is proved, level 4's 2^(t+s-4) entries take the halves the level
        // 3..0 layers fold into (2 x 2^(t+s-5) entries, shared: the layers run
        // one after another).
        const uint64_t h4 = 1ull << (t + s - 4), h5 = h4 >> 1;
        uint64_t total = 0;
        for (int j = 4; j < t; j++) total += 1ull << (t + s - j);
        if (!levels.alloc(total * 16)) return BITZ_ERR_ALLOC;
        std::vector<uint4*> lv(t + 1, nullptr);
        uint64_t off = 0;
        for (int j = 4; j < t; j++) {
            lv[j] = (uint4*)levels.ptr + off;
            off += 1ull << (t + s - j);
        }
        uint4* low = lv[4];
        if (!nibbles.alloc(h4 * 2)) return BITZ_ERR_ALLOC;  // 2^(t+s-4) u16
        int rc = gkr_nibble_rows(ctx, (const uint64_t*)packed.ptr, t, s, nibbles.ptr);
        if (rc) return rc;
        if ((rc = gkr_level4(ctx, nibbles.ptr, rowimg.ptr, t, s, lv[4]))) return rc;
        for (int j = 4; j + 1 < t; j++)
            if ((rc = product_level(ctx, lv[j], 1ull << (t + s - j), lv[j + 1]))) return rc;
        std::vector<G> internal(zeta.rbegin(), zeta.rend());
        G cl = gzero();
        for (int j = t - 1; j >= 4; j--) {
            uint64_t half = 1ull << (t + s - j - 1);
            if ((rc = gkr_layer(it, lv[j], lv[j] + half, kDense, internal, &cl))) return rc;
        }
        for (LayerKind kind : {kLevel3, kLevel2, kLevel1, kLeaf})
            if ((rc = gkr_layer(it, low, low + h5, kind, internal, &cl))) return rc;
        point->assign(internal.rbegin(), internal.rend());
        *claim = cl;
        return BITZ_OK;
    }

    // The rounds of one layer with the transcript on the device: every round's
    // launches (fold, sums, Fiat-Shamir step) are enqueued at once, the
    // challenges come from the device copy of the transcript, and a single
    // synchronisation at the end brings the messages and challenges back.
    // The host then replays them through its own transcript, which also
    // checks every device challenge. Kernel choices are those of gkr_layer.
    int gkr_layer_device(InnerTranscript& it, uint4* l, uint4* r, LayerKind kind, std::vector<G>& internal,
                         const std::vector<G>& corr, bool weigh, G* claim) {
        const bool leaf = kind == kLeaf;
        const int L = (int)internal.size();
        int rc;
        const size_t st_units = (sizeof(B3State) + 15) / 16;
        const size_t up_units = st_units + 1 + 2 * (size_t)L, down_units = 3 * (size_t)L + 2;
        if (!fsbuf.alloc((up_units + down_units) * 16)) return BITZ_ERR_ALLOC;
        {
            std::vector<uint8_t> up(up_units * 16, 0);
            B3State st;
            it.export_state(&st);
            memcpy(up.data(), &st, sizeof(st));
            G one = gone();
            memcpy(up.data() + st_units * 16, &one, 16);
            memcpy(up.data() + (st_units + 1) * 16, internal.data(), 16 * (size_t)L);
            memcpy(up.data() + (st_units + 1 + L) * 16, corr.data(), 16 * (size_t)L);
            BITZ_HIP(hipMemcpyAsync(fsbuf.ptr, up.data(), up.size(), hipMemcpyHostToDevice, ctx->stream));
            BITZ_HIP(hipStreamSynchronize(ctx->stream));  // `up` is freed on return
        }
        uint4* base = (uint4*)fsbuf.ptr;
        void* st_dev = base;
        uint4* factor = base + st_units;
        uint4* internal_dev = factor + 1;
        uint4* corr_dev = internal_dev + L;
        uint4* rho = corr_dev + L;
        uint4* msgs = rho + L;
        uint4* last = msgs + 2 * L;
        struct DeferGuard {
            bitz_ctx* c;
            ~DeferGuard() { c->defer_sums = false; }
        } guard{ctx};
        ctx->defer_sums = true;
        bool weighted = false, pending_fold = false;
        const uint4* pending_rho = nullptr;
        BITZ_REQUIRE(kind == kDense || (L >= 4 && t >= 7 && (!weigh || wsplit)), BITZ_ERR_SHAPE,
                     "GKR: layer too small for bit patterns");
        const int nround = kind == kDense ? -1 : nibble_round(kind);
        const int jround = kind == kDense ? -1 : table_round(kind);
        for (int round = 0; round < L; round++) {
            int rem = L - 1 - round;
            uint64_t hh = 1ull << rem;
            int s_c = std::min(s, rem);
            int send_one = geq(internal[round], gzero()) ? 1 : 0;
            // the layer's challenges r_0.. are rho[0..]: the row-word and table kernels read them there
            if (round == nround) {
                rc = gkr_nibble_round_sums(ctx, nibble_level(kind), nibbles.ptr, rowimg.ptr, t, s, nullptr, rho,
                                           eqc.ptr, eqy.ptr, send_one, nullptr);
            } else if (round == jround) {
                rc = gkr_jit_tables(ctx, nibble_level(kind), rowimg.ptr, t, nullptr, rho);
                if (!rc)
                    rc = gkr_jit_round_sums(ctx, nibble_level(kind), nibbles.ptr, t, s, eqc.ptr, eqy.ptr, send_one,
                                            nullptr);
            } else if (leaf && round == 0)
                rc = gkr_leaf_round_sums(ctx, (const uint64_t*)packed.ptr, rowimg.ptr, t, s, eqc.ptr, s_c, eqy.ptr,
                                         send_one, nullptr, nibbles.ptr);
            else if (leaf && round == 1)
                rc = gkr_leaf_round1_sums(ctx, (const uint64_t*)packed.ptr, rowimg.ptr, t, s, nullptr, eqc.ptr,
                                          eqy.ptr, send_one, nullptr, rho, nibbles.ptr);
            else if (kind == kLevel1 && round == 0)
                rc = gkr_level1_round0_sums(ctx, (const uint64_t*)packed.ptr, rowimg.ptr, t, s, eqc.ptr, eqy.ptr,
                                            send_one, nullptr, nibbles.ptr);
            else if (pending_fold && weighted && rem >= s) {
                rc = gkr_fold_round_sums(ctx, l, r, hh, nullptr, nullptr, s, eqy.ptr, send_one, nullptr, pending_rho,
                                         wsplit ? eqlo.ptr : nullptr, wlo);
            } else {
                if (pending_fold && (rc = fold_pair(ctx, l, r, 2 * hh, nullptr, pending_rho))) return rc;
                if (weighted && rem < s) {
                    if ((rc = mul_inplace(ctx, r, eqcinv.ptr, 1ull << s))) return rc;
                    weighted = false;
                }
                if (weighted) {
                    rc = gkr_round_sums_rowonly(ctx, l, r, hh, s, eqy.ptr, send_one, nullptr,
                                                wsplit ? eqlo.ptr : nullptr, wlo);
                } else if (weigh && round == 0) {
                    rc = premul_round(l, r, hh, send_one);
                    weighted = true;
                } else {
                    rc = gkr_round_sums(ctx, l, r, hh, eqc.ptr, s_c, eqy.ptr, send_one, nullptr);
                }
            }
            if (rc) return rc;
            if ((rc = gkr_fs_round(ctx, ctx->deferred_nb, st_dev, factor, internal_dev, corr_dev, rho, msgs,
                                   (uint32_t)round)))
                return rc;
            pending_fold = false;
            if (round == jround) {
                if ((rc = gkr_jit_fold(ctx, nibble_level(kind), nibbles.ptr, t, s, nullptr, rho, hi_weights(weigh),
                                       l, r)))
                    return rc;
                weighted = weigh;
            } else if (kind != kDense && round < jround) {
                // the level stays implicit: its next round reads the bits again
            } else {
                pending_fold = true;
                pending_rho = rho + round;
            }
        }
        if (pending_fold && (rc = fold_pair(ctx, l, r, 1, nullptr, pending_rho))) return rc;
        ctx->defer_sums = false;
        BITZ_HIP(hipMemcpyAsync(last, l, 16, hipMemcpyDeviceToDevice, ctx->stream));
        BITZ_HIP(hipMemcpyAsync(last + 1, r, 16, hipMemcpyDeviceToDevice, ctx->stream));
        std::vector<uint8_t> down(down_units * 16);
        if ((rc = download(rho, down.data(), down.size()))) return rc;
        // replay through the host transcript; every challenge must agree
        std::vector<G> next;
        const uint8_t* rho_h = down.data();
        const uint8_t* msg_h = down.data() + 16 * (size_t)L;
        for (int round = 0; round < L; round++) {
            it.prover_message(msg_h + 32 * (size_t)round, 32);
            uint8_t rb[16];
            it.verifier_gf(rb);
            if (memcmp(rb, rho_h + 16 * (size_t)round, 16) != 0) {
                set_error("GKR: device transcript diverged from the host transcript");
                return BITZ_ERR_DEVICE;
            }
            next.push_back(gfrom(rb));
        }
        const uint8_t* last_h = down.data() + 48 * (size_t)L;
        it.prover_message(last_h, 32);
        uint8_t rb[16];
        it.verifier_gf(rb);
        G rc_ = gfrom(rb);
        G el = gfrom(last_h), er = gfrom(last_h + 16);
        *claim = hgf_add(el, hgf_mul(rc_, hgf_add(er, el)));
        internal.clear();
        internal.push_back(rc_);
        internal.insert(internal.end(), next.begin(), next.end());
        return BITZ_OK;
    }

    // ---- inner-product sumcheck (src/bitz/sumcheck.rs:146-198, 372-400) ----
    int sumcheck_rounds(InnerTranscript& it, DevBuf& a, DevBuf& w, DevBuf& a2, DevBuf& w2, uint64_t n, G* target,
                        std::vector<G>& point) {
        int rc;
        void* A = a.ptr;
        void* W = w.ptr;
        void* A2 = a2.ptr;
        void* W2 = w2.ptr;
        hostgf::E* ha = nullptr;  // the tables on the host once they are small
        hostgf::E* hw = nullptr;
        while (n > 1) {
            G v[3];
            if (!ha && n <= (2ull << ctx->tune.host_tail_log)) {
                // the remaining rounds on the host (as host_tail for GKR)
                void* both;
                if ((rc = download_pair_mapped(ctx, A, W, n, &both))) return rc;
                ha = (hostgf::E*)both;
                hw = ha + n;
            }
            if (ha) {
                hostgf::E e[3];
                hostgf::inner_round(ha, hw, n / 2, e);
                for (int k = 0; k < 3; k++) v[k] = G{e[k].lo, e[k].hi};
            } else {
                if ((rc = round_poly(ctx, A, W, n, nullptr))) return rc;
                if ((rc = result(3, v))) return rc;
            }
            G c0 = v[0], c2 = v[2], c1 = hgf_add(hgf_add(v[0], v[1]), v[2]);
            if (!geq(hgf_add(c1, c2), *target)) {
                set_error("inner sumcheck: claim does not match the committed bits");
                return BITZ_ERR_ARG;
            }
            uint8_t msg[48];
            gput(c0, msg);
            gput(c1, msg + 16);
            gput(c2, msg + 32);
            it.prover_message(msg, 48);
            uint8_t rb[16];
            it.verifier_gf(rb);
            G r = gfrom(rb);
            *target = hgf_add(c0, hgf_mul(r, hgf_add(c1, hgf_mul(r, c2))));
            point.push_back(r);
            if (ha) {
                hostgf::fold_lsb(ha, n / 2, hostgf::E{r.lo, r.hi});
                hostgf::fold_lsb(hw, n / 2, hostgf::E{r.lo, r.hi});
            } else {
                if ((rc = fold_lsb(ctx, A, n, rb, A2))) return rc;
                if ((rc = fold_lsb(ctx, W, n, rb, W2))) return rc;
                std::swap(A, A2);
                std::swap(W, W2);
            }
            n >>= 1;
        }
        uint8_t fin[32];
        if (ha) {
            memcpy(fin, ha, 16);
            memcpy(fin + 16, hw, 16);
        } else {
            if ((rc = download(A, fin, 16))) return rc;
            if ((rc = download(W, fin + 16, 16))) return rc;
        }
        final_a = gfrom(fin);
        final_w = gfrom(fin + 16);
        return BITZ_OK;
    }
    G final_a{0, 0}, final_w{0, 0};  // the tables' last entries after the rounds

    // The inner-product sumcheck over a single column of committed bits
    // (n = 2^m rows, weights w): rounds 0..2 from the bits themselves
    // (bits_round: XOR accumulators weighed here by the eq table of the
    // earlier challenges, the weights folded on the way), then the table
    // after three folds from byte patterns (bits_materialize) and the
    // remaining rounds as sumcheck_rounds. Same messages as bits_to_gf
    // followed by sumcheck_rounds.
    int bit_sumcheck_rounds(InnerTranscript& it, DevBuf& w, uint64_t n, G* target, std::vector<G>& point) {
        int rc;
        const uint64_t* bits = (const uint64_t*)h->packed.ptr;
        DevBuf wa, wb, a3, a2, w2;
        if (!wa.alloc((n / 2) * 16) || !wb.alloc((n / 4) * 16) || !a3.alloc((n / 8) * 16) ||
            !a2.alloc((n / 16) * 16) || !w2.alloc((n / 16) * 16))
            return BITZ_ERR_ALLOC;
        std::vector<G> rs;
        uint8_t rb[16];
        for (int J = 0; J < 3; J++) {
            const void* win = J == 2 ? wa.ptr : w.ptr;
            void* wout = J == 0 ? nullptr : (J == 1 ? wa.ptr : wb.ptr);
            if ((rc = bits_round(ctx, J, bits, win, n >> J, J ? rb : nullptr, wout))) return rc;
            const int C = 1 << J;
            uint8_t acc[12 * 16];
            if ((rc = read_result(ctx, 3 * C, acc))) return rc;
            const std::vector<G> e = host_eq(rs);
            G v[3] = {gzero(), gzero(), gzero()};
            for (int sidx = 0; sidx < 3; sidx++)
                for (int c = 0; c < C; c++) v[sidx] = hgf_add(v[sidx], hgf_mul(e[c], gfrom(acc + 16 * (sidx * C + c))));
            G c0 = v[0], c2 = v[2], c1 = hgf_add(hgf_add(v[0], v[1]), v[2]);
            if (!geq(hgf_add(c1, c2), *target)) {
                set_error("inner sumcheck: claim does not match the committed bits");
                return BITZ_ERR_ARG;
            }
            uint8_t msg[48];
            gput(c0, msg);
            gput(c1, msg + 16);
            gput(c2, msg + 32);
            it.prover_message(msg, 48);
            it.verifier_gf(rb);
            const G r = gfrom(rb);
            *target = hgf_add(c0, hgf_mul(r, hgf_add(c1, hgf_mul(r, c2))));
            point.push_back(r);
            rs.push_back(r);
        }
        // A_3 from byte patterns over e(r0, r1, r2), W_3 = fold of W_2 by r2
        const std::vector<G> e8 = host_eq(rs);
        uint8_t eb[128];
        for (int c = 0; c < 8; c++) gput(e8[c], eb + 16 * c);
        if ((rc = bits_materialize(ctx, bits, wb.ptr, n / 8, rb, eb, a3.ptr, wa.ptr))) return rc;
        return sumcheck_rounds(it, a3, wa, a2, w2, n / 8, target, point);
    }

    int commit_level(const void* f, int log_msg, int log_lanes, int rate, DevBuf& cw, DevBuf& tree, uint8_t root[32]) {
        int k_code = log_msg - log_lanes + rate;
        if (!cw.alloc((16ull << (log_msg + rate))) || !tree.alloc((2ull << k_code) * 32)) return BITZ_ERR_ALLOC;
        int rc = commit_device(ctx, f, log_msg, log_lanes, rate, cw.ptr, tree.ptr);
        if (rc) return rc;
        return download((const uint8_t*)tree.ptr + ((2ull << k_code) - 2) * 32, root, 32);
    }

    // rows at sorted queries (lanes each) and the octopus multi-proof
    // (merkle.rs:698-745) from a device codeword/tree
    int open_level(const void* cw, const void* tree, uint64_t n_leaves, uint32_t lanes,
                   const std::vector<uint64_t>& queries, LevelOpen* out) {
        SubPhase sp(ctx, "sub_lig_open");
        out->lanes = lanes;
        out->rows.resize(queries.size() * lanes);
        // sibling node indices (host only: they depend on the queries alone)
        std::vector<uint64_t> active(queries), nodes;
        uint64_t level_start = 0, level_len = n_leaves;
        while (level_len > 1) {
            std::vector<uint64_t> nx;
            size_t i = 0;
            while (i < active.size()) {
                uint64_t p = active[i];
                bool sib = i + 1 < active.size() && active[i + 1] == (p ^ 1);
                if (sib)
                    i += 2;
                else {
                    nodes.push_back(level_start + (p ^ 1));
                    i += 1;
                }
                nx.push_back(p >> 1);
            }
            active.swap(nx);
            level_start += level_len;
            level_len >>= 1;
        }
        out->proof.resize(nodes.size() * 32);
        // one upload of both index lists, both gathers, one download round
        const size_t nq = queries.size(), nn = nodes.size();
        const size_t rows_bytes = nq * lanes * 16, nodes_bytes = nn * 32;
        int rc = ctx->scratch.ensure_stage(8 * (nq + nn) + rows_bytes + nodes_bytes);
        if (rc) return rc;
        uint8_t* st = (uint8_t*)ctx->scratch.stage;  // pinned: the copies stay asynchronous
        memcpy(st, queries.data(), 8 * nq);
        memcpy(st + 8 * nq, nodes.data(), 8 * nn);
        if (!qdev.alloc(8 * (nq + nn) + 16) || !tmp3.alloc(std::max<size_t>(rows_bytes + nodes_bytes, 16)))
            return BITZ_ERR_ALLOC;
        BITZ_HIP(hipMemcpyAsync(qdev.ptr, st, 8 * (nq + nn), hipMemcpyHostToDevice, ctx->stream));
        const uint64_t* qd = (const uint64_t*)qdev.ptr;
        uint8_t* gathered = (uint8_t*)tmp3.ptr;
        if ((rc = gather_rows(ctx, cw, qd, (uint32_t)nq, lanes, gathered))) return rc;
        if (nn && (rc = gather_nodes(ctx, tree, qd + nq, (uint32_t)nn, gathered + rows_bytes))) return rc;
        uint8_t* back = st + 8 * (nq + nn);
        BITZ_HIP(hipMemcpyAsync(back, gathered, rows_bytes + nodes_bytes, hipMemcpyDeviceToHost, ctx->stream));
        BITZ_HIP(hipStreamSynchronize(ctx->stream));
        memcpy(out->rows.data(), back, rows_bytes);
        if (nn) memcpy(out->proof.data(), back + rows_bytes, nodes_bytes);
        return BITZ_OK;
    }

    int run(const uint8_t* claimed_in, uint8_t* claimed_out, std::vector<uint8_t>* proof_out);
    int open_claim(InnerTranscript& it, PhaseClock& phase, u128 q, DevBuf& colw, u128 claimed,
                   const std::vector<uint8_t>& fold_bytes, const std::vector<G>& ood_point, G ood_y,
                   bool has_r0_nonce, uint64_t r0_nonce, std::vector<uint8_t>* proof_out);
    // the fold messages and the GKR over the bit grid `bits` (column-major
    // rows, this->t x this->s, row weights in roww): the factored
    // inner-product claim u1 (2^t), u2 (2^s), target
    int reduce_claim(InnerTranscript& it, PhaseClock& phase, const void* bits, const std::vector<uint8_t>& fold_bytes,
                     DevBuf& u1, DevBuf& u2, G* ip_target);
    // Pcs::prove_lin_rows of sum u1[b] u2[c] bit(c, b) = target over the
    // committed bits seen as the grid (lt, ls): (t, s) as committed, or
    // (m, 0), one column of every committed bit
    int prove_lin(InnerTranscript& it, PhaseClock& phase, DevBuf& u1, DevBuf& u2, int lt, int ls, G ip_target,
                  bool bind, const std::vector<G>& ood_point, G ood_y, bool has_r0_nonce, uint64_t r0_nonce,
                  std::vector<uint8_t>* proof_out);
};

// Ligerito challenger over the inner transcript (src/bitz/pcs.rs:1062-1175)
struct Challenger {
    InnerTranscript& it;
    bitz_ctx* ctx;
    bool first_f128 = true;
    void observe_label(const std::string& l) { it.public_message(l); }
    void observe_f128(G v) {
        uint8_t b[16];
        gput(v, b);
        if (first_f128) {
            first_f128 = false;
            it.public_message(b, 16);
            return;
        }
        it.prover_message(b, 16);
    }
    void observe_bytes(const uint8_t* p, size_t n) { it.prover_message_bytes(p, n); }
    G sample() {
        uint8_t b[16];
        it.verifier_gf(b);
        return gfrom(b);
    }
    std::vector<G> sample_vec(size_t n) {
        it.public_message(std::string("pcs/flock/sample-vector/v1"));
        uint64_t nn = n;
        it.public_message(&nn, 8);
        std::vector<G> v(n);
        for (auto& x : v) x = sample();
        return v;
    }
    int grind(uint32_t bits, uint64_t* nonce) {
        SubPhase sp(ctx, "sub_lig_grind");
        it.public_message(std::string("pcs/flock/pow/v1"));
        it.public_message(&bits, 4);
        uint8_t seed[16];
        it.verifier_gf(seed);
        static const uint8_t kPrefix[16] = {'b', 'i', 't', 'z', '-', 'p', 'c', 's', '-', 'p', 'o', 'w', '-', 'v', '1', 0};
        int rc = pow_find(ctx, kPrefix, seed, bits, nonce);
        if (rc) return rc;
        it.prover_message(nonce, 8);
        return BITZ_OK;
    }
    // sample_distinct_queries (ligerito.rs:3191-3222)
    std::vector<uint64_t> queries(uint64_t block_len, int count) {
        SubPhase sp(ctx, "sub_lig_queries");
        std::set<uint64_t> seen;
        std::vector<uint64_t> out;
        while ((int)out.size() < count) {
            G v = sample();
            uint64_t q = v.lo % block_len;
            if (seen.insert(q).second) out.push_back(q);
        }
        std::sort(out.begin(), out.end());
        return out;
    }
};

int ceil_log2(uint64_t n) {
    int k = 0;
    while ((1ull << k) < n) k++;
    return k;
}

int Prover::run(const uint8_t* claimed_in, uint8_t* claimed_out, std::vector<uint8_t>* proof_out) {
    int rc;
    const uint64_t nrows = 1ull << t, ncols = 1ull << s;
    PhaseClock phase(ctx);
    phase.mark(proof_out ? "statement" : "eval_statement");
    // ---------------- outer transcript: statement and Round 0 ----------------
    Transcript outer;
    uint8_t gb[16];
    gput(g, gb);
    bind_standalone_statement(outer, h->root, cfg, t, s, gb);
    phase.mark(proof_out ? "round0" : "eval_round0");
    const uint32_t r0bits = (uint32_t)cfg.round0_grinding_bits;
    ood_round_header(outer, mp, r0bits);
    uint64_t r0_nonce = 0;
    if (r0bits) {
        // grind_and_absorb at the Round-0 boundary (src/ligerito_flock.rs:1523-1537,
        // src/piop/spartan/grinding.rs:198-211): BLAKE3(seed || nonce), the
        // 32-byte seed split over the scan's prefix and seed words
        uint8_t seed[32];
        grinding_seed(outer, "bitz/core/ood-round-grinding/v1", 0, r0bits, seed);
        if ((rc = pow_find(ctx, seed, seed + 16, r0bits, &r0_nonce))) return rc;
        absorb_grinding_nonce(outer, r0_nonce);
    }
    uint8_t zb[16];
    outer.challenge(zb, 16);
    std::vector<G> ood_point(mp);
    {
        G cur = gfrom(zb);
        for (int i = 0; i < mp; i++) {
            ood_point[i] = cur;
            cur = hgf_mul(cur, cur);
        }
    }
    G ood_y;
    if ((rc = mle_packed(ood_point, &ood_y))) return rc;
    uint8_t yb[16];
    gput(ood_y, yb);
    {
        // absorb_ood_value (src/ligerito.rs:71-73): tag 0x50, then y
        uint8_t tagged[17];
        tagged[0] = 0x50;
        memcpy(tagged + 1, yb, 16);
        outer.absorb_slice(tagged, 17);
    }
    phase.mark(proof_out ? "instance" : "eval_instance");
    // ---------------- the standalone instance ----------------
    u128 q;
    {
        std::string err;
        if (!sample_prime_with_bits(outer, standalone_q_bits(t), &q, &err)) {
            set_error(err.c_str());
            return BITZ_ERR_DEVICE;
        }
    }
    std::vector<u128> r1(t), r2(s);
    for (auto& x : r1) x = sample_residue(outer, q);
    for (auto& x : r2) x = sample_residue(outer, q);
    phase.mark(proof_out ? "instance_weights" : "eval_instance_weights");
    // row and column weights (eq tables mod q) built on the device, where the
    // column folds and the LinearClaim statement read them
    DevBuf colw;
    if (!roww.alloc(nrows * 16) || !colw.alloc(ncols * 16) || !folds.alloc(ncols * 16)) return BITZ_ERR_ALLOC;
    if ((rc = eq_table_mod_q_dev(ctx, r1, q, roww.ptr))) return rc;
    if ((rc = eq_table_mod_q_dev(ctx, r2, q, colw.ptr))) return rc;
    phase.mark(proof_out ? "instance_folds" : "eval_instance_folds");
    // folds eta_c = sum_b w_b bit(c, b) (exact u128) on the device, down
    // through pinned staging with the column weights when the honest value is
    // wanted
    const bool want_honest = claimed_out != nullptr || claimed_in == nullptr;
    std::vector<uint8_t> fold_bytes(ncols * 16);
    u128 honest = 0;
    {
        if ((rc = ctx->scratch.ensure_stage(32 * ncols))) return rc;
        uint8_t* st = (uint8_t*)ctx->scratch.stage;
        if ((rc = fold_columns(ctx, (const uint64_t*)h->packed.ptr, t, s, roww.ptr, folds.ptr))) return rc;
        BITZ_HIP(hipMemcpyAsync(st, folds.ptr, ncols * 16, hipMemcpyDeviceToHost, ctx->stream));
        if (want_honest)
            BITZ_HIP(hipMemcpyAsync(st + 16 * ncols, colw.ptr, ncols * 16, hipMemcpyDeviceToHost, ctx->stream));
        BITZ_HIP(hipStreamSynchronize(ctx->stream));
        memcpy(fold_bytes.data(), st, ncols * 16);
        if (want_honest) {
            // sum_c fold_c col_w_c mod q: each term as a Montgomery product
            // fold_c col_w_c R^-1 (fold_c < 2^128, col_w_c < q), the sum times R
            ModN m;
            m.set(q);
            for (uint64_t c = 0; c < ncols; c++)
                honest = addmod(honest, m.mont(u128_from(st + 16 * c), u128_from(st + 16 * (ncols + c))), q);
            honest = m.to_mont(honest);
        }
    }
    u128 claimed = claimed_in ? u128_from(claimed_in) : honest;
    if (claimed_out) u128_put(honest, claimed_out);
    if (!proof_out) return BITZ_OK;  // evaluation only
    if (claimed >= q) {
        set_error("BitZ claim: value not reduced");
        return BITZ_ERR_ARG;
    }
    absorb_standalone_claim(outer, q, claimed);
    uint8_t tag[32];
    outer.challenge(tag, 16);
    outer.challenge(tag + 16, 16);

    InnerTranscript it("bitz/bitz-opener/standalone/v1", tag, 32);
    return open_claim(it, phase, q, colw, claimed, fold_bytes, ood_point, ood_y, r0bits != 0, r0_nonce, proof_out);
}

// BitZProver::prove (src/bitz/pcs.rs) of the claim
// sum_{c,b} roww[b] colw[c] bit(c, b) = claimed (mod q) on the inner
// transcript `it`: roww (member) and colw hold the canonical weights on the
// device, fold_bytes the column folds of roww; Round 0's point and value are
// already bound on the outer transcript.
int Prover::open_claim(InnerTranscript& it, PhaseClock& phase, u128 q, DevBuf& colw, u128 claimed,
                       const std::vector<uint8_t>& fold_bytes, const std::vector<G>& ood_point, G ood_y,
                       bool has_r0_nonce, uint64_t r0_nonce, std::vector<uint8_t>* proof_out) {
    int rc;
    const uint64_t nrows = 1ull << t, ncols = 1ull << s;
    phase.mark("fold_round");
    // ---------------- BitZ opening on the inner transcript ----------------
    it.public_message(h->root, 32);
    {
        uint8_t frame[48];
        uint64_t tt = t, ss = s;
        memcpy(frame, &tt, 8);
        memcpy(frame + 8, &ss, 8);
        u128_put(q, frame + 16);
        gput(g, frame + 32);
        it.public_message(frame, 48);
        // LinearClaim encoding (src/bitz/params.rs:273-285): len || row
        // weights || len || column weights || claim, hashed where the row
        // weights already are (the device copy made for the folds)
        G claim_g;
        {
            uint8_t b[16];
            u128_put(claimed, b);
            claim_g = gfrom(b);
        }
        if ((rc = public_weights(it, roww.ptr, nrows, colw.ptr, ncols, claim_g))) return rc;
    }
    DevBuf u1, u2;
    G ip_target;
    if ((rc = reduce_claim(it, phase, h->packed.ptr, fold_bytes, u1, u2, &ip_target))) return rc;
    return prove_lin(it, phase, u1, u2, t, s, ip_target, true, ood_point, ood_y, has_r0_nonce, r0_nonce, proof_out);
}

int Prover::reduce_claim(InnerTranscript& it, PhaseClock& phase, const void* bits,
                         const std::vector<uint8_t>& fold_bytes, DevBuf& u1, DevBuf& u2, G* ip_target) {
    int rc;
    const uint64_t nrows = 1ull << t, ncols = 1ull << s;
    uint8_t gb[16];
    gput(g, gb);
    // fold round (src/bitz/fold.rs:237-263): one prover message per column
    it.prover_messages(fold_bytes.data(), ncols, 16);
    std::vector<G> zeta(s);
    for (auto& z : zeta) {
        uint8_t b[16];
        it.verifier_gf(b);
        z = gfrom(b);
    }
    if (!rowimg.alloc(nrows * 16)) return BITZ_ERR_ALLOC;
    if ((rc = pow_images(ctx, gb, roww.ptr, nrows, rowimg.ptr))) return rc;
    // packed columns for the leaf layer and the combined rows
    uint64_t groups = (ncols + 63) / 64;
    if (!packed.alloc((groups << t) * 8)) return BITZ_ERR_ALLOC;
    if ((rc = pack_columns(ctx, (const uint64_t*)bits, t, s, (uint64_t*)packed.ptr))) return rc;
    // GKR (reduce.rs:100-114)
    std::vector<G> leaf_point;
    G leaf_claim;
    phase.mark("gkr");
    if ((rc = gkr(it, zeta, &leaf_point, &leaf_claim))) return rc;
    phase.mark("inner_sumcheck");
    levels = DevBuf();  // release the product tree
    // exit_from_terminal (reduce.rs:28-58)
    std::vector<G> alpha_c(leaf_point.begin(), leaf_point.begin() + s);
    std::vector<G> alpha_b(leaf_point.begin() + s, leaf_point.end());
    *ip_target = hgf_add(leaf_claim, gone());
    if ((rc = eq_dev(alpha_b, tmp0))) return rc;
    if (!u1.alloc(nrows * 16)) return BITZ_ERR_ALLOC;
    if ((rc = mul_elem(ctx, rowimg.ptr, tmp0.ptr, nrows, u1.ptr, 1))) return rc;
    return eq_dev(alpha_c, u2);
}

int Prover::prove_lin(InnerTranscript& it, PhaseClock& phase, DevBuf& u1, DevBuf& u2, int lt, int ls, G ip_target,
                      bool bind, const std::vector<G>& ood_point, G ood_y, bool has_r0_nonce, uint64_t r0_nonce,
                      std::vector<uint8_t>* proof_out) {
    int rc;
    const uint64_t nrows = 1ull << lt, ncols = 1ull << ls;
    const uint32_t r0bits = has_r0_nonce ? 1u : 0u;
    uint8_t yb[16];
    gput(ood_y, yb);
    // prove_lin / InnerProduct statement (src/bitz/pcs.rs:362-372, 878-888),
    // bound here unless the caller's statement already covers the claim
    if (bind) {
        it.public_message(std::string("bitz/pcs/bit-inner-product/v2"));
        it.public_message(h->root, 32);
        uint64_t pcs40[5] = {(uint64_t)m, (uint64_t)h->log_inv_rate, (uint64_t)h->log_batch, 0, 1};
        it.public_message(pcs40, 40);
        // the statement u1 || u2 (lengths framed): its whole chunks are
        // hashed on the device where u1, u2 live, the rest on the host
        if ((rc = public_weights(it, u1.ptr, nrows, u2.ptr, ncols, ip_target))) return rc;
    }
    it.public_message(std::string("bitz/pcs/inner-product-sumcheck/v1"));
    // row rounds on the column-combined table (a single column: the bits)
    std::vector<G> sc_point;
    G target = ip_target;
    if (ls == 0 && nrows >= 64) {
        // a single column: the committed bits themselves
        if ((rc = bit_sumcheck_rounds(it, u1, nrows, &target, sc_point))) return rc;
    } else {
        DevBuf comb, a2, w2;
        if (!comb.alloc(nrows * 16) || !a2.alloc(nrows * 16) || !w2.alloc(nrows * 16)) return BITZ_ERR_ALLOC;
        if (ls > 0) {
            if ((rc = xi_combined(ctx, (const uint64_t*)packed.ptr, lt, ls, u2.ptr, comb.ptr))) return rc;
        } else {
            if ((rc = bits_to_gf(ctx, (const uint64_t*)h->packed.ptr, nrows, comb.ptr))) return rc;
        }
        if ((rc = sumcheck_rounds(it, comb, u1, a2, w2, nrows, &target, sc_point))) return rc;
    }
    uint8_t evb[32];
    if (ls > 0) {
        uint8_t rs_bytes[16];
        gput(final_w, rs_bytes);
        // column rounds on the row-folded table
        if ((rc = eq_dev(sc_point, tmp0))) return rc;
        DevBuf folded, cols, f2, c2;
        if (!folded.alloc(ncols * 16) || !cols.alloc(ncols * 16) || !f2.alloc(ncols * 16) || !c2.alloc(ncols * 16))
            return BITZ_ERR_ALLOC;
        if ((rc = fold_rows_eq(ctx, (const uint64_t*)h->packed.ptr, lt, ls, tmp0.ptr, folded.ptr))) return rc;
        if ((rc = scale(ctx, u2.ptr, ncols, rs_bytes, cols.ptr))) return rc;
        if ((rc = sumcheck_rounds(it, folded, cols, f2, c2, ncols, &target, sc_point))) return rc;
    }
    gput(final_a, evb);
    gput(final_w, evb + 16);
    G evaluation = gfrom(evb);
    if (!geq(target, hgf_mul(evaluation, gfrom(evb + 16)))) {
        set_error("inner sumcheck: final check failed");
        return BITZ_ERR_ARG;
    }
    it.prover_message(evb, 16);
    // bind_mle_statement (src/bitz/pcs.rs:862-876)
    {
        it.public_message(std::string("bitz/pcs/mle-opening/v1"));
        it.public_message(h->root, 32);
        uint64_t pcs40[5] = {(uint64_t)m, (uint64_t)h->log_inv_rate, (uint64_t)h->log_batch, 0, 1};
        it.public_message(pcs40, 40);
        uint64_t n = sc_point.size();
        it.public_message(&n, 8);
        for (auto& c : sc_point) {
            uint8_t b[16];
            gput(c, b);
            it.public_message(b, 16);
        }
        it.public_message(evb, 16);
    }
    phase.mark("ring_switch");
    // ---------------- ring switch (prove_mle, src/bitz/pcs.rs:439-507) ----------------
    std::vector<G> prefix_pt(sc_point.begin(), sc_point.begin() + 7);
    std::vector<G> suffix_pt(sc_point.begin() + 7, sc_point.end());
    std::vector<G> prefix = host_eq(prefix_pt);
    DevBuf suffix, claims_dev, basis;
    if ((rc = eq_dev(suffix_pt, suffix))) return rc;
    const uint64_t np = 1ull << mp;
    if (!claims_dev.alloc(128 * 16)) return BITZ_ERR_ALLOC;
    if ((rc = fold_1b_rows(ctx, h->packed.ptr, suffix.ptr, np, claims_dev.ptr))) return rc;
    std::vector<G> claims(128);
    if ((rc = download(claims_dev.ptr, claims.data(), 128 * 16))) return rc;
    {
        G cc = gzero();
        for (int i = 0; i < 128; i++) cc = hgf_add(cc, hgf_mul(prefix[i], claims[i]));
        if (!geq(cc, evaluation)) {
            set_error("ring switch: claim check failed");
            return BITZ_ERR_ARG;
        }
    }
    it.public_message(std::string("bitz/pcs/mle-claims/v1"));
    it.prover_message(claims.data(), 128 * 16);
    it.public_message(std::string("bitz/pcs/ring-switch-challenges/v1"));
    std::vector<G> batch_pt(7);
    for (auto& x : batch_pt) {
        uint8_t b[16];
        it.verifier_gf(b);
        x = gfrom(b);
    }
    std::vector<G> bw = host_eq(batch_pt);
    // batch_claims: <transpose(claims), bw>
    G packed_target = gzero();
    for (int b = 0; b < 128; b++) {
        G u = gzero();
        for (int i = 0; i < 128; i++) {
            uint64_t bit = b < 64 ? (claims[i].lo >> b) & 1 : (claims[i].hi >> (b - 64)) & 1;
            if (!bit) continue;
            if (i < 64)
                u.lo |= 1ull << i;
            else
                u.hi |= 1ull << (i - 64);
        }
        packed_target = hgf_add(packed_target, hgf_mul(u, bw[b]));
    }
    if ((rc = upload_gf(bw, tmp1))) return rc;
    if (!basis.alloc(np * 16)) return BITZ_ERR_ALLOC;
    if ((rc = fold_b128(ctx, suffix.ptr, np, tmp1.ptr, basis.ptr))) return rc;
    // the Round-0 claim joins the batch (bind_ood_claim, pcs.rs:904-912)
    {
        it.public_message(std::string("bitz/bitz/round0-claim/v1"));
        uint64_t n = ood_point.size();
        it.public_message(&n, 8);
        for (auto& c : ood_point) {
            uint8_t b[16];
            gput(c, b);
            it.public_message(b, 16);
        }
        it.public_message(yb, 16);
        uint8_t eb[16];
        it.verifier_gf(eb);
        if ((rc = eq_dev(ood_point, tmp0))) return rc;
        if ((rc = axpy(ctx, basis.ptr, tmp0.ptr, np, eb))) return rc;
        packed_target = hgf_add(packed_target, hgf_mul(gfrom(eb), ood_y));
    }
    suffix = DevBuf();

    phase.mark("ligerito");
    // ---------------- Ligerito (recursive_prover_with_basis) ----------------
    Challenger ch{it, ctx};
    ch.observe_label("flock-ligerito-basis-v0");
    ch.observe_f128(packed_target);
    ch.observe_bytes(h->root, 32);
    std::vector<G> sc_msgs;  // (u0, u2) pairs flattened
    std::vector<uint64_t> grinding_nonces, fold_nonces;
    std::vector<G> ood_values;
    DevBuf F[2], B[2];
    int cur = 0;
    if (!F[0].alloc(np * 16) || !B[0].alloc(np * 16) || !F[1].alloc(np * 8) || !B[1].alloc(np * 8))
        return BITZ_ERR_ALLOC;
    BITZ_HIP(hipMemcpyAsync(F[0].ptr, h->packed.ptr, np * 16, hipMemcpyDeviceToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(B[0].ptr, basis.ptr, np * 16, hipMemcpyDeviceToDevice, ctx->stream));
    basis = DevBuf();
    uint64_t fn = np;
    auto msg = [&](const void* f, const void* b, uint64_t n, bool with_eval, G* out) -> int {
        int r_ = lig_msg(ctx, f, b, n, with_eval ? 1 : 0, nullptr);
        if (r_) return r_;
        return result(with_eval ? 3 : 2, out);
    };
    auto observe_msg = [&](G u0, G u2) {
        ch.observe_f128(u0);
        ch.observe_f128(u2);
        sc_msgs.push_back(u0);
        sc_msgs.push_back(u2);
    };
    auto fold_round = [&](int grind_bits, G* r_out) -> int {
        if (grind_bits > 0) {
            uint64_t nonce;
            int r_ = ch.grind((uint32_t)grind_bits, &nonce);
            if (r_) return r_;
            fold_nonces.push_back(nonce);
        }
        G r = ch.sample();
        uint8_t rb[16];
        gput(r, rb);
        int r_ = lig_fold(ctx, F[cur].ptr, B[cur].ptr, fn, rb, F[cur ^ 1].ptr, B[cur ^ 1].ptr);
        if (r_) return r_;
        cur ^= 1;
        fn >>= 1;
        G u[2];
        if ((r_ = msg(F[cur].ptr, B[cur].ptr, fn, false, u))) return r_;
        observe_msg(u[0], u[1]);
        *r_out = r;
        return BITZ_OK;
    };
    auto glue = [&](const void* bnew, G beta) -> int {
        uint8_t bb[16];
        gput(beta, bb);
        return axpy(ctx, B[cur].ptr, bnew, fn, bb);
    };
    G T = packed_target;
    {
        G u[2];
        if ((rc = msg(F[0].ptr, B[0].ptr, fn, false, u))) return rc;
        observe_msg(u[0], u[1]);
    }
    const int r_steps = (int)cfg.levels.size() - 1;
    std::vector<G> r_lane;
    for (int j = 0; j < cfg.initial_k; j++) {
        G r;
        if ((rc = fold_round(std::max(0, cfg.levels[0].fold_grinding_bits - j), &r))) return rc;
        r_lane.push_back(r);
    }
    // commit f1
    DevBuf cw_prev, tree_prev, cw_next, tree_next;
    int n1 = mp - cfg.initial_k;
    uint8_t root_prev[32];
    if ((rc = commit_level(F[cur].ptr, n1, cfg.levels[1].log_num_interleaved, cfg.levels[1].log_inv_rate, cw_prev,
                           tree_prev, root_prev)))
        return rc;
    ch.observe_bytes(root_prev, 32);
    std::vector<std::vector<uint8_t>> recursive_roots = {std::vector<uint8_t>(root_prev, root_prev + 32)};
    auto ood_round = [&](int level, int n_vars) -> int {
        for (int k = 0; k < cfg.levels[level].ood_samples; k++) {
            std::vector<G> z = ch.sample_vec(n_vars);
            int r_ = eq_dev(z, tmp0);
            if (r_) return r_;
            G u[3];
            if ((r_ = msg(F[cur].ptr, tmp0.ptr, fn, true, u))) return r_;
            ch.observe_f128(u[2]);
            ood_values.push_back(u[2]);
            observe_msg(u[0], u[1]);
            G beta = ch.sample();
            if ((r_ = glue(tmp0.ptr, beta))) return r_;
            T = hgf_add(T, hgf_mul(beta, u[2]));
        }
        return BITZ_OK;
    };
    if ((rc = ood_round(1, n1))) return rc;
    // induce a level's basis from its opened rows (induce_sumcheck_poly)
    auto induce_intro = [&](int log_n, const LevelOpen& op, const std::vector<G>& v_ch,
                            const std::vector<uint64_t>& queries, const std::vector<G>& alpha) -> int {
        std::optional<SubPhase> hp(std::in_place, ctx, "sub_lig_induce_host");
        std::vector<G> aw = host_eq(alpha);
        aw.resize(queries.size());
        std::vector<G> eqv = host_eq(v_ch);
        G enforced = gzero();
        for (size_t i = 0; i < queries.size(); i++) {
            G d = gzero();
            for (uint32_t l = 0; l < op.lanes; l++) d = hgf_add(d, hgf_mul(op.rows[i * op.lanes + l], eqv[l]));
            enforced = hgf_add(enforced, hgf_mul(d, aw[i]));
        }
        int r_ = upload_gf(aw, tmp1);
        if (r_) return r_;
        if (!qdev.alloc(queries.size() * 8 + 16)) return BITZ_ERR_ALLOC;
        if ((r_ = upload_async(ctx, queries.data(), queries.size() * 8, qdev.ptr))) return r_;
        // s_k(v_k) and inverses (eval_sk_at_vks)
        std::vector<G> sks, inv;
        sk_at_vks(log_n, &sks, &inv);
        std::vector<G> both(sks);
        both.insert(both.end(), inv.begin(), inv.end());
        if ((r_ = upload_gf(both, tmp2))) return r_;
        if (!tmp0.alloc(16ull << log_n)) return BITZ_ERR_ALLOC;
        hp.reset();
        if ((r_ = induce_basis(ctx, (const uint64_t*)qdev.ptr, (uint32_t)queries.size(), log_n, tmp2.ptr,
                               (const uint4*)tmp2.ptr + (log_n + 1), tmp1.ptr, tmp0.ptr)))
            return r_;
        G u[2];
        if ((r_ = msg(F[cur].ptr, tmp0.ptr, fn, false, u))) return r_;
        observe_msg(u[0], u[1]);
        G beta = ch.sample();
        if ((r_ = glue(tmp0.ptr, beta))) return r_;
        T = hgf_add(T, hgf_mul(beta, enforced));
        return BITZ_OK;
    };
    // L0 queries
    uint64_t nonce;
    if ((rc = ch.grind((uint32_t)cfg.levels[0].grinding_bits, &nonce))) return rc;
    grinding_nonces.push_back(nonce);
    const LigLevel& L0 = cfg.levels[0];
    uint64_t block0 = 1ull << (L0.log_msg_cols + L0.log_inv_rate);
    std::vector<uint64_t> q0 = ch.queries(block0, L0.queries);
    std::vector<G> alpha0 = ch.sample_vec(ceil_log2(L0.queries));
    LevelOpen open0;
    if ((rc = open_level(h->codeword.ptr, h->tree.ptr, block0, 1u << cfg.initial_k, q0, &open0))) return rc;
    if ((rc = induce_intro(n1, open0, r_lane, q0, alpha0))) return rc;

    std::vector<LevelOpen> rec_opens;
    LevelOpen final_open;
    std::vector<G> yr;
    int prev_level = 1;  // level index of the commitment held in cw_prev
    for (int i = 0; i < r_steps; i++) {
        int k_i = cfg.levels[i + 1].k_recursive;
        std::vector<G> level_rs;
        for (int j = 0; j < k_i; j++) {
            G r;
            if ((rc = fold_round(std::max(0, cfg.levels[i + 1].fold_grinding_bits - j), &r))) return rc;
            level_rs.push_back(r);
        }
        const LigLevel& Lp = cfg.levels[prev_level];
        uint64_t block_prev = 1ull << (Lp.log_msg_cols + Lp.log_inv_rate);
        if (i == r_steps - 1) {
            yr.resize(fn);
            if ((rc = download(F[cur].ptr, yr.data(), fn * 16))) return rc;
            for (auto& v : yr) ch.observe_f128(v);
            if ((rc = ch.grind((uint32_t)cfg.levels[i + 1].grinding_bits, &nonce))) return rc;
            grinding_nonces.push_back(nonce);
            std::vector<uint64_t> ql = ch.queries(block_prev, cfg.levels[i + 1].queries);
            if ((rc = open_level(cw_prev.ptr, tree_prev.ptr, block_prev, 1u << Lp.log_num_interleaved, ql,
                                 &final_open)))
                return rc;
            break;
        }
        int n_next = ceil_log2(fn);
        const LigLevel& Ln = cfg.levels[i + 2];
        uint8_t root_next[32];
        if ((rc = commit_level(F[cur].ptr, n_next, Ln.log_num_interleaved, Ln.log_inv_rate, cw_next, tree_next,
                               root_next)))
            return rc;
        ch.observe_bytes(root_next, 32);
        recursive_roots.emplace_back(root_next, root_next + 32);
        if ((rc = ood_round(i + 2, n_next))) return rc;
        if ((rc = ch.grind((uint32_t)cfg.levels[i + 1].grinding_bits, &nonce))) return rc;
        grinding_nonces.push_back(nonce);
        std::vector<uint64_t> qi = ch.queries(block_prev, cfg.levels[i + 1].queries);
        std::vector<G> alphai = ch.sample_vec(ceil_log2(cfg.levels[i + 1].queries));
        LevelOpen op;
        if ((rc = open_level(cw_prev.ptr, tree_prev.ptr, block_prev, 1u << Lp.log_num_interleaved, qi, &op)))
            return rc;
        if ((rc = induce_intro(n_next, op, level_rs, qi, alphai))) return rc;
        rec_opens.push_back(std::move(op));
        std::swap(cw_prev, cw_next);
        std::swap(tree_prev, tree_next);
        prev_level = i + 2;
    }
    phase.mark("proof_bytes");
    // ---------------- LigeritoProof as a hint (write_opening_proof) ----------------
    Writer w;
    w.bytes(h->root, 32);
    write_open(w, open0);
    w.u64(recursive_roots.size());
    for (auto& r : recursive_roots) w.bytes(r.data(), 32);
    w.u64(rec_opens.size());
    for (auto& o : rec_opens) write_open(w, o);
    w.u64(yr.size());
    for (auto& v : yr) w.gf(v);
    write_open(w, final_open);
    w.u64(sc_msgs.size() / 2);
    for (auto& v : sc_msgs) w.gf(v);
    w.u64(grinding_nonces.size());
    for (auto v : grinding_nonces) w.u64(v);
    w.u64(ood_values.size());
    for (auto& v : ood_values) w.gf(v);
    w.u64(fold_nonces.size());
    for (auto v : fold_nonces) w.u64(v);
    it.hint_bytes(w.b.data(), w.b.size());
    // BitZOpeningProof::to_bytes (bitz_opener.rs:300-323)
    Writer out;
    out.u64(it.narg().size());
    out.bytes(it.narg().data(), it.narg().size());
    out.u64(it.hints().size());
    out.bytes(it.hints().data(), it.hints().size());
    out.b.push_back(1);
    out.bytes(yb, 16);
    if (r0bits) {
        out.b.push_back(1);
        out.u64(r0_nonce);
    } else {
        out.b.push_back(0);
    }
    proof_out->swap(out.b);
    return BITZ_OK;
}

// ------------------------------------------------------------- the verifier
// verify_standalone (src/piop/spartan/protocol/bitz_opener.rs:709-747) ->
// BitZVerifier::verify (src/bitz/mod.rs:218-255) -> the succinct recursive
// Ligerito verifier (pcs/ligerito.rs:3967-4440, src/bitz/pcs.rs:508-618).
// Transcript checks, Merkle multi-proofs and the residual (every basis
// component in closed form, residual.h) run on the host; the tables the size
// of the claim's grid (fold images, eq tables over its rows and columns and
// their folds) are built by the prover's kernels on the context's device.
#define V_CHECK(cond, msg)                                  \
    do {                                                    \
        if (!(cond)) {                                      \
            set_error(std::string("verify: ") + (msg));     \
            return BITZ_ERR_VERIFY;                         \
        }                                                   \
    } while (0)

// the inner transcript on the verifier's side: prover messages are read
// from the narg string and absorbed
struct InnerReader {
    Transcript t;
    const uint8_t* narg;
    size_t len, pos = 0;
    InnerReader(const std::string& session, const uint8_t* inst, size_t ilen, const uint8_t* n, size_t l)
        : narg(n), len(l) {
        t.absorb_slice(std::string("bitz/v1"));
        t.absorb_slice(session);
        t.absorb_slice(inst, ilen);
    }
    void public_message(const void* d, size_t n) { t.absorb_slice(d, n); }
    void public_message(const std::string& s) { t.absorb_slice(s); }
    Transcript::SlicePlan public_plan(size_t l) const { return t.slice_plan(l); }
    void public_message_runs(size_t l, const Transcript::SlicePlan& plan, const uint8_t* lead, const uint32_t* cvs,
                             const uint8_t* tail) {
        t.absorb_slice_runs(l, plan, lead, cvs, tail);
    }
    const uint8_t* read(size_t n) {
        if (n > len - pos) return nullptr;
        const uint8_t* p = narg + pos;
        t.absorb_slice(p, n);
        pos += n;
        return p;
    }
    // count prover messages of `each` bytes
    const uint8_t* read_many(size_t count, size_t each) {
        if (count > (len - pos) / std::max<size_t>(each, 1)) return nullptr;
        const uint8_t* p = narg + pos;
        t.absorb_slices(p, count, each);
        pos += count * each;
        return p;
    }
    G gf() {
        uint8_t b[16];
        t.challenge(b, 16);
        return gfrom(b);
    }
};

struct ByteReader {
    const uint8_t* p;
    size_t n, pos = 0;
    bool ok = true;
    const uint8_t* take(size_t k) {
        if (!ok || k > n - pos) {
            ok = false;
            return nullptr;
        }
        const uint8_t* r = p + pos;
        pos += k;
        return r;
    }
    uint64_t u64() {
        const uint8_t* r = take(8);
        uint64_t v = 0;
        if (r) memcpy(&v, r, 8);
        return v;
    }
    G gf() {
        const uint8_t* r = take(16);
        return r ? gfrom(r) : gzero();
    }
    // a count of items of `each` bytes that the remaining bytes can hold
    bool count(uint64_t c, size_t each) { return ok && c <= (n - pos) / std::max<size_t>(each, 1); }
};

struct OpenData {
    std::vector<std::vector<G>> rows;
    std::vector<uint8_t> proof;  // 32-byte nodes
};

bool read_open(ByteReader& rd, OpenData* o) {
    const uint64_t nq = rd.u64();
    if (!rd.count(nq, 8)) return false;
    o->rows.resize(nq);
    for (auto& row : o->rows) {
        const uint64_t lanes = rd.u64();
        if (!rd.count(lanes, 16)) return false;
        row.resize(lanes);
        for (auto& v : row) v = rd.gf();
    }
    const uint64_t np = rd.u64();
    if (!rd.count(np, 32)) return false;
    const uint8_t* b = rd.take(32 * np);
    if (!b && np) return false;
    o->proof.assign(b, b + 32 * np);
    return rd.ok;
}

// the Ligerito proof of the hint stream (field order of ligerito.rs:1528-1551)
struct LigProofData {
    uint8_t initial_root[32];
    OpenData initial, fin;
    std::vector<std::array<uint8_t, 32>> roots;
    std::vector<OpenData> rec;
    std::vector<G> yr, ood;
    std::vector<std::pair<G, G>> msgs;
    std::vector<uint64_t> grind, fold_nonces;
};

bool parse_ligerito(const uint8_t* b, size_t n, LigProofData* p) {
    ByteReader rd{b, n};
    const uint8_t* r = rd.take(32);
    if (!r) return false;
    memcpy(p->initial_root, r, 32);
    if (!read_open(rd, &p->initial)) return false;
    uint64_t c = rd.u64();
    if (!rd.count(c, 32)) return false;
    p->roots.resize(c);
    for (auto& x : p->roots) memcpy(x.data(), rd.take(32), 32);
    c = rd.u64();
    if (!rd.count(c, 16)) return false;
    p->rec.resize(c);
    for (auto& o : p->rec)
        if (!read_open(rd, &o)) return false;
    c = rd.u64();
    if (!rd.count(c, 16)) return false;
    p->yr.resize(c);
    for (auto& v : p->yr) v = rd.gf();
    if (!read_open(rd, &p->fin)) return false;
    c = rd.u64();
    if (!rd.count(c, 32)) return false;
    p->msgs.resize(c);
    for (auto& m : p->msgs) {
        m.first = rd.gf();
        m.second = rd.gf();
    }
    c = rd.u64();
    if (!rd.count(c, 8)) return false;
    p->grind.resize(c);
    for (auto& v : p->grind) v = rd.u64();
    c = rd.u64();
    if (!rd.count(c, 16)) return false;
    p->ood.resize(c);
    for (auto& v : p->ood) v = rd.gf();
    c = rd.u64();
    if (!rd.count(c, 8)) return false;
    p->fold_nonces.resize(c);
    for (auto& v : p->fold_nonces) v = rd.u64();
    return rd.ok && rd.pos == n;
}

// BLAKE3 chaining value of a leaf of at most 1024 bytes hashed as chunk 0
// without the ROOT flag (blake3_leaf_cv, merkle.rs:192-196)
void leaf_cv(const uint8_t* data, size_t len, uint8_t out[32]) {
    static const uint32_t kIV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                    0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
    uint32_t cv[8], o[16];
    memcpy(cv, kIV, 32);
    const size_t blocks = len ? (len + 63) / 64 : 1;
    for (size_t b = 0; b < blocks; b++) {
        uint8_t block[64] = {0};
        const size_t bl = std::min<size_t>(64, len - 64 * b);
        if (len) memcpy(block, data + 64 * b, bl);
        const uint32_t flags = (b == 0 ? 1u : 0u) | (b + 1 == blocks ? 2u : 0u);  // CHUNK_START, CHUNK_END
        blake3_compress_host(cv, block, 0, (uint32_t)(len ? bl : 0), flags, o);
        memcpy(cv, o, 32);
    }
    memcpy(out, cv, 32);
}

// verify_merkle_multi_proof (merkle.rs:749-828): sorted distinct leaf
// positions, their chaining values, and the sibling nodes the octopus proof
// lists level by level
bool merkle_verify_multi(const uint8_t root[32], uint64_t num_leaves, const std::vector<uint64_t>& pos,
                         std::vector<std::array<uint8_t, 32>> hashes, const std::vector<uint8_t>& proof) {
    if (num_leaves == 0 || (num_leaves & (num_leaves - 1))) return false;
    if (pos.empty()) return proof.empty();
    for (size_t i = 0; i < pos.size(); i++)
        if (pos[i] >= num_leaves || (i && pos[i - 1] >= pos[i])) return false;
    std::vector<uint64_t> ap(pos);
    size_t used = 0;
    const size_t n_proof = proof.size() / 32;
    for (uint64_t level = num_leaves; level > 1; level >>= 1) {
        std::vector<uint64_t> np;
        std::vector<std::array<uint8_t, 32>> nh;
        for (size_t i = 0; i < ap.size();) {
            const uint64_t p = ap[i];
            uint32_t l[8], r[8], c[8];
            if (i + 1 < ap.size() && ap[i + 1] == (p ^ 1)) {
                memcpy(l, hashes[i].data(), 32);
                memcpy(r, hashes[i + 1].data(), 32);
                i += 2;
            } else {
                if (used >= n_proof) return false;
                const uint8_t* sib = proof.data() + 32 * used++;
                if ((p & 1) == 0) {
                    memcpy(l, hashes[i].data(), 32);
                    memcpy(r, sib, 32);
                } else {
                    memcpy(l, sib, 32);
                    memcpy(r, hashes[i].data(), 32);
                }
                i += 1;
            }
            blake3_parent_cv(l, r, c);
            std::array<uint8_t, 32> h;
            memcpy(h.data(), c, 32);
            np.push_back(p >> 1);
            nh.push_back(h);
        }
        ap.swap(np);
        hashes.swap(nh);
    }
    return used == n_proof && memcmp(hashes[0].data(), root, 32) == 0;
}

// BLAKE3(prefix16 || seed16 || nonce_le) has at least `bits` leading zero
// bits (bytes in order, most significant bit first)
bool pow_valid(const uint8_t prefix[16], const uint8_t seed[16], uint64_t nonce, uint32_t bits) {
    uint8_t m[40], h[32];
    memcpy(m, prefix, 16);
    memcpy(m + 16, seed, 16);
    memcpy(m + 32, &nonce, 8);
    blake3_hash_host(m, 40, h);
    uint32_t z = 0;
    for (int i = 0; i < 32 && z < bits; i++) {
        if (h[i] == 0) {
            z += 8;
            continue;
        }
        z += (uint32_t)__builtin_clz((unsigned)h[i]) - 24;
        break;
    }
    return z >= bits;
}

struct Verifier {
    bitz_ctx* ctx;
    DevBuf scratch_a, scratch_b, pt;

    int download(const void* dev, void* host, size_t bytes) {
        BITZ_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
        BITZ_HIP(hipStreamSynchronize(ctx->stream));
        return BITZ_OK;
    }
    int eq_into(const std::vector<G>& r, DevBuf& buf) {
        BITZ_REQUIRE((int)r.size() <= kMaxTensorN, BITZ_ERR_SHAPE, "verify: eq table over too many coordinates");
        if (!buf.alloc(16ull << r.size())) return BITZ_ERR_ALLOC;
        return eq_args(ctx, r.data(), (int)r.size(), buf.ptr);
    }
    int dot_value(const void* a, const void* b, uint64_t n, G* out) {
        int rc = dot(ctx, a, b, n, nullptr);
        if (rc) return rc;
        uint8_t v[16];
        if ((rc = read_result(ctx, 1, v))) return rc;
        *out = gfrom(v);
        return BITZ_OK;
    }
    // v (2^n entries) folded LSB first by rs (fold_lsb, ligerito.rs:2458-2497)
    // into 2^(n - |rs|) host values
    int fold_chain(const void* v, uint64_t n_entries, const std::vector<G>& rs, size_t first, std::vector<G>* out) {
        uint64_t n = n_entries;
        BITZ_REQUIRE(rs.size() - first < 64 && (n >> (rs.size() - first)) >= 1, BITZ_ERR_SHAPE, "verify: fold chain");
        if (!scratch_a.alloc(16 * std::max<uint64_t>(n / 2, 1)) || !scratch_b.alloc(16 * std::max<uint64_t>(n / 4, 1)))
            return BITZ_ERR_ALLOC;
        const void* cur = v;
        int rc;
        for (size_t k = first; k < rs.size(); k++) {
            uint8_t rb[16];
            gput(rs[k], rb);
            void* dst = ((k - first) & 1) ? scratch_b.ptr : scratch_a.ptr;
            if ((rc = fold_lsb(ctx, cur, n, rb, dst))) return rc;
            cur = dst;
            n >>= 1;
        }
        out->resize(n);
        return download(cur, out->data(), 16 * n);
    }
};

int verify_claim_impl(bitz_ctx* ctx, InnerReader& it, const uint8_t root[32], int t, int s, const LigConfig& cfg,
                      u128 q, DevBuf& roww, DevBuf& colw, u128 claimed, const std::vector<G>& opt, G ood_y,
                      uint64_t narg_len, const uint8_t* hints, uint64_t hints_len);
// receive_fold and the GKR verification over a grid (t, s): the factored
// claim u1 (2^t), u2 (2^s) and its target
int verify_reduce_impl(bitz_ctx* ctx, InnerReader& it, int t, int s, u128 q, DevBuf& roww, DevBuf& colw,
                       u128 claimed, DevBuf& u1, DevBuf& u2, G* target_out);
// Pcs::verify_lin of sum u1[b] u2[c] bit(c, b) = target over the committed
// bits seen as the grid (t, s); the statement bound only when `bind`
int verify_lin_impl(bitz_ctx* ctx, InnerReader& it, const uint8_t root[32], int t, int s, const LigConfig& cfg,
                    DevBuf& u1, DevBuf& u2, G target, bool bind, const std::vector<G>& opt, G ood_y,
                    uint64_t narg_len, const uint8_t* hints, uint64_t hints_len);

int verify_standalone_impl(bitz_ctx* ctx, const uint8_t root[32], int t, int s, int r0, int k0, u128 claimed,
                           const uint8_t* proof, size_t plen) {
    int rc;
    const int m = t + s, mp = m - 7;
    LigConfig cfg;
    {
        std::string err = custom_johnson_config(m, r0, k0, 100, &cfg);
        BITZ_REQUIRE(err.empty(), BITZ_ERR_SHAPE, "Ligerito configuration: " + err);
        BITZ_REQUIRE(t >= 7 && s >= 1 && cfg.levels.size() >= 2 && mp <= kMaxTensorN, BITZ_ERR_SHAPE,
                     "verify: unsupported shape");
    }
    static G gen = smallest_generator();
    const G g = gen;
    Verifier V{ctx};
    // ---- proof tail (BitZOpeningProof::from_bytes, bitz_opener.rs:325-372) ----
    ByteReader pr{proof, plen};
    const uint64_t narg_len = pr.u64();
    V_CHECK(pr.count(narg_len, 1), "truncated proof");
    const uint8_t* narg = pr.take(narg_len);
    const uint64_t hints_len = pr.u64();
    V_CHECK(pr.count(hints_len, 1), "truncated proof");
    const uint8_t* hints = pr.take(hints_len);
    const uint8_t* tag = pr.take(1);
    V_CHECK(tag && tag[0] == 1, "missing Round-0 claim");
    const uint8_t* yb = pr.take(16);
    V_CHECK(yb != nullptr, "truncated proof");
    const G ood_y = gfrom(yb);
    const uint8_t* tag2 = pr.take(1);
    V_CHECK(tag2 && (tag2[0] == 0 || tag2[0] == 1), "truncated proof");
    bool has_nonce = tag2[0] == 1;
    uint64_t r0_nonce = has_nonce ? pr.u64() : 0;
    V_CHECK(pr.ok && pr.pos == plen, "trailing proof bytes");
    // ---- outer transcript: statement, Round 0, instance ----
    Transcript outer;
    uint8_t gb[16];
    gput(g, gb);
    bind_standalone_statement(outer, root, cfg, t, s, gb);
    const uint32_t r0bits = (uint32_t)cfg.round0_grinding_bits;
    ood_round_header(outer, mp, r0bits);
    if (r0bits) {
        // verify_and_absorb (src/piop/spartan/grinding.rs:219-236)
        V_CHECK(has_nonce, "missing Round-0 nonce");
        uint8_t seed[32];
        grinding_seed(outer, "bitz/core/ood-round-grinding/v1", 0, r0bits, seed);
        const bool ok = pow_valid(seed, seed + 16, r0_nonce, r0bits);
        absorb_grinding_nonce(outer, r0_nonce);
        V_CHECK(ok, "Round-0 proof of work");
    } else {
        V_CHECK(!has_nonce, "unexpected Round-0 nonce");
    }
    uint8_t zb[16];
    outer.challenge(zb, 16);
    std::vector<G> opt(mp);
    {
        G cur = gfrom(zb);
        for (int i = 0; i < mp; i++) {
            opt[i] = cur;
            cur = hgf_mul(cur, cur);
        }
    }
    {
        uint8_t tagged[17];
        tagged[0] = 0x50;
        memcpy(tagged + 1, yb, 16);
        outer.absorb_slice(tagged, 17);
    }
    u128 q;
    {
        std::string err;
        V_CHECK(sample_prime_with_bits(outer, standalone_q_bits(t), &q, &err), "prime search");
    }
    std::vector<u128> r1(t), r2(s);
    for (auto& x : r1) x = sample_residue(outer, q);
    for (auto& x : r2) x = sample_residue(outer, q);
    V_CHECK(claimed < q, "claim not reduced");
    absorb_standalone_claim(outer, q, claimed);
    uint8_t tagb[32];
    outer.challenge(tagb, 16);
    outer.challenge(tagb + 16, 16);
    const uint64_t nrows = 1ull << t, ncols = 1ull << s;
    DevBuf roww, colw;
    if (!roww.alloc(nrows * 16) || !colw.alloc(ncols * 16)) return BITZ_ERR_ALLOC;
    if ((rc = eq_table_mod_q_dev(ctx, r1, q, roww.ptr))) return rc;
    if ((rc = eq_table_mod_q_dev(ctx, r2, q, colw.ptr))) return rc;
    InnerReader it("bitz/bitz-opener/standalone/v1", tagb, 32, narg, narg_len);
    return verify_claim_impl(ctx, it, root, t, s, cfg, q, roww, colw, claimed, opt, ood_y, narg_len, hints,
                             hints_len);
}

// BitZVerifier::verify (src/bitz/mod.rs:218-255) of the claim
// sum_{c,b} roww[b] colw[c] bit(c, b) = claimed (mod q), the canonical
// weights on the device, on the inner transcript `it` (reading the narg
// string of narg_len bytes); Round 0's point and value are bound on the
// outer transcript.
int verify_claim_impl(bitz_ctx* ctx, InnerReader& it, const uint8_t root[32], int t, int s, const LigConfig& cfg,
                      u128 q, DevBuf& roww, DevBuf& colw, u128 claimed, const std::vector<G>& opt, G ood_y,
                      uint64_t narg_len, const uint8_t* hints, uint64_t hints_len) {
    int rc;
    static G gen = smallest_generator();
    // ---- inner transcript: instance, fold round ----
    it.public_message(root, 32);
    {
        uint8_t frame[48];
        uint64_t tt = t, ss = s;
        memcpy(frame, &tt, 8);
        memcpy(frame + 8, &ss, 8);
        u128_put(q, frame + 16);
        gput(gen, frame + 32);
        it.public_message(frame, 48);
    }
    {
        uint8_t b[16];
        u128_put(claimed, b);
        if ((rc = public_weights_dev(ctx, it, roww.ptr, 1ull << t, colw.ptr, 1ull << s, gfrom(b)))) return rc;
    }
    DevBuf u1, u2;
    G target;
    if ((rc = verify_reduce_impl(ctx, it, t, s, q, roww, colw, claimed, u1, u2, &target))) return rc;
    return verify_lin_impl(ctx, it, root, t, s, cfg, u1, u2, target, true, opt, ood_y, narg_len, hints, hints_len);
}

int verify_reduce_impl(bitz_ctx* ctx, InnerReader& it, int t, int s, u128 q, DevBuf& roww, DevBuf& colw,
                       u128 claimed, DevBuf& u1, DevBuf& u2, G* target_out) {
    int rc;
    static G gen = smallest_generator();
    const G g = gen;
    Verifier V{ctx};
    uint8_t gb[16];
    gput(g, gb);
    const uint64_t nrows = 1ull << t, ncols = 1ull << s;
    DevBuf folds, img, tmp;
    // img holds the column images (2^s) and later the row images (2^t)
    if (!folds.alloc(ncols * 16) || !img.alloc(std::max(nrows, ncols) * 16)) return BITZ_ERR_ALLOC;
    // fold round (fold.rs receive_fold): range bound, reconstruction
    const uint8_t* fb = it.read_many(ncols, 16);
    V_CHECK(fb != nullptr, "narg string exhausted");
    {
        std::vector<uint8_t> cw(ncols * 16);
        if ((rc = V.download(colw.ptr, cw.data(), cw.size()))) return rc;
        ModN mq;
        mq.set(q);
        u128 acc = 0;
        for (uint64_t c = 0; c < ncols; c++) {
            const u128 f = u128_from(fb + 16 * c);
            // f <= 2^t (q - 1)
            const u128 hi = f >> t, lo = f & (((u128)1 << t) - 1);
            V_CHECK(hi < q - 1 || (hi == q - 1 && lo == 0), "fold out of range");
            acc = addmod(acc, mq.mont(f, u128_from(&cw[16 * c])), q);
        }
        V_CHECK(mq.to_mont(acc) == claimed, "folds do not reconstruct the claim");
    }
    std::vector<G> zeta(s);
    for (auto& z : zeta) z = it.gf();
    if ((rc = upload_async(ctx, fb, ncols * 16, folds.ptr))) return rc;
    if ((rc = pow_images(ctx, gb, folds.ptr, ncols, img.ptr))) return rc;
    G claim;
    if ((rc = V.eq_into(zeta, tmp))) return rc;
    if ((rc = V.dot_value(img.ptr, tmp.ptr, ncols, &claim))) return rc;
    // ---- GKR (gpgkr_verify, gkr.rs:337-383) ----
    std::vector<G> internal(zeta.rbegin(), zeta.rend());
    for (int layer = 0; layer < t; layer++) {
        G prefix = gone();
        std::vector<G> nxt;
        for (const G z : internal) {
            const uint8_t* b = it.read(32);
            V_CHECK(b != nullptr, "narg string exhausted");
            const G se = gfrom(b), si = gfrom(b + 16);
            G s0, s1;
            if (geq(z, gzero())) {
                s0 = claim;
                s1 = se;
            } else {
                s0 = se;
                s1 = hgf_mul(hgf_add(claim, hgf_mul(hgf_add(gone(), z), se)), hgf_inv(z));
            }
            const G rr = it.gf();
            nxt.push_back(rr);
            const G f = eq_factor(rr, z);
            claim = hgf_mul(f, hgf_add(s0, hgf_mul(rr, hgf_add(hgf_add(s1, s0), hgf_mul(hgf_add(rr, gone()), si)))));
            prefix = hgf_mul(prefix, f);
        }
        const uint8_t* b = it.read(32);
        V_CHECK(b != nullptr, "narg string exhausted");
        const G el = gfrom(b), er = gfrom(b + 16);
        V_CHECK(geq(hgf_mul(prefix, hgf_mul(el, er)), claim), "GKR layer check");
        const G rc_ = it.gf();
        internal.clear();
        internal.push_back(rc_);
        internal.insert(internal.end(), nxt.begin(), nxt.end());
        claim = hgf_add(el, hgf_mul(rc_, hgf_add(er, el)));
    }
    std::vector<G> point(internal.rbegin(), internal.rend());
    std::vector<G> alpha_c(point.begin(), point.begin() + s), alpha_b(point.begin() + s, point.end());
    // ---- inner-product statement (exit_from_terminal, reduce.rs:28-58) ----
    if ((rc = pow_images(ctx, gb, roww.ptr, nrows, img.ptr))) return rc;
    if ((rc = V.eq_into(alpha_b, tmp))) return rc;
    if (!u1.alloc(nrows * 16)) return BITZ_ERR_ALLOC;
    if ((rc = mul_elem(ctx, img.ptr, tmp.ptr, nrows, u1.ptr, 1))) return rc;
    if ((rc = V.eq_into(alpha_c, u2))) return rc;
    *target_out = hgf_add(claim, gone());
    return BITZ_OK;
}

int verify_lin_impl(bitz_ctx* ctx, InnerReader& it, const uint8_t root[32], int t, int s, const LigConfig& cfg,
                    DevBuf& u1, DevBuf& u2, G target, bool bind, const std::vector<G>& opt, G ood_y,
                    uint64_t narg_len, const uint8_t* hints, uint64_t hints_len) {
    int rc;
    const int m = cfg.m, mp = m - 7;
    const int r0 = cfg.levels[0].log_inv_rate, k0 = cfg.initial_k;
    Verifier V{ctx};
    uint8_t yb[16];
    gput(ood_y, yb);
    const uint64_t nrows = 1ull << t, ncols = 1ull << s;
    const uint64_t pcs40[5] = {(uint64_t)m, (uint64_t)r0, (uint64_t)k0, 0, 1};
    if (bind) {
        it.public_message(std::string("bitz/pcs/bit-inner-product/v2"));
        it.public_message(root, 32);
        it.public_message(pcs40, 40);
        if ((rc = public_weights_dev(ctx, it, u1.ptr, nrows, u2.ptr, ncols, target))) return rc;
    }
    it.public_message(std::string("bitz/pcs/inner-product-sumcheck/v1"));
    std::vector<G> sc_point;
    for (int round = 0; round < t + s; round++) {
        const uint8_t* b = it.read(48);
        V_CHECK(b != nullptr, "narg string exhausted");
        const G c0 = gfrom(b), c1 = gfrom(b + 16), c2 = gfrom(b + 32);
        V_CHECK(geq(hgf_add(c1, c2), target), "inner sumcheck round");
        const G r = it.gf();
        target = hgf_add(c0, hgf_mul(r, hgf_add(c1, hgf_mul(r, c2))));
        sc_point.push_back(r);
    }
    std::vector<G> rows_w, cols_w;
    if ((rc = V.fold_chain(u1.ptr, nrows, std::vector<G>(sc_point.begin(), sc_point.begin() + t), 0, &rows_w)))
        return rc;
    if ((rc = V.fold_chain(u2.ptr, ncols, std::vector<G>(sc_point.begin() + t, sc_point.end()), 0, &cols_w)))
        return rc;
    const uint8_t* evb = it.read(16);
    V_CHECK(evb != nullptr, "narg string exhausted");
    const G evaluation = gfrom(evb);
    V_CHECK(geq(target, hgf_mul(evaluation, hgf_mul(rows_w[0], cols_w[0]))), "inner sumcheck final check");
    it.public_message(std::string("bitz/pcs/mle-opening/v1"));
    it.public_message(root, 32);
    it.public_message(pcs40, 40);
    {
        uint64_t n = sc_point.size();
        it.public_message(&n, 8);
        for (auto& c : sc_point) {
            uint8_t b[16];
            gput(c, b);
            it.public_message(b, 16);
        }
        it.public_message(evb, 16);
    }
    // ---- ring switch (verify_mle, pcs.rs:508-560) ----
    const std::vector<G> prefix = host_eq(std::vector<G>(sc_point.begin(), sc_point.begin() + 7));
    it.public_message(std::string("bitz/pcs/mle-claims/v1"));
    const uint8_t* cb = it.read(128 * 16);
    V_CHECK(cb != nullptr, "narg string exhausted");
    std::vector<G> claims(128);
    G cc = gzero();
    for (int i = 0; i < 128; i++) {
        claims[i] = gfrom(cb + 16 * i);
        cc = hgf_add(cc, hgf_mul(prefix[i], claims[i]));
    }
    V_CHECK(geq(cc, evaluation), "ring-switch claim check");
    it.public_message(std::string("bitz/pcs/ring-switch-challenges/v1"));
    std::vector<G> batch_pt(7);
    for (auto& x : batch_pt) x = it.gf();
    const std::vector<G> bw = host_eq(batch_pt);
    G packed_target = gzero();
    for (int b = 0; b < 128; b++) {
        G u = gzero();
        for (int i = 0; i < 128; i++) {
            const uint64_t bit = b < 64 ? (claims[i].lo >> b) & 1 : (claims[i].hi >> (b - 64)) & 1;
            if (!bit) continue;
            if (i < 64)
                u.lo |= 1ull << i;
            else
                u.hi |= 1ull << (i - 64);
        }
        packed_target = hgf_add(packed_target, hgf_mul(u, bw[b]));
    }
    it.public_message(std::string("bitz/bitz/round0-claim/v1"));
    {
        uint64_t n = opt.size();
        it.public_message(&n, 8);
        for (auto& c : opt) {
            uint8_t b[16];
            gput(c, b);
            it.public_message(b, 16);
        }
        it.public_message(yb, 16);
    }
    const G eta = it.gf();
    packed_target = hgf_add(packed_target, hgf_mul(eta, ood_y));
    // the initial basis b(x) = sum_i bit_i(eq(suffix, x)) bw[i] + eta eq(opt, x)
    // over mp coordinates is never built: the final check evaluates it at the
    // residual in closed form (residual.h)
    const std::vector<G> suffix_pt(sc_point.begin() + 7, sc_point.end());
    V_CHECK((int)suffix_pt.size() == mp && (int)opt.size() == mp, "Round-0 point length");
    // ---- the Ligerito proof of the hint stream ----
    V_CHECK(hints_len >= 4, "missing Ligerito hint");
    uint32_t ln;
    memcpy(&ln, hints, 4);
    V_CHECK(4 + (uint64_t)ln == hints_len, "hint length");
    LigProofData lp;
    V_CHECK(parse_ligerito(hints + 4, ln, &lp), "malformed Ligerito proof");
    const auto& lv = cfg.levels;
    V_CHECK(memcmp(lp.initial_root, root, 32) == 0, "initial root");
    size_t mi = 0, fni = 0, gni = 0, oi = 0;
    bool first_obs = true;
    auto observe = [&](G v) -> bool {
        uint8_t b[16];
        gput(v, b);
        if (first_obs) {
            first_obs = false;
            it.public_message(b, 16);
            return true;
        }
        const uint8_t* r = it.read(16);
        return r && memcmp(r, b, 16) == 0;
    };
    struct Quad {
        G c, b, a;
    };
    auto quad = [](std::pair<G, G> msg, G t_r) { return Quad{msg.first, hgf_add(t_r, msg.second), msg.second}; };
    auto qeval = [](const Quad& qd, G r) { return hgf_add(qd.c, hgf_add(hgf_mul(r, qd.b), hgf_mul(hgf_mul(r, r), qd.a))); };
    auto qfold = [](const Quad& p1, const Quad& p2, G al) {
        return Quad{hgf_add(p1.c, hgf_mul(al, p2.c)), hgf_add(p1.b, hgf_mul(al, p2.b)), hgf_add(p1.a, hgf_mul(al, p2.a))};
    };
    static const uint8_t kPcsPow[16] = {'b', 'i', 't', 'z', '-', 'p', 'c', 's', '-', 'p', 'o', 'w', '-', 'v', '1', 0};
    std::string fail;
    auto next_msg = [&](std::pair<G, G>* m_) -> bool {
        if (mi >= lp.msgs.size()) {
            fail = "sumcheck transcript exhausted";
            return false;
        }
        *m_ = lp.msgs[mi++];
        if (!observe(m_->first) || !observe(m_->second)) {
            fail = "observed value differs from the narg string";
            return false;
        }
        return true;
    };
    auto check_pow = [&](const std::vector<uint64_t>& nonces, size_t* idx, uint32_t bits) -> bool {
        it.public_message(std::string("pcs/flock/pow/v1"));
        it.public_message(&bits, 4);
        uint8_t seed[16];
        it.t.challenge(seed, 16);
        if (*idx >= nonces.size()) {
            fail = "missing nonce";
            return false;
        }
        const uint64_t nonce = nonces[(*idx)++];
        const uint8_t* r = it.read(8);
        uint64_t rn = 0;
        if (r) memcpy(&rn, r, 8);
        if (!r || rn != nonce || !pow_valid(kPcsPow, seed, nonce, bits)) {
            fail = "proof of work";
            return false;
        }
        return true;
    };
    auto sample_vec = [&](size_t n) {
        it.public_message(std::string("pcs/flock/sample-vector/v1"));
        uint64_t nn = n;
        it.public_message(&nn, 8);
        std::vector<G> v(n);
        for (auto& x : v) x = it.gf();
        return v;
    };
    auto sample_queries = [&](uint64_t block_len, int count) {
        std::set<uint64_t> seen;
        std::vector<uint64_t> out;
        while ((int)out.size() < count) {
            const uint64_t qq = it.gf().lo % block_len;
            if (seen.insert(qq).second) out.push_back(qq);
        }
        std::sort(out.begin(), out.end());
        return out;
    };
    auto check_root_bytes = [&](const uint8_t* expected) -> bool {
        const uint8_t* b = it.read(36);
        uint32_t l32 = 0;
        if (b) memcpy(&l32, b, 4);
        return b && l32 == 32 && memcmp(b + 4, expected, 32) == 0;
    };
    auto verify_open = [&](const OpenData& o, const uint8_t* root_, uint64_t block_len, uint64_t lanes,
                           const std::vector<uint64_t>& queries) -> bool {
        if (o.rows.size() != queries.size()) return false;
        std::vector<std::array<uint8_t, 32>> leaves(o.rows.size());
        std::vector<uint8_t> bytes;
        for (size_t i = 0; i < o.rows.size(); i++) {
            if (o.rows[i].size() != lanes) return false;
            bytes.resize(16 * lanes);
            for (uint64_t l = 0; l < lanes; l++) gput(o.rows[i][l], &bytes[16 * l]);
            leaf_cv(bytes.data(), bytes.size(), leaves[i].data());
        }
        return merkle_verify_multi(root_, block_len, queries, leaves, o.proof);
    };
    auto enforced_sum = [&](const OpenData& o, const std::vector<G>& v_ch, const std::vector<uint64_t>& queries,
                            const std::vector<G>& alpha) {
        std::vector<G> aw = host_eq(alpha);
        const std::vector<G> eqv = host_eq(v_ch);
        G tot = gzero();
        for (size_t i = 0; i < queries.size() && i < aw.size(); i++) {
            G d = gzero();
            for (size_t x = 0; x < o.rows[i].size() && x < eqv.size(); x++) d = hgf_add(d, hgf_mul(o.rows[i][x], eqv[x]));
            tot = hgf_add(tot, hgf_mul(d, aw[i]));
        }
        return tot;
    };
    // basis components introduced by the recursion, each folded by the
    // challenges after its introduction (index `start` into ris) and weighted
    // by its batching challenge: the induced bases of the query sets and the
    // eq bases of the OOD samples (pcs/ligerito.rs:4362-4412)
    struct InducedComp {
        int log_n;
        std::vector<uint64_t> queries;
        std::vector<G> aw;
        size_t start;
        G w;
    };
    struct EqComp {
        std::vector<G> z;
        size_t start;
        G w;
    };
    std::vector<InducedComp> induced_comps;
    std::vector<EqComp> eq_comps;
    auto induce_comp = [&](int log_n, const std::vector<uint64_t>& queries, const std::vector<G>& alpha, size_t start,
                           G w) -> int {
        std::vector<G> aw = host_eq(alpha);
        aw.resize(queries.size());
        induced_comps.push_back(InducedComp{log_n, queries, std::move(aw), start, w});
        return BITZ_OK;
    };
    it.public_message(std::string("flock-ligerito-basis-v0"));
    V_CHECK(observe(packed_target), "observed value differs from the narg string");
    V_CHECK(check_root_bytes(root), "root message");
    G t_r = packed_target;
    std::pair<G, G> msg;
    V_CHECK(next_msg(&msg), fail);
    Quad rq = quad(msg, t_r);
    std::vector<G> ris;
    for (int j = 0; j < k0; j++) {
        const int bits = std::max(0, lv[0].fold_grinding_bits - j);
        if (bits > 0) V_CHECK(check_pow(lp.fold_nonces, &fni, (uint32_t)bits), fail);
        const G r = it.gf();
        ris.push_back(r);
        t_r = qeval(rq, r);
        V_CHECK(next_msg(&msg), fail);
        rq = quad(msg, t_r);
    }
    V_CHECK(!lp.roots.empty(), "missing recursive roots");
    V_CHECK(check_root_bytes(lp.roots[0].data()), "root message");
    int n_cur = mp - k0;
    auto ood_mirror = [&](int level, int n_vars) -> int {
        for (int o = 0; o < lv[level].ood_samples; o++) {
            const std::vector<G> z = sample_vec((size_t)n_vars);
            V_CHECK(oi < lp.ood.size(), "missing OOD value");
            const G y = lp.ood[oi++];
            V_CHECK(observe(y), "observed value differs from the narg string");
            V_CHECK(next_msg(&msg), fail);
            const Quad iq = quad(msg, y);
            const G beta = it.gf();
            rq = qfold(rq, iq, beta);
            t_r = hgf_add(t_r, hgf_mul(beta, y));
            eq_comps.push_back(EqComp{z, ris.size(), beta});
        }
        return BITZ_OK;
    };
    if ((rc = ood_mirror(1, n_cur))) return rc;
    V_CHECK(check_pow(lp.grind, &gni, (uint32_t)lv[0].grinding_bits), fail);
    const uint64_t block0 = 1ull << (lv[0].log_msg_cols + lv[0].log_inv_rate);
    const std::vector<uint64_t> q0 = sample_queries(block0, lv[0].queries);
    const std::vector<G> a0 = sample_vec((size_t)ceil_log2((uint64_t)lv[0].queries));
    V_CHECK(verify_open(lp.initial, root, block0, 1ull << k0, q0), "Merkle multi-proof");
    {
        const G es = enforced_sum(lp.initial, ris, q0, a0);
        V_CHECK(next_msg(&msg), fail);
        const Quad iq = quad(msg, es);
        const G beta = it.gf();
        rq = qfold(rq, iq, beta);
        t_r = hgf_add(t_r, hgf_mul(beta, es));
        if ((rc = induce_comp(n_cur, q0, a0, ris.size(), beta))) return rc;
    }
    const uint8_t* prev_root = lp.roots[0].data();
    int prev = 1;
    const int r_steps = (int)lv.size() - 1;
    size_t next_root = 1, reci = 0;
    for (int i = 0; i < r_steps; i++) {
        std::vector<G> level_rs;
        for (int j = 0; j < lv[i + 1].k_recursive; j++) {
            const int bits = std::max(0, lv[i + 1].fold_grinding_bits - j);
            if (bits > 0) V_CHECK(check_pow(lp.fold_nonces, &fni, (uint32_t)bits), fail);
            const G r = it.gf();
            ris.push_back(r);
            level_rs.push_back(r);
            t_r = qeval(rq, r);
            V_CHECK(next_msg(&msg), fail);
            rq = quad(msg, t_r);
        }
        n_cur -= lv[i + 1].k_recursive;
        const LigLevel& Lp = lv[prev];
        const uint64_t block_prev = 1ull << (Lp.log_msg_cols + Lp.log_inv_rate);
        if (i == r_steps - 1) {
            V_CHECK(mi == lp.msgs.size() && oi == lp.ood.size() && fni == lp.fold_nonces.size(),
                    "unconsumed Ligerito transcript");
            V_CHECK(n_cur >= 0 && n_cur < 32 && lp.yr.size() == (1ull << n_cur), "yr length");
            for (const G v : lp.yr) V_CHECK(observe(v), "observed value differs from the narg string");
            V_CHECK(check_pow(lp.grind, &gni, (uint32_t)lv[i + 1].grinding_bits), fail);
            const std::vector<uint64_t> ql = sample_queries(block_prev, lv[i + 1].queries);
            const std::vector<G> al = sample_vec((size_t)ceil_log2((uint64_t)lv[i + 1].queries));
            V_CHECK(verify_open(lp.fin, prev_root, block_prev, 1ull << Lp.log_num_interleaved, ql),
                    "Merkle multi-proof");
            const G esl = enforced_sum(lp.fin, level_rs, ql, al);
            const G beta_l = it.gf();
            t_r = hgf_add(t_r, hgf_mul(beta_l, esl));
            if ((rc = induce_comp(n_cur, ql, al, ris.size(), beta_l))) return rc;
            // succinct residual: every component at (its folding challenges, y)
            // for the binary points y of the last n_cur coordinates
            using residual::E;
            const uint64_t ny = lp.yr.size();
            V_CHECK((int)ris.size() + n_cur == mp, "residual dimension");
            std::vector<G> combined(ny), part(ny);
            residual::ring_switch((const E*)suffix_pt.data(), mp, (const E*)ris.data(), (int)ris.size(),
                                  (const E*)bw.data(), (E*)combined.data());
            residual::eq_point((const E*)opt.data(), mp, (const E*)ris.data(), n_cur, E{eta.lo, eta.hi},
                               (E*)part.data());
            for (uint64_t y = 0; y < ny; y++) combined[y] = hgf_add(combined[y], part[y]);
            for (const EqComp& c : eq_comps) {
                V_CHECK(c.start + c.z.size() == ris.size() + n_cur, "residual dimension");
                residual::eq_point((const E*)c.z.data(), (int)c.z.size(), (const E*)ris.data() + c.start, n_cur,
                                   E{c.w.lo, c.w.hi}, (E*)part.data());
                for (uint64_t y = 0; y < ny; y++) combined[y] = hgf_add(combined[y], part[y]);
            }
            for (const InducedComp& c : induced_comps) {
                V_CHECK(c.start + c.log_n == ris.size() + n_cur, "residual dimension");
                std::vector<G> sks, inv;
                sk_at_vks(c.log_n, &sks, &inv);
                residual::induced(c.log_n, c.queries.data(), (int)c.queries.size(), (const E*)c.aw.data(),
                                  (const E*)sks.data(), (const E*)inv.data(), (const E*)ris.data() + c.start, n_cur,
                                  (E*)part.data());
                for (uint64_t y = 0; y < ny; y++) combined[y] = hgf_add(combined[y], hgf_mul(c.w, part[y]));
            }
            G inner = gzero();
            for (size_t y = 0; y < combined.size(); y++) inner = hgf_add(inner, hgf_mul(lp.yr[y], combined[y]));
            V_CHECK(geq(inner, t_r), "Ligerito final check");
            V_CHECK(gni == lp.grind.size(), "unconsumed grinding nonces");
            // every recursive root and opening is consumed (the flock
            // verifier rejects recursive_roots.len() != r, ligerito.rs:5298)
            V_CHECK(next_root == lp.roots.size() && reci == lp.rec.size(),
                    "unconsumed recursive roots or openings");
            V_CHECK(it.pos == it.len, "trailing narg bytes");
            return BITZ_OK;
        }
        V_CHECK(next_root < lp.roots.size(), "missing recursive root");
        const uint8_t* root_n = lp.roots[next_root++].data();
        V_CHECK(check_root_bytes(root_n), "root message");
        if ((rc = ood_mirror(i + 2, n_cur))) return rc;
        V_CHECK(check_pow(lp.grind, &gni, (uint32_t)lv[i + 1].grinding_bits), fail);
        const std::vector<uint64_t> qi = sample_queries(block_prev, lv[i + 1].queries);
        const std::vector<G> ai = sample_vec((size_t)ceil_log2((uint64_t)lv[i + 1].queries));
        V_CHECK(reci < lp.rec.size(), "missing recursive opening");
        const OpenData& oi_ = lp.rec[reci++];
        V_CHECK(verify_open(oi_, prev_root, block_prev, 1ull << Lp.log_num_interleaved, qi), "Merkle multi-proof");
        const G es = enforced_sum(oi_, level_rs, qi, ai);
        V_CHECK(next_msg(&msg), fail);
        const Quad iq = quad(msg, es);
        const G beta = it.gf();
        rq = qfold(rq, iq, beta);
        t_r = hgf_add(t_r, hgf_mul(beta, es));
        if ((rc = induce_comp(n_cur, qi, ai, ris.size(), beta))) return rc;
        prev_root = root_n;
        prev = i + 2;
    }
    V_CHECK(false, "Ligerito ladder without levels");
}

int make_prover(bitz_ctx* ctx, const bitz_commit_handle* h, int target_bits, Prover* p) {
    p->ctx = ctx;
    p->h = h;
    p->t = (int)h->t;
    p->s = (int)h->s;
    p->m = p->t + p->s;
    p->mp = p->m - 7;
    std::string err = custom_johnson_config(p->m, (int)h->log_inv_rate, (int)h->log_batch, target_bits, &p->cfg);
    BITZ_REQUIRE(err.empty(), BITZ_ERR_SHAPE, "Ligerito configuration: " + err);
    BITZ_REQUIRE(p->t >= 7 && p->s >= 1, BITZ_ERR_SHAPE, "standalone opening needs t >= 7 and s >= 1");
    BITZ_REQUIRE(p->cfg.levels.size() >= 2, BITZ_ERR_SHAPE, "Ligerito ladder needs at least two levels");
    static G gen = smallest_generator();
    p->g = gen;
    return BITZ_OK;
}

}  // namespace

// ---- the BitZ opening of a linear claim, for composed proofs ----
hgf bitz_generator() {
    static const G gen = smallest_generator();
    return gen;
}

int round0_prove(bitz_ctx* ctx, const bitz_commit_handle* h, Transcript& outer, Round0Claim* r0) {
    Prover p;
    int rc = make_prover(ctx, h, 100, &p);
    if (rc) return rc;
    const uint32_t bits = (uint32_t)p.cfg.round0_grinding_bits;
    ood_round_header(outer, p.mp, bits);
    r0->has_nonce = bits != 0;
    r0->nonce = 0;
    if (bits) {
        uint8_t seed[32];
        grinding_seed(outer, "bitz/core/ood-round-grinding/v1", 0, bits, seed);
        if ((rc = pow_find(ctx, seed, seed + 16, bits, &r0->nonce))) return rc;
        absorb_grinding_nonce(outer, r0->nonce);
    }
    uint8_t zb[16];
    outer.challenge(zb, 16);
    r0->point.resize(p.mp);
    G cur = gfrom(zb);
    for (int i = 0; i < p.mp; i++) {
        r0->point[i] = cur;
        cur = hgf_mul(cur, cur);
    }
    if ((rc = p.mle_packed(r0->point, &r0->y))) return rc;
    uint8_t tagged[17];
    tagged[0] = 0x50;
    gput(r0->y, tagged + 1);
    outer.absorb_slice(tagged, 17);
    return BITZ_OK;
}

int open_linear_claim(bitz_ctx* ctx, const bitz_commit_handle* h, const char* session, const uint8_t tag[32],
                      const uint8_t* lead, size_t lead_len, u128 q, const void* roww, const void* colw_in,
                      u128 claimed, const Round0Claim& r0, std::vector<uint8_t>* proof) {
    Prover p;
    int rc = make_prover(ctx, h, 100, &p);
    if (rc) return rc;
    BITZ_REQUIRE(claimed < q, BITZ_ERR_ARG, "BitZ claim: value not reduced");
    const uint64_t nrows = 1ull << p.t, ncols = 1ull << p.s;
    DevBuf colw;
    if (!p.roww.alloc(nrows * 16) || !colw.alloc(ncols * 16) || !p.folds.alloc(ncols * 16)) return BITZ_ERR_ALLOC;
    BITZ_HIP(hipMemcpyAsync(p.roww.ptr, roww, nrows * 16, hipMemcpyDeviceToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(colw.ptr, colw_in, ncols * 16, hipMemcpyDeviceToDevice, ctx->stream));
    std::vector<uint8_t> fold_bytes(ncols * 16);
    if ((rc = fold_columns(ctx, (const uint64_t*)h->packed.ptr, p.t, p.s, p.roww.ptr, p.folds.ptr))) return rc;
    if ((rc = p.download(p.folds.ptr, fold_bytes.data(), fold_bytes.size()))) return rc;
    InnerTranscript it(session, tag, 32);
    if (lead_len) it.public_message(lead, lead_len);
    PhaseClock phase(ctx);
    return p.open_claim(it, phase, q, colw, claimed, fold_bytes, r0.point, r0.y, r0.has_nonce, r0.nonce, proof);
}

int open_virtual_claim(bitz_ctx* ctx, const bitz_commit_handle* h, const char* session, const uint8_t tag[32],
                       const std::vector<std::vector<uint8_t>>& lead, const void* grid, int th, int sh, u128 q,
                       const void* roww, const void* colw_in, u128 claimed, TransposeFn transpose, void* user,
                       const Round0Claim& r0, std::vector<uint8_t>* proof) {
    Prover p;
    int rc = make_prover(ctx, h, 100, &p);
    if (rc) return rc;
    BITZ_REQUIRE(claimed < q, BITZ_ERR_ARG, "BitZ claim: value not reduced");
    BITZ_REQUIRE(th >= 7 && sh >= 1 && th + sh <= 35, BITZ_ERR_SHAPE, "virtual opening: derived grid shape");
    // the fold and the GKR run over the derived grid
    p.t = th;
    p.s = sh;
    const uint64_t nrows = 1ull << th, ncols = 1ull << sh;
    DevBuf colw;
    if (!p.roww.alloc(nrows * 16) || !colw.alloc(ncols * 16) || !p.folds.alloc(ncols * 16)) return BITZ_ERR_ALLOC;
    BITZ_HIP(hipMemcpyAsync(p.roww.ptr, roww, nrows * 16, hipMemcpyDeviceToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(colw.ptr, colw_in, ncols * 16, hipMemcpyDeviceToDevice, ctx->stream));
    std::vector<uint8_t> fold_bytes(ncols * 16);
    if ((rc = fold_columns(ctx, (const uint64_t*)grid, th, sh, p.roww.ptr, p.folds.ptr))) return rc;
    if ((rc = p.download(p.folds.ptr, fold_bytes.data(), fold_bytes.size()))) return rc;
    InnerTranscript it(session, tag, 32);
    for (auto& msg : lead) it.public_message(msg.data(), msg.size());
    {
        uint8_t b[16];
        u128_put(claimed, b);
        if ((rc = p.public_weights(it, p.roww.ptr, nrows, colw.ptr, ncols, gfrom(b)))) return rc;
    }
    PhaseClock phase(ctx);
    phase.mark("fold_round");
    DevBuf u1, u2, wf, one;
    G target;
    if ((rc = p.reduce_claim(it, phase, grid, fold_bytes, u1, u2, &target))) return rc;
    phase.mark("transpose");
    if (!wf.alloc((16ull << p.m))) return BITZ_ERR_ALLOC;
    if ((rc = transpose(user, u1.ptr, u2.ptr, wf.ptr))) return rc;
    if ((rc = p.upload_gf(std::vector<G>{gone()}, one))) return rc;
    phase.mark("inner_sumcheck");
    return p.prove_lin(it, phase, wf, one, p.m, 0, target, false, r0.point, r0.y, r0.has_nonce, r0.nonce, proof);
}

int verify_linear_claim(bitz_ctx* ctx, const uint8_t root[32], int t, int s, int r0, int k0, const char* session,
                        const uint8_t tag[32], const uint8_t* lead, size_t lead_len, u128 q, const void* roww,
                        const void* colw, u128 claimed, const std::vector<G>& opt, G ood_y, const uint8_t* narg,
                        uint64_t narg_len, const uint8_t* hints, uint64_t hints_len) {
    LigConfig cfg;
    std::string err = custom_johnson_config(t + s, r0, k0, 100, &cfg);
    BITZ_REQUIRE(err.empty(), BITZ_ERR_SHAPE, "Ligerito configuration: " + err);
    BITZ_REQUIRE(t >= 7 && s >= 1 && cfg.levels.size() >= 2 && t + s - 7 <= kMaxTensorN, BITZ_ERR_SHAPE,
                 "verify: unsupported shape");
    if (claimed >= q) {
        set_error("verify: claim not reduced");
        return BITZ_ERR_VERIFY;
    }
    const uint64_t nrows = 1ull << t, ncols = 1ull << s;
    DevBuf rw, cw;
    if (!rw.alloc(nrows * 16) || !cw.alloc(ncols * 16)) return BITZ_ERR_ALLOC;
    BITZ_HIP(hipMemcpyAsync(rw.ptr, roww, nrows * 16, hipMemcpyDeviceToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(cw.ptr, colw, ncols * 16, hipMemcpyDeviceToDevice, ctx->stream));
    InnerReader it(session, tag, 32, narg, narg_len);
    if (lead_len) it.public_message(lead, lead_len);
    return verify_claim_impl(ctx, it, root, t, s, cfg, q, rw, cw, claimed, opt, ood_y, narg_len, hints, hints_len);
}

int verify_virtual_claim(bitz_ctx* ctx, const uint8_t root[32], int tf, int sf, const char* session,
                         const uint8_t tag[32], const std::vector<std::vector<uint8_t>>& lead, int th, int sh, u128 q,
                         const void* roww, const void* colw, u128 claimed, TransposeFn transpose, void* user,
                         const std::vector<G>& opt, G ood_y, const uint8_t* narg, uint64_t narg_len,
                         const uint8_t* hints, uint64_t hints_len) {
    LigConfig cfg;
    std::string err = custom_johnson_config(tf + sf, 1, 4, 100, &cfg);
    BITZ_REQUIRE(err.empty(), BITZ_ERR_SHAPE, "Ligerito configuration: " + err);
    BITZ_REQUIRE(th >= 7 && sh >= 1 && th + sh <= 35 && cfg.levels.size() >= 2 && tf + sf - 7 <= kMaxTensorN,
                 BITZ_ERR_SHAPE, "verify: unsupported shape");
    if (claimed >= q) {
        set_error("verify: claim not reduced");
        return BITZ_ERR_VERIFY;
    }
    const uint64_t nrows = 1ull << th, ncols = 1ull << sh;
    DevBuf rw, cw, u1, u2, wf, one;
    if (!rw.alloc(nrows * 16) || !cw.alloc(ncols * 16)) return BITZ_ERR_ALLOC;
    BITZ_HIP(hipMemcpyAsync(rw.ptr, roww, nrows * 16, hipMemcpyDeviceToDevice, ctx->stream));
    BITZ_HIP(hipMemcpyAsync(cw.ptr, colw, ncols * 16, hipMemcpyDeviceToDevice, ctx->stream));
    InnerReader it(session, tag, 32, narg, narg_len);
    for (auto& msg : lead) it.public_message(msg.data(), msg.size());
    int rc;
    {
        uint8_t b[16];
        u128_put(claimed, b);
        if ((rc = public_weights_dev(ctx, it, rw.ptr, nrows, cw.ptr, ncols, gfrom(b)))) return rc;
    }
    G target;
    if ((rc = verify_reduce_impl(ctx, it, th, sh, q, rw, cw, claimed, u1, u2, &target))) return rc;
    if (!wf.alloc(16ull << cfg.m) || !one.alloc(16)) return BITZ_ERR_ALLOC;
    if ((rc = transpose(user, u1.ptr, u2.ptr, wf.ptr))) return rc;
    {
        const G o = gone();
        BITZ_HIP(hipMemcpyAsync(one.ptr, &o, 16, hipMemcpyHostToDevice, ctx->stream));
        BITZ_HIP(hipStreamSynchronize(ctx->stream));
    }
    return verify_lin_impl(ctx, it, root, cfg.m, 0, cfg, wf, one, target, false, opt, ood_y, narg_len, hints,
                           hints_len);
}

bool round0_pow_valid(const uint8_t seed[32], uint64_t nonce, uint32_t bits) {
    return pow_valid(seed, seed + 16, nonce, bits);
}

}  // namespace bitz

extern "C" int bitz_standalone_evaluation(bitz_ctx* ctx, bitz_commit_handle* h, uint8_t claimed_out[16]) {
    BITZ_REQUIRE(ctx && h && claimed_out, BITZ_ERR_ARG, "bitz_standalone_evaluation: null argument");
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    bitz::Prover p;
    int rc = bitz::make_prover(ctx, h, 100, &p);
    if (rc) return rc;
    return p.run(nullptr, claimed_out, nullptr);
}

extern "C" int bitz_prove_standalone(bitz_ctx* ctx, bitz_commit_handle* h, const uint8_t claimed[16],
                                     uint8_t* proof, uint64_t capacity, uint64_t* proof_len) {
    BITZ_REQUIRE(ctx && h && claimed && proof_len, BITZ_ERR_ARG, "bitz_prove_standalone: null argument");
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    bitz::Prover p;
    int rc = bitz::make_prover(ctx, h, 100, &p);
    if (rc) return rc;
    std::vector<uint8_t> bytes;
    rc = p.run(claimed, nullptr, &bytes);
    if (rc) return rc;
    *proof_len = bytes.size();
    BITZ_REQUIRE(proof && capacity >= bytes.size(), BITZ_ERR_ARG, "bitz_prove_standalone: proof buffer too small");
    memcpy(proof, bytes.data(), bytes.size());
    return BITZ_OK;
}

extern "C" int bitz_verify_standalone(bitz_ctx* ctx, const uint8_t root[32], uint32_t t, uint32_t s,
                                      uint32_t log_inv_rate, uint32_t initial_k, const uint8_t claimed[16],
                                      const uint8_t* proof, uint64_t proof_len) {
    BITZ_REQUIRE(ctx && root && claimed && (proof || proof_len == 0), BITZ_ERR_ARG,
                 "bitz_verify_standalone: null argument");
    BITZ_REQUIRE(t <= 40 && s <= 40, BITZ_ERR_SHAPE, "bitz_verify_standalone: shape out of range");
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    bitz::u128 c = 0;
    for (int i = 15; i >= 0; i--) c = (c << 8) | claimed[i];
    return bitz::verify_standalone_impl(ctx, root, (int)t, (int)s, (int)log_inv_rate, (int)initial_k, c, proof,
                                        (size_t)proof_len);
}

extern "C" int bitz_ligerito_config(uint32_t m, uint32_t log_inv_rate, uint32_t initial_k, uint32_t target_bits,
                                    int32_t* out, uint32_t capacity) {
    BITZ_REQUIRE(out, BITZ_ERR_ARG, "bitz_ligerito_config: null out");
    bitz::LigConfig cfg;
    std::string err = bitz::custom_johnson_config((int)m, (int)log_inv_rate, (int)initial_k, (int)target_bits, &cfg);
    BITZ_REQUIRE(err.empty(), BITZ_ERR_SHAPE, "Ligerito configuration: " + err);
    std::vector<int32_t> v = {cfg.m, cfg.log_n, cfg.initial_k, cfg.yr_log_n, cfg.round0_grinding_bits,
                              (int32_t)cfg.levels.size()};
    for (auto& lv : cfg.levels)
        for (int x : {lv.log_inv_rate, lv.log_msg_cols, lv.log_num_interleaved, lv.k_recursive, lv.queries,
                      lv.grinding_bits, lv.fold_grinding_bits, lv.ood_samples})
            v.push_back(x);
    BITZ_REQUIRE(capacity >= v.size(), BITZ_ERR_ARG, "bitz_ligerito_config: capacity too small");
    memcpy(out, v.data(), v.size() * 4);
    return (int)v.size();
}

extern "C" void bitz_host_blake3(const uint8_t* data, uint64_t len, uint8_t out[32]) {
    bitz::blake3_hash_host(data, len, out);
}

extern "C" void bitz_host_blake3_parts(const uint8_t* data, uint64_t len, uint64_t split, uint8_t out[32]) {
    bitz::Blake3 h;
    if (split > len) split = len;
    h.update(data, split);
    h.update(data + split, len - split);
    h.finalize_xof(out, 32);
}
