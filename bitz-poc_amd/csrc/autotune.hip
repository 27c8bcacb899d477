// Autotuning of launch shapes and thresholds (bitz_ctx_autotune).
//
// None of the tuned choices changes a result, only how the work is cut:
//  * round-sum kernels: slots per thread (round_r_max), the chunk count below
//    which that halves (round_min_chunks), blocks per launch (round_block_cap),
//    chosen by timing the fused fold + row-only sums kernel, the opening's
//    dominant kernel, at the round sizes of a layer;
//  * host_tail_log: the largest round size whose host round (carry-less
//    multiply loops, host_gf.cpp) is not slower than a device round (fold,
//    sums, reduction, mailbox round trip), both timed here;
//  * ntt_radix_log: butterfly layers per global NTT pass, chosen by timing the
//    commit of the given shape.
// Every candidate runs on synthetic data of the real sizes; the choice is
// kept in the context (ctx->tune) for the jobs that follow.
#include <algorithm>
#include <chrono>
#include <vector>

#include "../../include/bitz_hip.h"
#include "host_gf.h"
#include "open_impl.h"

namespace bitz {
namespace {

// pseudo-random fill (SplitMix64 of the index): timing data must not be
// uniform, or table lookups would all hit one entry
__global__ void k_fill_random(uint64_t* __restrict__ p, uint64_t n, uint64_t seed) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        p[i] = z ^ (z >> 31);
    }
}

int fill_random(bitz_ctx* ctx, void* p, uint64_t words, uint64_t seed) {
    const unsigned blocks = (unsigned)std::min<uint64_t>(4096, (words + 255) / 256);
    hipLaunchKernelGGL(k_fill_random, dim3(std::max(1u, blocks)), dim3(256), 0, ctx->stream, (uint64_t*)p, words,
                       seed);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Milliseconds of `fn` (launches on ctx->stream), the best of `reps`.
template <class Fn>
int time_device(bitz_ctx* ctx, int reps, Fn fn, double* best) {
    hipEvent_t a, b;
    BITZ_HIP(hipEventCreate(&a));
    BITZ_HIP(hipEventCreate(&b));
    *best = 1e30;
    int rc = BITZ_OK;
    for (int i = 0; i < reps && rc == BITZ_OK; i++) {
        (void)hipEventRecord(a, ctx->stream);
        rc = fn();
        (void)hipEventRecord(b, ctx->stream);
        (void)hipEventSynchronize(b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        *best = std::min(*best, (double)ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

}  // namespace
}  // namespace bitz

extern "C" int bitz_ctx_set_tune(bitz_ctx* ctx, uint32_t round_r_max, uint32_t round_min_chunks,
                                 uint32_t round_block_cap, uint32_t host_tail_log, uint32_t ntt_radix_log) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_ctx_set_tune: null ctx");
    BITZ_REQUIRE(round_r_max >= 1 && round_r_max <= 32 && (round_r_max & (round_r_max - 1)) == 0, BITZ_ERR_ARG,
                 "bitz_ctx_set_tune: round_r_max must be a power of two in [1, 32]");
    BITZ_REQUIRE(round_min_chunks >= 1 && round_min_chunks <= (1u << 20), BITZ_ERR_ARG,
                 "bitz_ctx_set_tune: round_min_chunks out of range");
    BITZ_REQUIRE(round_block_cap >= 64 && round_block_cap <= 65536, BITZ_ERR_ARG,
                 "bitz_ctx_set_tune: round_block_cap must be in [64, 65536]");
    BITZ_REQUIRE(host_tail_log >= 4 && host_tail_log <= 14, BITZ_ERR_ARG,
                 "bitz_ctx_set_tune: host_tail_log must be in [4, 14]");
    BITZ_REQUIRE(ntt_radix_log >= 1 && ntt_radix_log <= 4, BITZ_ERR_ARG,
                 "bitz_ctx_set_tune: ntt_radix_log must be in [1, 4]");
    ctx->tune.round_r_max = round_r_max;
    ctx->tune.round_min_chunks = round_min_chunks;
    ctx->tune.round_block_cap = round_block_cap;
    ctx->tune.host_tail_log = (int)host_tail_log;
    ctx->tune.ntt_radix_log = (int)ntt_radix_log;
    ctx->tune.tuned = false;
    return BITZ_OK;
}

extern "C" int bitz_ctx_get_tune(bitz_ctx* ctx, uint32_t out[6]) {
    BITZ_REQUIRE(ctx && out, BITZ_ERR_ARG, "bitz_ctx_get_tune: null argument");
    out[0] = ctx->tune.round_r_max;
    out[1] = ctx->tune.round_min_chunks;
    out[2] = ctx->tune.round_block_cap;
    out[3] = (uint32_t)ctx->tune.host_tail_log;
    out[4] = (uint32_t)ctx->tune.ntt_radix_log;
    out[5] = ctx->tune.tuned ? 1u : 0u;
    return BITZ_OK;
}

extern "C" int bitz_ctx_autotune(bitz_ctx* ctx, uint32_t t, uint32_t s, uint32_t log_batch, uint32_t log_inv_rate,
                                 double* tune_ms) {
    using namespace bitz;
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_ctx_autotune: null ctx");
    const uint64_t fp = bitz_job_footprint(t, s, log_batch, log_inv_rate);
    BITZ_REQUIRE(fp != 0 && t <= 28 && s <= 20, BITZ_ERR_SHAPE, "bitz_ctx_autotune: shape rejected");
    BITZ_HIP(hipSetDevice(ctx->device));
    PoolScope pool_scope(&ctx->pool);
    const double t0 = now_ms();
    const int n = (int)(t + s);
    Tune best = ctx->tune;
    int rc;
    uint8_t rho[16];
    for (int i = 0; i < 16; i++) rho[i] = (uint8_t)(0x5A + 37 * i);

    // ---- 1. round-sum launch shape (fused fold + row-only sums) ----
    // sizes of a large layer's device rounds, every third one
    std::vector<int> logs;
    for (int k = std::min(n - 4, 24); k > 10; k -= 3) logs.push_back(k);
    if (!logs.empty()) {
        const uint64_t hmax = 1ull << logs.front();
        const int sy = std::max(0, logs.front() - (int)s);
        DevBuf l, r, eqy;
        if (!l.alloc(64 * hmax) || !r.alloc(64 * hmax) || !eqy.alloc(16ull << sy)) return BITZ_ERR_ALLOC;
        if ((rc = fill_random(ctx, l.ptr, 8 * hmax, 1))) return rc;
        if ((rc = fill_random(ctx, r.ptr, 8 * hmax, 2))) return rc;
        if ((rc = fill_random(ctx, eqy.ptr, 2ull << sy, 3))) return rc;
        double best_ms = 1e30;
        for (uint32_t rmax : {8u, 16u, 32u})
            for (uint32_t minch : {1024u, 2048u, 4096u})
                for (uint32_t cap : {2048u, 4096u, 8192u}) {
                    ctx->tune.round_r_max = rmax;
                    ctx->tune.round_min_chunks = minch;
                    ctx->tune.round_block_cap = cap;
                    double total = 0;
                    for (int k : logs) {
                        double ms;
                        rc = time_device(ctx, 3, [&] {
                            return gkr_fold_round_sums(ctx, l.ptr, r.ptr, 1ull << k, rho, nullptr, (int)s, eqy.ptr,
                                                       0, ctx->scratch.result.ptr);
                        }, &ms);
                        if (rc) {
                            ctx->tune = best;
                            return rc;
                        }
                        total += ms;
                    }
                    if (total < best_ms) {
                        best_ms = total;
                        best.round_r_max = rmax;
                        best.round_min_chunks = minch;
                        best.round_block_cap = cap;
                    }
                }
        ctx->tune = best;
    }

    // ---- 2. host tail threshold ----
    {
        const int xmax = std::min(13, n - 3);
        const uint64_t hh_max = 1ull << xmax;
        DevBuf l, r, eqc, eqy;
        if (!l.alloc(64 * hh_max) || !r.alloc(64 * hh_max) || !eqc.alloc(16 * hh_max) || !eqy.alloc(16))
            return BITZ_ERR_ALLOC;
        if ((rc = fill_random(ctx, l.ptr, 8 * hh_max, 4))) return rc;
        if ((rc = fill_random(ctx, r.ptr, 8 * hh_max, 5))) return rc;
        if ((rc = fill_random(ctx, eqc.ptr, 2 * hh_max, 6))) return rc;
        const uint64_t one[2] = {1, 0};
        BITZ_HIP(hipMemcpyAsync(eqy.ptr, one, 16, hipMemcpyHostToDevice, ctx->stream));
        if ((rc = ctx->scratch.ensure_stage(64 * hh_max))) return rc;
        BITZ_HIP(hipStreamSynchronize(ctx->stream));
        std::vector<hostgf::E> hl(4 * hh_max), hr(4 * hh_max), hw(2 * hh_max);
        for (size_t i = 0; i < hl.size(); i++) {
            hl[i] = hostgf::E{0x9E3779B97F4A7C15ull * (i + 1), 0xC2B2AE3D27D4EB4Full * (i + 7)};
            hr[i] = hostgf::E{0x165667B19E3779F9ull * (i + 3), 0x27D4EB2F165667C5ull * (i + 5)};
        }
        for (size_t i = 0; i < hw.size(); i++) hw[i] = hostgf::E{i * 0x9E37ull + 1, i};
        int tail = 0;
        for (int x = 6; x <= xmax; x++) {
            const uint64_t hh = 1ull << x;
            // device round below the row-only rounds: the previous round's
            // fold, this round's sums, the mailbox read of the two sums
            double dev = 1e30;
            for (int rep = 0; rep < 5; rep++) {
                const double a = now_ms();
                if ((rc = fold_pair(ctx, l.ptr, r.ptr, 2 * hh, rho))) return rc;
                if ((rc = gkr_round_sums(ctx, l.ptr, r.ptr, hh, eqc.ptr, x, eqy.ptr, 0, nullptr))) return rc;
                uint8_t sums[32];
                if ((rc = read_result(ctx, 2, sums))) return rc;
                dev = std::min(dev, now_ms() - a);
            }
            // host round of the same size
            double host = 1e30;
            for (int rep = 0; rep < 5; rep++) {
                const double a = now_ms();
                hostgf::E sums[2];
                hostgf::gkr_round_sums(hl.data(), hr.data(), hh, hw.data(), 0, sums);
                hostgf::fold(hl.data(), hh, hw[1]);
                hostgf::fold(hr.data(), hh, hw[1]);
                hostgf::drop_top(hw.data(), hh / 2);
                host = std::min(host, now_ms() - a);
            }
            // entering the tail one round earlier downloads 2 x 2hh more entries
            double copy = 1e30;
            for (int rep = 0; rep < 3; rep++) {
                const double a = now_ms();
                BITZ_HIP(hipMemcpyAsync(ctx->scratch.stage, l.ptr, 32 * hh, hipMemcpyDeviceToHost, ctx->stream));
                BITZ_HIP(hipStreamSynchronize(ctx->stream));
                copy = std::min(copy, now_ms() - a);
            }
            if (host + copy <= dev)
                tail = x;
            else
                break;
        }
        best.host_tail_log = std::max(tail, 6);
        ctx->tune = best;
    }

    // ---- 3. NTT radix: the commit of this shape ----
    {
        const uint64_t words = (1ull << s) << (t - 6);
        DevBuf rows;
        if (!rows.alloc(8 * words)) return BITZ_ERR_ALLOC;
        if ((rc = fill_random(ctx, rows.ptr, words, 8))) return rc;
        BITZ_HIP(hipStreamSynchronize(ctx->stream));
        double best_ms = 1e30;
        int best_radix = best.ntt_radix_log;
        for (int radix : {3, 4}) {
            ctx->tune.ntt_radix_log = radix;
            double ms = 1e30;
            for (int rep = 0; rep < 3; rep++) {
                uint8_t root[32];
                bitz_commit_handle* h = nullptr;
                const double a = now_ms();
                rc = bitz_commit_device_rows(ctx, rows.ptr, t, s, log_batch, log_inv_rate, root, &h);
                ms = std::min(ms, now_ms() - a);
                bitz_commit_free(h);
                if (rc) {
                    ctx->tune = best;
                    return rc;
                }
            }
            if (ms < best_ms) {
                best_ms = ms;
                best_radix = radix;
            }
        }
        best.ntt_radix_log = best_radix;
    }
    best.tuned = true;
    ctx->tune = best;
    if (tune_ms) *tune_ms = now_ms() - t0;
    return BITZ_OK;
}
