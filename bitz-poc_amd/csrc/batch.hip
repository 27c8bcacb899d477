// Batched commitments in one launch: the kernel sequences of many
// independent commits (Reed-Solomon encode + BLAKE3 Merkle tree, each of its
// own size and rate: commit_rs_flock_from_rows, src/ligerito_flock.rs:198-296
// -> pcs/commit.rs:180-315) captured into one HIP graph whose branches run
// side by side, then launched once. The instantiated graph is kept per batch
// (its items' pointers and shapes), so a batch of resident job slots that
// comes back replays it without capture or instantiation.
#include <string.h>

#include <algorithm>
#include <set>

#include "common.h"
#include "open_impl.h"

namespace bitz {
namespace {

constexpr int kBranches = 8;
constexpr size_t kKeptGraphs = 4;

uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h;
}

uint64_t batch_key(uint32_t n, const bitz_commit_item* it) {
    uint64_t h = mix(0, n);
    for (uint32_t i = 0; i < n; i++) {
        h = mix(h, (uint64_t)(uintptr_t)it[i].packed);
        h = mix(h, (uint64_t)(uintptr_t)it[i].codeword);
        h = mix(h, (uint64_t)(uintptr_t)it[i].tree);
        h = mix(h, ((uint64_t)it[i].t << 48) | ((uint64_t)it[i].s << 32) | ((uint64_t)it[i].log_batch << 16) |
                       it[i].log_inv_rate);
    }
    return h;
}

int check_item(const bitz_commit_item& it, int* log_msg, int* ell) {
    BITZ_REQUIRE(it.packed && it.codeword && it.tree, BITZ_ERR_ARG, "bitz_commit_batch: null buffer");
    BITZ_REQUIRE(it.t >= 7 && it.t <= 35 && it.s <= 35 - it.t, BITZ_ERR_SHAPE,
                 "bitz_commit_batch: shape needs 7 <= t and t + s <= 35");
    *log_msg = (int)(it.t + it.s) - 7;
    BITZ_REQUIRE(it.log_batch <= 6 && (int)it.log_batch < *log_msg && it.log_inv_rate >= 1 &&
                     it.log_inv_rate <= 8,
                 BITZ_ERR_SHAPE, "bitz_commit_batch: need log_batch <= 6, log_batch < t+s-7, 1 <= log_inv_rate <= 8");
    *ell = *log_msg - (int)it.log_batch + (int)it.log_inv_rate;
    return BITZ_OK;
}

// capture every item's commit kernels, item i on branch i mod kBranches;
// the capture runs on side streams (the context's stream may be the null
// stream, which cannot be captured), the graph is launched on the
// context's stream
int capture(bitz_ctx* ctx, uint32_t n, const bitz_commit_item* items, hipGraphExec_t* exec) {
    while (ctx->branch_streams.size() < kBranches) {
        hipStream_t st;
        BITZ_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        ctx->branch_streams.push_back(st);
    }
    const hipStream_t saved = ctx->stream;
    const hipStream_t origin = ctx->branch_streams[0];
    const int branches = (int)std::min<uint32_t>(n, kBranches);
    hipEvent_t fork, join[kBranches];
    BITZ_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (int b = 0; b < branches; b++) BITZ_HIP(hipEventCreateWithFlags(&join[b], hipEventDisableTiming));
    int rc = BITZ_OK;
    hipGraph_t graph = nullptr;
    BITZ_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal));
    if (hipEventRecord(fork, origin) != hipSuccess) rc = BITZ_ERR_DEVICE;
    for (int b = 1; b < branches && !rc; b++)
        if (hipStreamWaitEvent(ctx->branch_streams[b], fork, 0) != hipSuccess) rc = BITZ_ERR_DEVICE;
    const bool timing = ctx->timing;
    ctx->timing = false;  // no event records inside the graph
    for (uint32_t i = 0; i < n && !rc; i++) {
        const bitz_commit_item& it = items[i];
        ctx->stream = ctx->branch_streams[i % kBranches];
        rc = commit_device(ctx, it.packed, (int)(it.t + it.s) - 7, (int)it.log_batch, (int)it.log_inv_rate,
                           it.codeword, it.tree);
    }
    ctx->stream = saved;
    ctx->timing = timing;
    for (int b = 1; b < branches && !rc; b++) {
        if (hipEventRecord(join[b], ctx->branch_streams[b]) != hipSuccess ||
            hipStreamWaitEvent(origin, join[b], 0) != hipSuccess)
            rc = BITZ_ERR_DEVICE;
    }
    const hipError_t end = hipStreamEndCapture(origin, &graph);
    (void)hipEventDestroy(fork);
    for (int b = 0; b < branches; b++) (void)hipEventDestroy(join[b]);
    if (!rc && end != hipSuccess) {
        set_error(std::string("bitz_commit_batch: capture failed: ") + hipGetErrorString(end));
        rc = BITZ_ERR_DEVICE;
    }
    if (!rc) {
        const hipError_t e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            set_error(std::string("bitz_commit_batch: instantiate failed: ") + hipGetErrorString(e));
            rc = BITZ_ERR_DEVICE;
        }
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) (void)hipGetLastError();
    return rc;
}

}  // namespace
}  // namespace bitz

extern "C" int bitz_commit_batch(bitz_ctx* ctx, uint32_t n, bitz_commit_item* items) {
    BITZ_REQUIRE(ctx && (items || n == 0), BITZ_ERR_ARG, "bitz_commit_batch: null argument");
    if (n == 0) return BITZ_OK;
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    // shapes checked and every NTT domain's twiddles built before the capture
    std::set<int> ells;
    for (uint32_t i = 0; i < n; i++) {
        int log_msg, ell;
        if (int rc = bitz::check_item(items[i], &log_msg, &ell)) return rc;
        ells.insert(ell);
    }
    for (int ell : ells) {
        const uint4* tw;
        if (int rc = bitz::ensure_twiddles(ctx, ell, &tw)) return rc;
    }
    const uint64_t key = bitz::batch_key(n, items);
    hipGraphExec_t exec = nullptr;
    for (auto it = ctx->batch_graphs.begin(); it != ctx->batch_graphs.end(); ++it)
        if (it->first == key) {
            exec = it->second;
            break;
        }
    if (!exec) {
        if (int rc = bitz::capture(ctx, n, items, &exec)) return rc;
        if (ctx->batch_graphs.size() >= bitz::kKeptGraphs) {
            // the oldest graph's last launch is ordered before anything later
            // on this stream; its executable can go once that launch is done
            BITZ_HIP(hipStreamSynchronize(ctx->stream));
            (void)hipGraphExecDestroy(ctx->batch_graphs.front().second);
            ctx->batch_graphs.erase(ctx->batch_graphs.begin());
        }
        ctx->batch_graphs.emplace_back(key, exec);
    }
    BITZ_HIP(hipGraphLaunch(exec, ctx->stream));
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t leaves = 1ull << ((int)(items[i].t + items[i].s) - 7 - (int)items[i].log_batch +
                                          (int)items[i].log_inv_rate);
        BITZ_HIP(hipMemcpyAsync(items[i].root, (const uint8_t*)items[i].tree + (2 * leaves - 2) * 32, 32,
                                hipMemcpyDeviceToHost, ctx->stream));
    }
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    return BITZ_OK;
}
