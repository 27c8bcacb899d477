// Shared host-side plumbing of libbitz_hip: context, error reporting,
// checked HIP calls, and host GF(2^128) arithmetic for small setup tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "host_gf.h"

#include "../../include/bitz_hip.h"

namespace bitz {

void set_error(const std::string& msg);

// Caching device allocator of one context. Every launch of a context goes to
// its one stream, so a block released here and handed out again is only
// touched by later work of that stream: no synchronisation is needed. Large
// blocks are rounded up to 2 MiB so that the buffers of repeated proofs of
// one shape are served from the cache instead of hipMalloc/hipFree.
struct Pool {
    std::multimap<size_t, void*> free_blocks;
    size_t cached = 0;
    static size_t round_up(size_t n) {
        const size_t big = 2u << 20, small = 256;
        return n >= big / 2 ? (n + big - 1) / big * big : (n + small - 1) / small * small;
    }
    void* get(size_t n, size_t* got) {
        n = round_up(n);
        auto it = free_blocks.lower_bound(n);
        if (it != free_blocks.end() && it->first <= n + n / 4) {
            void* p = it->second;
            *got = it->first;
            cached -= it->first;
            free_blocks.erase(it);
            return p;
        }
        void* p = nullptr;
        if (hipMalloc(&p, n) != hipSuccess) {
            (void)hipGetLastError();
            trim();
            if (hipMalloc(&p, n) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
        }
        *got = n;
        return p;
    }
    void put(void* p, size_t n) {
        free_blocks.emplace(n, p);
        cached += n;
    }
    void trim() {
        for (auto& kv : free_blocks) (void)hipFree(kv.second);
        free_blocks.clear();
        cached = 0;
    }
    ~Pool() { trim(); }
};

// The pool that DevBuf allocations of the calling thread draw from (set by
// PoolScope at the library's entry points; none = plain hipMalloc).
Pool*& current_pool();
struct PoolScope {
    Pool* prev;
    explicit PoolScope(Pool* p) : prev(current_pool()) { current_pool() = p; }
    ~PoolScope() { current_pool() = prev; }
};

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    Pool* pool = nullptr;  // where ptr came from and goes back to
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes), pool(o.pool) {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            bytes = o.bytes;
            pool = o.pool;
            o.ptr = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) {
            if (pool)
                pool->put(ptr, bytes);
            else
                (void)hipFree(ptr);
        }
        ptr = nullptr;
        bytes = 0;
    }
    bool alloc(size_t n) {
        if (ptr && bytes >= n) return true;
        release();
        if (n == 0) return true;
        pool = current_pool();
        if (pool) {
            size_t got = 0;
            ptr = pool->get(n, &got);
            if (!ptr) return false;
            bytes = got;
            return true;
        }
        if (hipMalloc(&ptr, n) != hipSuccess) {
            ptr = nullptr;
            return false;
        }
        bytes = n;
        return true;
    }
};

}  // namespace bitz

namespace bitz {
// Per-context device scratch reused across calls (partial sums, small
// tables) plus a pinned host staging buffer for round results.
struct Scratch {
    DevBuf partials, result, eq_lo, fold_partial, comb, what, misc, leaf_tab;
    // the row-word rounds' per-row value tables and product / fold tables
    DevBuf nib_t, nib_tab;
    // row words built for a pattern-sums call that was not given them
    DevBuf nib_rows;
    // the proof-of-work search's best nonce: all ones between searches (each
    // search's final report resets it)
    DevBuf pow_best;
    bool pow_best_ready = false;
    void* host = nullptr;
    // Round results the host waits for: pinned, device-mapped, coherent host
    // memory ([0]: sequence flag, values from byte 64) written by the final
    // reduction kernel itself, so the host polls instead of copying + syncing.
    void* mbox = nullptr;
    void* mbox_dev = nullptr;
    // pinned, device-mapped, coherent host memory the host tails of the GKR
    // layers arrive in (written by a kernel, announced through the mailbox flag)
    void* tail = nullptr;
    void* tail_dev = nullptr;
    size_t tail_bytes = 0;
    DevBuf tail_done;  // the block count of the kernel that writes `tail`
    uint32_t seq = 0;
    bool mbox_pending = false;
    // pinned host staging for the host-side last rounds of a sumcheck
    void* stage = nullptr;
    size_t stage_bytes = 0;
    int ensure_stage(size_t bytes);
    // pinned upload ring: host data is copied in and sent with an
    // asynchronous copy, no stream synchronisation; a region is written again
    // only after a synchronisation at the wrap-around
    void* up = nullptr;
    size_t up_bytes = 0, up_off = 0;
    uint8_t comb_base[16] = {0};
    bool comb_ready = false;
    int ensure(size_t partial_bytes);
    ~Scratch();
};
}  // namespace bitz

namespace bitz {
// Optional per-kernel timing: HIP events around instrumented launches, with
// the launch's algorithmic work (bytes and integer ops) attached.
struct KernelStat {
    uint64_t launches = 0;
    double ms = 0, bytes = 0, ops = 0;
};
// Launch-shape choices that do not change any result. The defaults are what
// bitz_ctx_autotune picks most often for the 2^24-coefficient shape (n = 30)
// on one MI355X; bitz_ctx_autotune replaces them by choices timed for the
// caller's shape.
struct Tune {
    uint32_t round_r_max = 32;        // slots per thread in round-sum kernels
    uint32_t round_min_chunks = 1024; // chunks below which R halves
    uint32_t round_block_cap = 2048;  // blocks per round-sum launch
    int host_tail_log = 11;           // sumcheck rounds of at most 2^x slots run on the host
    int ntt_radix_log = 4;            // butterfly layers per global NTT pass
    bool tuned = false;
};

struct PendingTiming {
    std::string name;
    hipEvent_t start, stop;
    double bytes, ops;
};
}  // namespace bitz

struct bitz_ctx {
    // first member: destroyed last, after every buffer that returns to it
    bitz::Pool pool;
    int device = 0;
    bitz::Scratch scratch;
    bool timing = false;
    std::string timing_only;  // non-empty: time only launches of this name
    std::vector<bitz::PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, bitz::KernelStat> stats;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // device twiddle tables per NTT domain dimension ell:
    // tw[(1 << layer) - 1 + block] = twiddle(layer, block)
    std::map<int, std::unique_ptr<bitz::DevBuf>> twiddles;
    // device buffers handed out through bitz_dev_alloc, by address
    std::map<void*, bitz::DevBuf> user_bufs;
    // set while the GKR rounds reduce their block partials on the device:
    // sums launches then leave the partials in scratch and record the count
    bool defer_sums = false;
    unsigned deferred_nb = 0;
    // GKR rounds draw their challenges from a device copy of the transcript
    // (off by default: on MI355X one wave hashes a round's ~8 BLAKE3
    // compressions in ~20 us, no faster than the host round trip)
    bool device_fs = false;
    bitz::Tune tune;  // launch shapes and thresholds (defaults or bitz_ctx_autotune)
    // bitz_commit_batch: side streams of the captured graphs' parallel
    // branches (default priority) and the instantiated graphs of recent
    // batches, keyed by their items (pointers and shapes)
    std::vector<hipStream_t> branch_streams;
    std::vector<std::pair<uint64_t, hipGraphExec_t>> batch_graphs;
};

// A committed matrix kept on the device: the bit rows (= packed message
// words), the L0 codeword and its Merkle tree.
struct bitz_commit_handle {
    bitz_ctx* ctx;
    bitz::DevBuf packed, codeword, tree;
    uint32_t t, s, log_batch, log_inv_rate;
    uint64_t n_leaves;
    uint8_t root[32];
};

namespace bitz {
// Integer work per primitive, for the op counts attached to timed launches:
// the VALU instructions of each primitive in this gfx950 build (one
// instruction = 64 lane operations of the wave), counted in the assembly of
// the parity kernels k_gf_mul / k_gf_mul_const: a variable x variable
// GF(2^128) product (~430 with its loads and stores), a product by a launch
// constant through its 4-bit window table (32 LDS lookups, ~125 VALU), a
// BLAKE3 compression (7 rounds x 8 G x 14 + 16 = 800).
constexpr double kGfMulOps = 430.0;
constexpr double kGfTabMulOps = 125.0;
constexpr double kBlake3Ops = 800.0;

// bitz_ctx_set_timing_only with this name records the host phases and times
// no launch (no kernel has this name)
constexpr const char* kPhasesOnly = "phases";

// RAII bracket: records start/stop events on ctx->stream when timing is on.
struct TimedLaunch {
    bitz_ctx* ctx;
    PendingTiming p;
    bool on;
    TimedLaunch(bitz_ctx* c, const char* name, double bytes, double ops);
    ~TimedLaunch();
};
}  // namespace bitz

namespace bitz {
// Host wall-clock phases of a call: mark(name) closes the running phase and
// opens the next; with timing on, each closed phase adds its milliseconds
// to ctx->stats under "phase:<name>" (host work plus any wait on the GPU).
struct PhaseClock {
    bitz_ctx* ctx;
    std::string cur;
    std::chrono::steady_clock::time_point t0;
    explicit PhaseClock(bitz_ctx* c) : ctx(c) {}
    void mark(const char* name) {
        close();
        cur = name;
        t0 = std::chrono::steady_clock::now();
    }
    void close();
    ~PhaseClock() { close(); }
};
// Accumulates the lifetime of the scope into "phase:<name>" (as PhaseClock),
// for sub-phases that recur inside one phase (GKR layer set-up, host tails,
// mailbox waits).
struct SubPhase {
    bitz_ctx* ctx;
    const char* name;
    std::chrono::steady_clock::time_point t0;
    SubPhase(bitz_ctx* c, const char* n) : ctx(c), name(n), t0(std::chrono::steady_clock::now()) {}
    ~SubPhase();
};
}  // namespace bitz

#define BITZ_HIP(expr)                                                                   \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            bitz::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));          \
            return BITZ_ERR_DEVICE;                                                      \
        }                                                                                \
    } while (0)

#define BITZ_REQUIRE(cond, code, msg)  \
    do {                               \
        if (!(cond)) {                 \
            bitz::set_error(msg);      \
            return (code);             \
        }                              \
    } while (0)

#define BITZ_LAUNCH_CHECK()                                                              \
    do {                                                                                 \
        hipError_t _e = hipGetLastError();                                               \
        if (_e != hipSuccess) {                                                          \
            bitz::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));     \
            return BITZ_ERR_DEVICE;                                                      \
        }                                                                                \
    } while (0)

namespace bitz {

// ---- host GF(2^128) values (round messages, challenges, small tables) ----
// Products go through host_gf.cpp: PCLMULQDQ where the CPU has it.
struct hgf {
    uint64_t lo, hi;
};
inline hgf hgf_add(hgf a, hgf b) { return hgf{a.lo ^ b.lo, a.hi ^ b.hi}; }
inline hgf hgf_mul(hgf a, hgf b) {
    const hostgf::E r = hostgf::mul(hostgf::E{a.lo, a.hi}, hostgf::E{b.lo, b.hi});
    return hgf{r.lo, r.hi};
}
inline hgf hgf_inv(hgf a) {
    hgf acc{1, 0}, sq = a;
    for (int i = 1; i < 128; i++) {
        sq = hgf_mul(sq, sq);
        acc = hgf_mul(acc, sq);
    }
    if (a.lo == 0 && a.hi == 0) acc = hgf{0, 0};
    return acc;
}

}  // namespace bitz
