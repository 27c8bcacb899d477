// Context, error reporting and the small arithmetic parity entry points.
#include "common.h"
#include "gf128_dev.h"
#include "blake3_dev.h"
#include "open_impl.h"

namespace bitz {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
Pool*& current_pool() {
    static thread_local Pool* pool = nullptr;
    return pool;
}
}  // namespace bitz

extern "C" const char* bitz_last_error(void) { return bitz::g_err.c_str(); }

namespace bitz {
static hipEvent_t take_event(bitz_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

TimedLaunch::TimedLaunch(bitz_ctx* c, const char* name, double bytes, double ops)
    : ctx(c), on(c->timing && (c->timing_only.empty() || c->timing_only == name)) {
    if (!on) return;
    p.name = name;
    p.bytes = bytes;
    p.ops = ops;
    p.start = take_event(ctx);
    p.stop = take_event(ctx);
    (void)hipEventRecord(p.start, ctx->stream);
}

TimedLaunch::~TimedLaunch() {
    if (!on) return;
    (void)hipEventRecord(p.stop, ctx->stream);
    ctx->pending.push_back(p);
}

void PhaseClock::close() {
    if (!cur.empty() && ctx->timing && (ctx->timing_only.empty() || ctx->timing_only == kPhasesOnly)) {
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        auto& st = ctx->stats["phase:" + cur];
        st.launches++;
        st.ms += ms;
    }
    cur.clear();
}

SubPhase::~SubPhase() {
    if (ctx->timing && (ctx->timing_only.empty() || ctx->timing_only == kPhasesOnly)) {
        auto& st = ctx->stats[std::string("phase:") + name];
        st.launches++;
        st.ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

static void resolve_timings(bitz_ctx* ctx) {
    for (auto& p : ctx->pending) {
        (void)hipEventSynchronize(p.stop);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.start, p.stop);
        auto& st = ctx->stats[p.name];
        st.launches++;
        st.ms += ms;
        st.bytes += p.bytes;
        st.ops += p.ops;
        ctx->event_pool.push_back(p.start);
        ctx->event_pool.push_back(p.stop);
    }
    ctx->pending.clear();
}
}  // namespace bitz

extern "C" int bitz_ctx_set_timing(bitz_ctx* ctx, int enable) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_ctx_set_timing: null ctx");
    bitz::resolve_timings(ctx);
    ctx->timing = enable != 0;
    ctx->timing_only.clear();
    if (enable) ctx->stats.clear();
    return BITZ_OK;
}

extern "C" int bitz_ctx_set_timing_only(bitz_ctx* ctx, const char* kernel) {
    BITZ_REQUIRE(ctx && kernel && kernel[0], BITZ_ERR_ARG, "bitz_ctx_set_timing_only: need ctx and a kernel name");
    bitz::resolve_timings(ctx);
    ctx->timing = true;
    ctx->timing_only = kernel;
    ctx->stats.clear();
    return BITZ_OK;
}

extern "C" int bitz_ctx_kernel_stats(bitz_ctx* ctx, char* names, uint64_t names_cap, double* values,
                                     uint64_t values_cap) {
    BITZ_REQUIRE(ctx && names && values, BITZ_ERR_ARG, "bitz_ctx_kernel_stats: null argument");
    bitz::resolve_timings(ctx);
    std::string all;
    uint64_t k = 0;
    for (auto& kv : ctx->stats) {
        if ((k + 1) * 4 > values_cap) break;
        all += kv.first + "\n";
        values[4 * k] = (double)kv.second.launches;
        values[4 * k + 1] = kv.second.ms;
        values[4 * k + 2] = kv.second.bytes;
        values[4 * k + 3] = kv.second.ops;
        k++;
    }
    BITZ_REQUIRE(all.size() + 1 <= names_cap, BITZ_ERR_ARG, "bitz_ctx_kernel_stats: names buffer too small");
    memcpy(names, all.c_str(), all.size() + 1);
    return (int)k;
}

extern "C" const char* bitz_version(void) { return "bitz_hip gfx950 0.1"; }

extern "C" int bitz_ctx_create(int device, bitz_ctx** out) {
    BITZ_REQUIRE(out != nullptr, BITZ_ERR_ARG, "bitz_ctx_create: null out");
    int count = 0;
    BITZ_HIP(hipGetDeviceCount(&count));
    BITZ_REQUIRE(device >= 0 && device < count, BITZ_ERR_ARG, "bitz_ctx_create: no such device");
    BITZ_HIP(hipSetDevice(device));
    // the no-allocation entry point's device array, resolved once per device
    if (int rc = bitz::noalloc_setup(device)) return rc;
    auto* ctx = new bitz_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        bitz::set_error("bitz_ctx_create: hipStreamCreate failed");
        return BITZ_ERR_DEVICE;
    }
    ctx->stream = ctx->own_stream;
    *out = ctx;
    return BITZ_OK;
}

extern "C" void bitz_ctx_destroy(bitz_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    // work of the context (copies out of its pinned buffers included) ends
    // before those buffers are freed
    if (ctx->stream && ctx->stream != ctx->own_stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->own_stream) {
        (void)hipStreamSynchronize(ctx->own_stream);
        (void)hipStreamDestroy(ctx->own_stream);
    }
    bitz::resolve_timings(ctx);
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto& g : ctx->batch_graphs) (void)hipGraphExecDestroy(g.second);
    for (auto st : ctx->branch_streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    delete ctx;
}

extern "C" int bitz_ctx_set_stream(bitz_ctx* ctx, void* stream) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_ctx_set_stream: null ctx");
    // blocks cached by the context's pool may still be in use by work on the
    // old stream: drain it before the new stream can be handed those blocks
    if ((hipStream_t)stream != ctx->stream) BITZ_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stream = (hipStream_t)stream;
    return BITZ_OK;
}

extern "C" int bitz_ctx_set_device_transcript(bitz_ctx* ctx, int enable) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_ctx_set_device_transcript: null ctx");
    ctx->device_fs = enable != 0;
    return BITZ_OK;
}

extern "C" int bitz_ctx_trim(bitz_ctx* ctx) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_ctx_trim: null ctx");
    BITZ_HIP(hipSetDevice(ctx->device));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    ctx->pool.trim();
    return BITZ_OK;
}

extern "C" uint64_t bitz_ctx_cached_bytes(bitz_ctx* ctx) { return ctx ? (uint64_t)ctx->pool.cached : 0; }

// ---- device buffers for callers without a HIP binding of their own ----

extern "C" int bitz_dev_alloc(bitz_ctx* ctx, uint64_t bytes, void** out) {
    BITZ_REQUIRE(ctx && out, BITZ_ERR_ARG, "bitz_dev_alloc: null argument");
    BITZ_HIP(hipSetDevice(ctx->device));
    bitz::PoolScope scope(&ctx->pool);
    bitz::DevBuf b;
    BITZ_REQUIRE(b.alloc(bytes ? bytes : 1), BITZ_ERR_ALLOC, "bitz_dev_alloc: device allocation failed");
    void* p = b.ptr;
    ctx->user_bufs.emplace(p, std::move(b));
    *out = p;
    return BITZ_OK;
}

extern "C" int bitz_dev_free(bitz_ctx* ctx, void* ptr) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_dev_free: null ctx");
    if (!ptr) return BITZ_OK;
    auto it = ctx->user_bufs.find(ptr);
    BITZ_REQUIRE(it != ctx->user_bufs.end(), BITZ_ERR_ARG, "bitz_dev_free: not a buffer of this context");
    ctx->user_bufs.erase(it);  // back to the pool, in stream order
    return BITZ_OK;
}

extern "C" int bitz_dev_upload(bitz_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    BITZ_REQUIRE(ctx && dst && (src || bytes == 0), BITZ_ERR_ARG, "bitz_dev_upload: null argument");
    if (bytes == 0) return BITZ_OK;
    BITZ_HIP(hipSetDevice(ctx->device));
    BITZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    return BITZ_OK;
}

extern "C" int bitz_dev_download(bitz_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    BITZ_REQUIRE(ctx && (dst || bytes == 0) && src, BITZ_ERR_ARG, "bitz_dev_download: null argument");
    if (bytes == 0) return BITZ_OK;
    BITZ_HIP(hipSetDevice(ctx->device));
    BITZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    return BITZ_OK;
}

extern "C" void* bitz_ctx_stream(bitz_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int bitz_ctx_synchronize(bitz_ctx* ctx) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_ctx_synchronize: null ctx");
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    return BITZ_OK;
}

// ---- parity helpers ----

__global__ void k_gf_mul(const uint4* __restrict__ a, const uint4* __restrict__ b,
                         uint4* __restrict__ out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gf_store(&out[i], gf_mul(gf_load(&a[i]), gf_load(&b[i])));
}

__global__ void k_gf_mul_const(const uint4* __restrict__ a, uint64_t c_lo, uint64_t c_hi,
                               uint4* __restrict__ out, uint64_t n) {
    __shared__ uint4 tab[512];
    gf_table_build(tab, gf_from_u64(c_lo, c_hi), threadIdx.x, blockDim.x);
    __syncthreads();
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gf_store(&out[i], gf_mul_tab(tab, gf_load(&a[i])));
}

// One thread per leaf: non-root chunk CV of `leaf_size` bytes (<= 1024).
__global__ void k_blake3_leaves(const uint32_t* __restrict__ data, uint64_t n, uint32_t leaf_words,
                                uint32_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* leaf = data + i * leaf_words;
    uint32_t cv[8];
    b3::iv_init(cv);
    uint32_t nblocks = leaf_words == 0 ? 1 : (leaf_words + 15) / 16;
    for (uint32_t b = 0; b < nblocks; b++) {
        uint32_t m[16];
        uint32_t avail = leaf_words - b * 16;
        for (int k = 0; k < 16; k++) m[k] = (uint32_t)k < avail ? leaf[b * 16 + k] : 0u;
        uint32_t flags = (b == 0 ? b3::CHUNK_START : 0u) | (b + 1 == nblocks ? b3::CHUNK_END : 0u);
        uint32_t blen = b + 1 == nblocks ? (avail >= 16 ? 64u : avail * 4u) : 64u;
        b3::compress(cv, m, 0, blen, flags);
    }
    for (int k = 0; k < 8; k++) out[i * 8 + k] = cv[k];
}

extern "C" int bitz_gf_mul_dev(bitz_ctx* ctx, const void* a, const void* b, void* out, uint64_t n) {
    BITZ_REQUIRE(ctx && a && b && out, BITZ_ERR_ARG, "bitz_gf_mul_dev: null argument");
    if (n == 0) return BITZ_OK;
    BITZ_HIP(hipSetDevice(ctx->device));
    uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_gf_mul, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       (const uint4*)a, (const uint4*)b, (uint4*)out, n);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

extern "C" int bitz_gf_mul_const_dev(bitz_ctx* ctx, const void* a, const uint8_t c[16], void* out,
                                     uint64_t n) {
    BITZ_REQUIRE(ctx && a && c && out, BITZ_ERR_ARG, "bitz_gf_mul_const_dev: null argument");
    if (n == 0) return BITZ_OK;
    BITZ_HIP(hipSetDevice(ctx->device));
    uint64_t c_lo, c_hi;
    memcpy(&c_lo, c, 8);
    memcpy(&c_hi, c + 8, 8);
    uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_gf_mul_const, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       (const uint4*)a, c_lo, c_hi, (uint4*)out, n);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}

extern "C" int bitz_blake3_leaves_dev(bitz_ctx* ctx, const void* data, uint64_t n,
                                      uint64_t leaf_size, void* out) {
    BITZ_REQUIRE(ctx && data && out, BITZ_ERR_ARG, "bitz_blake3_leaves_dev: null argument");
    BITZ_REQUIRE(leaf_size <= 1024 && leaf_size % 4 == 0, BITZ_ERR_SHAPE,
                 "bitz_blake3_leaves_dev: leaf_size must be a multiple of 4 and <= 1024");
    if (n == 0) return BITZ_OK;
    BITZ_HIP(hipSetDevice(ctx->device));
    uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_blake3_leaves, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       (const uint32_t*)data, n, (uint32_t)(leaf_size / 4), (uint32_t*)out);
    BITZ_LAUNCH_CHECK();
    return BITZ_OK;
}
