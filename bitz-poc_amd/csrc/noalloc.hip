// The ring-switch bit fold for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// allocation). bitz_fold_1b_rows_noalloc runs the kernels of
// bitz_fold_1b_rows_dev, with the block partials in a fixed device array of
// this file instead of the context's pool: 512 chunks x 128 sums x 16 bytes
// = 1 MiB per GPU, held until the process ends.
//
// Every call on one GPU shares that GPU's array, which keeps the last
// call's partials. Each call writes every partial it reads, so nothing needs
// clearing between calls; calls that overlap on one GPU (two streams, two
// threads) race on the array and must be serialised by the caller.
#include <mutex>

#include "open_impl.h"

// Device memory of this file. In host code the name is a stand-in, not a
// device address: kernels take the address hipGetSymbolAddress returns.
static __device__ __align__(256) uint4 scratch[(size_t)bitz::kFold1bMaxChunks * 128];

namespace bitz {
namespace {

constexpr size_t kScratchEntries = (size_t)kFold1bMaxChunks * 128;  // uint4 entries
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

// Resolves the array's address on `device` (which must be current), once per
// device. Called by bitz_ctx_create.
int noalloc_setup(int device) {
    BITZ_REQUIRE(device >= 0 && device < kMaxDevices, BITZ_ERR_ARG, "noalloc_setup: device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return BITZ_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
        set_error(std::string("noalloc_setup: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
        return BITZ_ERR_DEVICE;
    }
    g_scratch[device] = ptr;
    return BITZ_OK;
}

}  // namespace bitz

extern "C" int bitz_fold_1b_rows_noalloc(bitz_ctx* ctx, const void* p, const void* suffix, uint64_t n, void* out) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "bitz_fold_1b_rows_noalloc: null context");
    BITZ_REQUIRE(p && suffix && out, BITZ_ERR_ARG, "bitz_fold_1b_rows_noalloc: null argument");
    BITZ_REQUIRE(ctx->device >= 0 && ctx->device < bitz::kMaxDevices, BITZ_ERR_ARG,
                 "bitz_fold_1b_rows_noalloc: device index out of range");
    BITZ_HIP(hipSetDevice(ctx->device));
    void* partials;
    {
        std::lock_guard<std::mutex> lock(bitz::g_mu);
        partials = bitz::g_scratch[ctx->device];
    }
    BITZ_REQUIRE(partials != nullptr, BITZ_ERR_DEVICE,
                 "bitz_fold_1b_rows_noalloc: scratch array not set up on this device");
    const size_t need = (size_t)bitz::fold_1b_rows_chunks(n) * 128;
    BITZ_REQUIRE(need <= bitz::kScratchEntries, BITZ_ERR_SHAPE,
                 "bitz_fold_1b_rows_noalloc: partials exceed the 1 MiB scratch array");
    return bitz::fold_1b_rows_into(ctx, p, suffix, n, out, partials);
}
