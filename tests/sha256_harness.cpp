// Test harness: the library's host SHA-256 compression (sha256_host.cpp),
// both paths exported, so the CPU tests can check each against hashlib.
#include "../bitz-poc_amd/csrc/sha256_host.cpp"

extern "C" {
int sh_has_shani(void) { return bitz::chain::using_shani() ? 1 : 0; }
void sh_compress(int path, const uint32_t* state, const uint32_t* block, uint32_t* out) {
    if (path == 1)
        bitz::chain::compress_shani(state, block, out);
    else if (path == 2)
        bitz::chain::compress(state, block, out);
    else
        bitz::chain::compress_portable(state, block, out);
}
}
