// Test harness: the host transcript of libbitz_hip (host_proto.cpp), built
// into a small shared library, so the CPU tests can compare absorbing a long
// message whole with absorbing it as lead bytes, subtree chaining values of
// its whole chunks and tail bytes.
#include <vector>

#include "../bitz-poc_amd/csrc/host_proto.cpp"

extern "C" {
// absorb `pre` bytes, then the message either whole (runs = 0) or through
// subtree runs computed here on the host (runs = 1); out = 32 challenge bytes
void hp_transcript(const uint8_t* pre, uint64_t pre_len, const uint8_t* msg, uint64_t len, int runs, uint8_t* out) {
    bitz::Transcript t;
    t.absorb_slice(pre, pre_len);
    if (!runs) {
        t.absorb_slice(msg, len);
    } else {
        bitz::Transcript::SlicePlan p = t.slice_plan(len);
        std::vector<int> ks = bitz::blake3_subtree_runs(p.c0, p.chunks);
        std::vector<uint32_t> cvs(8 * ks.size());
        const uint8_t* at = msg + p.lead;
        uint64_t c = p.c0;
        for (size_t i = 0; i < ks.size(); i++) {
            uint64_t n = 1ull << ks[i];
            std::vector<uint32_t> level(8 * n);
            for (uint64_t j = 0; j < n; j++) bitz::blake3_chunk_cv(at + 1024 * j, c + j, &level[8 * j]);
            for (; n > 1; n /= 2)
                for (uint64_t j = 0; j < n / 2; j++)
                    bitz::blake3_parent_cv(&level[16 * j], &level[16 * j + 8], &level[8 * j]);
            std::copy(level.begin(), level.begin() + 8, cvs.begin() + 8 * i);
            at += 1024ull << ks[i];
            c += 1ull << ks[i];
        }
        t.absorb_slice_runs(len, p, msg, cvs.data(), at);
    }
    t.challenge(out, 32);
}

int hp_runs(uint64_t c0, uint64_t m, int* ks, int cap) {
    std::vector<int> v = bitz::blake3_subtree_runs(c0, m);
    for (size_t i = 0; i < v.size() && (int)i < cap; i++) ks[i] = v[i];
    return (int)v.size();
}
}

extern "C" {
// the BLAKE3 compression: 0 portable, 1 SSE4.1 rows, 2 the dispatched entry
int hp_has_sse41(void) { return bitz::blake3_using_sse41() ? 1 : 0; }
void hp_compress(int path, const uint32_t* cv, const uint8_t* block, uint64_t counter, uint32_t block_len,
                 uint32_t flags, uint32_t* out) {
    if (path == 1)
        bitz::blake3_compress_sse41(cv, block, counter, block_len, flags, out);
    else if (path == 2)
        bitz::blake3_compress_host(cv, block, counter, block_len, flags, out);
    else
        bitz::blake3_compress_portable(cv, block, counter, block_len, flags, out);
}
}
