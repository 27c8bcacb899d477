// Host-side protocol pieces (see host_proto.h).
#include "host_proto.h"

#include <cpuid.h>
#include <immintrin.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>

namespace bitz {

// ------------------------------------------------------------------ BLAKE3
static const uint32_t kIV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static constexpr uint8_t kPerm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { kChunkStart = 1, kChunkEnd = 2, kParent = 4, kRoot = 8 };

static inline uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// message word order of each round: round r reads m[kSched.idx[r][i]] where
// idx[r+1][i] = idx[r][kPerm[i]] (a compile-time table, so the unrolled
// rounds index the message words with constants)
struct Sched {
    uint8_t idx[7][16];
};
static constexpr Sched make_sched() {
    Sched s{};
    for (int i = 0; i < 16; i++) s.idx[0][i] = (uint8_t)i;
    for (int r = 1; r < 7; r++)
        for (int i = 0; i < 16; i++) s.idx[r][i] = s.idx[r - 1][kPerm[i]];
    return s;
}
static constexpr Sched kSched = make_sched();

void blake3_compress_portable(const uint32_t cv[8], const uint8_t block[64], uint64_t counter, uint32_t block_len,
                              uint32_t flags, uint32_t out[16]) {
    uint32_t m[16], s[16];
    for (int i = 0; i < 16; i++)
        m[i] = (uint32_t)block[4 * i] | ((uint32_t)block[4 * i + 1] << 8) | ((uint32_t)block[4 * i + 2] << 16) |
               ((uint32_t)block[4 * i + 3] << 24);
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    s[8] = kIV[0];
    s[9] = kIV[1];
    s[10] = kIV[2];
    s[11] = kIV[3];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    auto g = [&](int a, int b, int c, int d, uint32_t x, uint32_t y) {
        s[a] = s[a] + s[b] + x;
        s[d] = rotr32(s[d] ^ s[a], 16);
        s[c] = s[c] + s[d];
        s[b] = rotr32(s[b] ^ s[c], 12);
        s[a] = s[a] + s[b] + y;
        s[d] = rotr32(s[d] ^ s[a], 8);
        s[c] = s[c] + s[d];
        s[b] = rotr32(s[b] ^ s[c], 7);
    };
#pragma GCC unroll 7
    for (int r = 0; r < 7; r++) {
        const uint8_t* q = kSched.idx[r];
        g(0, 4, 8, 12, m[q[0]], m[q[1]]);
        g(1, 5, 9, 13, m[q[2]], m[q[3]]);
        g(2, 6, 10, 14, m[q[4]], m[q[5]]);
        g(3, 7, 11, 15, m[q[6]], m[q[7]]);
        g(0, 5, 10, 15, m[q[8]], m[q[9]]);
        g(1, 6, 11, 12, m[q[10]], m[q[11]]);
        g(2, 7, 8, 13, m[q[12]], m[q[13]]);
        g(3, 4, 9, 14, m[q[14]], m[q[15]]);
    }
    for (int i = 0; i < 8; i++) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

// The same compression with the state's four rows in 128-bit registers: the
// column step runs the four G functions lane-parallel, the rows are then
// rotated by 1, 2, 3 lanes so the diagonals become columns, and back
// (the 16- and 8-bit rotations are byte shuffles).
__attribute__((target("sse4.1,ssse3"), always_inline)) static inline void b3_g4(__m128i& r0, __m128i& r1, __m128i& r2,
                                                                              __m128i& r3, __m128i mx, __m128i my) {
    const __m128i rot16 = _mm_setr_epi8(2, 3, 0, 1, 6, 7, 4, 5, 10, 11, 8, 9, 14, 15, 12, 13);
    const __m128i rot8 = _mm_setr_epi8(1, 2, 3, 0, 5, 6, 7, 4, 9, 10, 11, 8, 13, 14, 15, 12);
    r0 = _mm_add_epi32(_mm_add_epi32(r0, r1), mx);
    r3 = _mm_shuffle_epi8(_mm_xor_si128(r3, r0), rot16);
    r2 = _mm_add_epi32(r2, r3);
    __m128i x = _mm_xor_si128(r1, r2);
    r1 = _mm_or_si128(_mm_srli_epi32(x, 12), _mm_slli_epi32(x, 20));
    r0 = _mm_add_epi32(_mm_add_epi32(r0, r1), my);
    r3 = _mm_shuffle_epi8(_mm_xor_si128(r3, r0), rot8);
    r2 = _mm_add_epi32(r2, r3);
    x = _mm_xor_si128(r1, r2);
    r1 = _mm_or_si128(_mm_srli_epi32(x, 7), _mm_slli_epi32(x, 25));
}

__attribute__((target("sse4.1,ssse3"))) void blake3_compress_sse41(const uint32_t cv[8], const uint8_t block[64],
                                                                   uint64_t counter, uint32_t block_len,
                                                                   uint32_t flags, uint32_t out[16]) {
    uint32_t m[16];
    memcpy(m, block, 64);  // little-endian words
    const __m128i c0 = _mm_loadu_si128((const __m128i*)cv), c1 = _mm_loadu_si128((const __m128i*)(cv + 4));
    __m128i r0 = c0, r1 = c1;
    __m128i r2 = _mm_setr_epi32((int)kIV[0], (int)kIV[1], (int)kIV[2], (int)kIV[3]);
    __m128i r3 = _mm_setr_epi32((int)(uint32_t)counter, (int)(uint32_t)(counter >> 32), (int)block_len, (int)flags);
    for (int r = 0; r < 7; r++) {
        const uint8_t* q = kSched.idx[r];
        b3_g4(r0, r1, r2, r3, _mm_setr_epi32((int)m[q[0]], (int)m[q[2]], (int)m[q[4]], (int)m[q[6]]),
              _mm_setr_epi32((int)m[q[1]], (int)m[q[3]], (int)m[q[5]], (int)m[q[7]]));
        r1 = _mm_shuffle_epi32(r1, _MM_SHUFFLE(0, 3, 2, 1));
        r2 = _mm_shuffle_epi32(r2, _MM_SHUFFLE(1, 0, 3, 2));
        r3 = _mm_shuffle_epi32(r3, _MM_SHUFFLE(2, 1, 0, 3));
        b3_g4(r0, r1, r2, r3, _mm_setr_epi32((int)m[q[8]], (int)m[q[10]], (int)m[q[12]], (int)m[q[14]]),
              _mm_setr_epi32((int)m[q[9]], (int)m[q[11]], (int)m[q[13]], (int)m[q[15]]));
        r1 = _mm_shuffle_epi32(r1, _MM_SHUFFLE(2, 1, 0, 3));
        r2 = _mm_shuffle_epi32(r2, _MM_SHUFFLE(1, 0, 3, 2));
        r3 = _mm_shuffle_epi32(r3, _MM_SHUFFLE(0, 3, 2, 1));
    }
    _mm_storeu_si128((__m128i*)out, _mm_xor_si128(r0, r2));
    _mm_storeu_si128((__m128i*)(out + 4), _mm_xor_si128(r1, r3));
    _mm_storeu_si128((__m128i*)(out + 8), _mm_xor_si128(r2, c0));
    _mm_storeu_si128((__m128i*)(out + 12), _mm_xor_si128(r3, c1));
}

bool blake3_using_sse41() {
    static const bool v = [] {
        unsigned a, b, c, d;
        // leaf 1 ECX: bit 9 SSSE3, bit 19 SSE4.1
        return __get_cpuid(1, &a, &b, &c, &d) && ((c >> 9) & 1) && ((c >> 19) & 1);
    }();
    return v;
}

void blake3_compress_host(const uint32_t cv[8], const uint8_t block[64], uint64_t counter, uint32_t block_len,
                          uint32_t flags, uint32_t out[16]) {
    if (blake3_using_sse41())
        blake3_compress_sse41(cv, block, counter, block_len, flags, out);
    else
        blake3_compress_portable(cv, block, counter, block_len, flags, out);
}

static void words_le(const uint32_t* w, int n, uint8_t* out) {
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) out[4 * i + k] = (uint8_t)(w[i] >> (8 * k));
}

Blake3::Blake3() { chunk_reset(0); }

void Blake3::export_state(B3State* s) const {
    memset(s, 0, sizeof(*s));
    memcpy(s->stack, stack_, sizeof(stack_));
    memcpy(s->cv, chunk_.cv, 32);
    s->counter = chunk_.counter;
    memcpy(s->buf, chunk_.buf, 64);
    s->buf_len = chunk_.buf_len;
    s->blocks = chunk_.blocks;
    s->stack_len = stack_len_;
}

void Blake3::chunk_reset(uint64_t counter) {
    memcpy(chunk_.cv, kIV, 32);
    chunk_.counter = counter;
    chunk_.buf_len = 0;
    chunk_.blocks = 0;
    memset(chunk_.buf, 0, 64);
}

void Blake3::push_cv(uint32_t cv[8], uint64_t total_chunks) {
    uint32_t cur[8];
    memcpy(cur, cv, 32);
    while ((total_chunks & 1) == 0) {
        uint8_t block[64];
        words_le(stack_[stack_len_ - 1], 8, block);
        words_le(cur, 8, block + 32);
        uint32_t out[16];
        blake3_compress_host(kIV, block, 0, 64, kParent, out);
        memcpy(cur, out, 32);
        stack_len_--;
        total_chunks >>= 1;
    }
    memcpy(stack_[stack_len_++], cur, 32);
}

void blake3_parent_cv(const uint32_t left[8], const uint32_t right[8], uint32_t cv[8]) {
    uint8_t block[64];
    words_le(left, 8, block);
    words_le(right, 8, block + 32);
    uint32_t out[16];
    blake3_compress_host(kIV, block, 0, 64, kParent, out);
    memcpy(cv, out, 32);
}

std::vector<int> blake3_subtree_runs(uint64_t c0, uint64_t m) {
    std::vector<int> ks;
    uint64_t c = c0, end = c0 + m;
    while (c < end) {
        int k = 0;
        while (k < 62 && ((c >> k) & 1) == 0 && c + (2ull << k) <= end) k++;
        ks.push_back(k);
        c += 1ull << k;
    }
    return ks;
}

void Blake3::absorb_subtree(const uint32_t cv[8], int k) {
    if (chunk_len() == 1024) {  // finish the buffered chunk: input follows
        uint32_t out[16];
        blake3_compress_host(chunk_.cv, chunk_.buf, chunk_.counter, chunk_.buf_len,
                             kChunkEnd | (chunk_.blocks == 0 ? kChunkStart : 0), out);
        uint64_t total = chunk_.counter + 1;
        push_cv(out, total);
        chunk_reset(total);
    }
    // at a chunk boundary, at a multiple of 2^k chunks (checked by the caller's plan)
    uint32_t cur[8];
    memcpy(cur, cv, 32);
    uint64_t t = (chunk_.counter >> k) + 1;
    while ((t & 1) == 0) {
        blake3_parent_cv(stack_[stack_len_ - 1], cur, cur);
        stack_len_--;
        t >>= 1;
    }
    memcpy(stack_[stack_len_++], cur, 32);
    chunk_reset(chunk_.counter + (1ull << k));
}

// CV of one whole, non-final 1024-byte chunk
void blake3_chunk_cv(const uint8_t* chunk, uint64_t counter, uint32_t cv[8]) {
    memcpy(cv, kIV, 32);
    for (int b = 0; b < 16; b++) {
        uint32_t out[16];
        blake3_compress_host(cv, chunk + 64 * b, counter, 64, (b == 0 ? kChunkStart : 0) | (b == 15 ? kChunkEnd : 0),
                             out);
        memcpy(cv, out, 32);
    }
}

void Blake3::update(const void* data, size_t len) {
    const uint8_t* p = (const uint8_t*)data;
    // Long inputs: once at a chunk boundary, the whole chunks that more input
    // follows are independent (a chunk's CV depends on its bytes and counter
    // only), so their CVs are computed on several threads and pushed in order.
    const size_t kBulkChunks = 64;
    if (len >= (kBulkChunks + 1) * 1024) {
        size_t lead = chunk_len() == 1024 ? 0 : (1024 - chunk_len()) % 1024;
        if (chunk_len() == 0) lead = 0;
        if (lead) {
            update(p, lead);
            p += lead;
            len -= lead;
        }
        if (chunk_len() == 1024) {  // finish the buffered chunk: input follows
            uint32_t out[16];
            blake3_compress_host(chunk_.cv, chunk_.buf, chunk_.counter, chunk_.buf_len,
                                 kChunkEnd | (chunk_.blocks == 0 ? kChunkStart : 0), out);
            uint64_t total = chunk_.counter + 1;
            push_cv(out, total);
            chunk_reset(total);
        }
        size_t m = (len - 1) / 1024;  // keep at least one byte for the buffered chunk
        if (m >= kBulkChunks && chunk_len() == 0) {
            std::vector<uint32_t> cvs(m * 8);
            const uint64_t c0 = chunk_.counter;
            unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            for (unsigned w = 0; w < nt; w++)
                pool.emplace_back([&, w]() {
                    for (size_t i = w; i < m; i += nt) blake3_chunk_cv(p + 1024 * i, c0 + i, &cvs[8 * i]);
                });
            for (auto& th : pool) th.join();
            for (size_t i = 0; i < m; i++) push_cv(&cvs[8 * i], c0 + i + 1);
            chunk_reset(c0 + m);
            p += 1024 * m;
            len -= 1024 * m;
        }
    }
    while (len > 0) {
        if (chunk_len() == 1024) {
            // finish the chunk (its last block is still buffered)
            uint32_t out[16];
            blake3_compress_host(chunk_.cv, chunk_.buf, chunk_.counter, chunk_.buf_len,
                                 kChunkEnd | (chunk_.blocks == 0 ? kChunkStart : 0), out);
            uint64_t total = chunk_.counter + 1;
            push_cv(out, total);
            chunk_reset(total);
        }
        if (chunk_.buf_len == 64) {
            uint32_t out[16];
            blake3_compress_host(chunk_.cv, chunk_.buf, chunk_.counter, 64, chunk_.blocks == 0 ? kChunkStart : 0,
                                 out);
            memcpy(chunk_.cv, out, 32);
            chunk_.blocks++;
            chunk_.buf_len = 0;
            memset(chunk_.buf, 0, 64);
        }
        size_t want = 64 - chunk_.buf_len;
        size_t room = 1024 - chunk_len();
        size_t take = std::min(std::min(want, room), len);
        memcpy(chunk_.buf + chunk_.buf_len, p, take);
        chunk_.buf_len += (uint8_t)take;
        p += take;
        len -= take;
    }
}

void Blake3::finalize_xof(uint8_t* out, size_t len) const {
    // output node of the current chunk, merged up the CV stack
    uint32_t cv[8];
    memcpy(cv, chunk_.cv, 32);
    uint8_t block[64];
    memcpy(block, chunk_.buf, 64);
    uint32_t block_len = chunk_.buf_len;
    uint64_t counter = chunk_.counter;
    uint32_t flags = kChunkEnd | (chunk_.blocks == 0 ? kChunkStart : 0);
    for (int i = stack_len_ - 1; i >= 0; i--) {
        uint32_t o[16];
        blake3_compress_host(cv, block, counter, block_len, flags, o);
        uint8_t pblock[64];
        words_le(stack_[i], 8, pblock);
        words_le(o, 8, pblock + 32);
        memcpy(cv, kIV, 32);
        memcpy(block, pblock, 64);
        block_len = 64;
        counter = 0;
        flags = kParent;
    }
    uint64_t oc = 0;
    size_t done = 0;
    while (done < len) {
        uint32_t o[16];
        blake3_compress_host(cv, block, oc, block_len, flags | kRoot, o);
        uint8_t bytes[64];
        words_le(o, 16, bytes);
        size_t take = std::min<size_t>(64, len - done);
        memcpy(out + done, bytes, take);
        done += take;
        oc++;
    }
}

void blake3_hash_host(const void* data, size_t len, uint8_t out[32]) {
    Blake3 h;
    h.update(data, len);
    h.finalize_xof(out, 32);
}

// -------------------------------------------------------------- transcripts
void Transcript::absorb_slice(const void* data, size_t len) {
    uint8_t head[9];
    head[0] = 0x06;
    uint64_t l = len;
    memcpy(head + 1, &l, 8);
    put(head, 9);
    put(data, len);
    uint8_t tail = 0x07;
    put(&tail, 1);
}

void Transcript::absorb_slices(const void* data, size_t n, size_t len) {
    const size_t framed = len + 10;
    std::vector<uint8_t> stream(n * framed);
    const uint8_t* p = (const uint8_t*)data;
    const uint64_t l = len;
    for (size_t i = 0; i < n; i++) {
        uint8_t* q = stream.data() + i * framed;
        q[0] = 0x06;
        memcpy(q + 1, &l, 8);
        memcpy(q + 9, p + i * len, len);
        q[9 + len] = 0x07;
    }
    put(stream.data(), stream.size());
}

Transcript::SlicePlan Transcript::slice_plan(size_t len) const {
    const uint64_t start = h_.position() + 9;  // after the 0x06 || len frame
    const uint64_t lead = std::min<uint64_t>((1024 - start % 1024) % 1024, len);
    const uint64_t chunks = (len - lead) / 1024;
    return SlicePlan{lead, chunks, (start + lead) / 1024};
}

void Transcript::absorb_slice_runs(size_t len, const SlicePlan& plan, const uint8_t* lead_bytes,
                                   const uint32_t* run_cvs, const uint8_t* tail_bytes) {
    uint8_t head[9];
    head[0] = 0x06;
    uint64_t l = len;
    memcpy(head + 1, &l, 8);
    h_.update(head, 9);
    h_.update(lead_bytes, plan.lead);
    if (plan.chunks) {
        std::vector<int> ks = blake3_subtree_runs(plan.c0, plan.chunks);
        for (size_t i = 0; i < ks.size(); i++) h_.absorb_subtree(run_cvs + 8 * i, ks[i]);
    }
    h_.update(tail_bytes, len - plan.lead - 1024 * plan.chunks);
    uint8_t tail = 0x07;
    h_.update(&tail, 1);
}

void Transcript::challenge(uint8_t* out, size_t len) {
    h_.finalize_xof(out, len);
    uint8_t a = 0x12, b = 0x34;
    put(&a, 1);
    put(out, len);
    put(&b, 1);
}

void Transcript::sampling_bytes(uint8_t* out, size_t len) {
    h_.finalize_xof(out, len);
    static const char kFrame[] = "bitz/sampling-read/v1";
    put(kFrame, sizeof(kFrame) - 1);
    uint64_t l = len;
    put(&l, 8);
    put(out, len);
}

InnerTranscript::InnerTranscript(const std::string& session, const uint8_t* instance, size_t instance_len) {
    t_.absorb_slice(std::string("bitz/v1"));
    t_.absorb_slice(session);
    t_.absorb_slice(instance, instance_len);
}

void InnerTranscript::prover_message_bytes(const void* data, size_t len) {
    std::vector<uint8_t> m(4 + len);
    uint32_t l = (uint32_t)len;
    memcpy(m.data(), &l, 4);
    memcpy(m.data() + 4, data, len);
    prover_message(m.data(), m.size());
}

void InnerTranscript::hint_bytes(const void* data, size_t len) {
    uint32_t l = (uint32_t)len;
    const uint8_t* lp = (const uint8_t*)&l;
    hints_.insert(hints_.end(), lp, lp + 4);
    const uint8_t* p = (const uint8_t*)data;
    hints_.insert(hints_.end(), p, p + len);
}

// ------------------------------------------------------------ Ligerito ladder
// Formulas of vendor/flock-mod/crates/flock-core/src/pcs/ligerito.rs:700-930
// (Johnson regime with OOD) at the slim template's defaults: eta 0.02,
// query grinding 16 bits, target 100.
static const double kLogQ = 128.0;

static double johnson_log_a(int log_inv_rate, double eta, int log_msg_cols, int log_interleaved) {
    double rho = exp2(-(double)log_inv_rate);
    double sqrt_rho = sqrt(rho);
    double gamma = 1.0 - sqrt_rho - eta;
    double m_param = std::max<double>(ceil(sqrt_rho / (2.0 * eta)), 3.0);
    double half = m_param + 0.5;
    double half5 = half * half * half * half * half;
    double numerator = 2.0 * half5 + 3.0 * half * gamma * rho;
    double denominator = 3.0 * pow(rho, 1.5);
    double n = exp2((double)(log_msg_cols + log_inv_rate));
    double a = (numerator / denominator) * n + half / sqrt_rho;
    return log2(a) + std::max(0.0, (double)log_interleaved - 1.0);
}

static double per_query_bits(int log_inv_rate, double eta) {
    double rho = exp2(-(double)log_inv_rate);
    double gamma = 1.0 - sqrt(rho) - eta;
    return log2(1.0 / (1.0 - gamma));
}

static double ood_bits(int log_inv_rate, double eta, int mu, int samples) {
    double rho = exp2(-(double)log_inv_rate);
    double log2_l = log2(1.0 / (2.0 * eta * sqrt(rho)));
    double log2_mu = log2((double)mu);
    if (samples == 0) return kLogQ - log2_l - log2_mu;
    return samples * (kLogQ - log2_mu) - (2.0 * log2_l - 1.0);
}

std::string custom_johnson_config(int m, int r0, int k0, int target_bits, LigConfig* out) {
    if (m < 20 || m > 35) return "custom Johnson ladders need m in [20, 35]";
    if (r0 < 1 || r0 > 4) return "custom Johnson ladders need log_inv_rate in [1, 4]";
    int log_n = m - 7;
    if (k0 <= 0 || k0 >= log_n) return "custom initial_k out of range";
    LigConfig cfg;
    cfg.m = m;
    cfg.log_n = log_n;
    cfg.initial_k = k0;
    cfg.target_bits = target_bits;
    cfg.eta = 0.02;
    const int grinding = 16;
    struct Shape {
        int mc, il, kr, r;
    };
    std::vector<Shape> shapes = {{log_n - k0, k0, k0, r0}};
    int n_run = log_n - k0, rate = r0;
    while (n_run > 5) {
        int kr = std::min(3, n_run);
        rate++;
        shapes.push_back({n_run - kr, kr, kr, rate});
        n_run -= kr;
    }
    cfg.yr_log_n = n_run;
    for (size_t i = 0; i < shapes.size(); i++) {
        LigLevel lv;
        lv.log_inv_rate = shapes[i].r;
        lv.log_msg_cols = shapes[i].mc;
        lv.log_num_interleaved = shapes[i].il;
        lv.k_recursive = shapes[i].kr;
        lv.grinding_bits = grinding;
        lv.ood_samples = i == 0 ? 0 : 1;
        double need_q = (double)(target_bits - grinding);
        double pq = per_query_bits(lv.log_inv_rate, cfg.eta);
        int q = 1;
        for (; q <= 10000; q++)
            if (q * pq + 1e-3 >= need_q) break;
        if (q > 10000) return "Johnson query search did not converge";
        lv.queries = q;
        long positions = 1L << (lv.log_msg_cols + lv.log_inv_rate);
        if (lv.queries > positions) return "custom Johnson level is too thin for its query count";
        double pg = kLogQ - johnson_log_a(lv.log_inv_rate, cfg.eta, lv.log_msg_cols, lv.log_num_interleaved);
        lv.fold_grinding_bits = (int)std::max(0.0, ceil((double)target_bits - pg));
        lv.eps_pg = pg;
        lv.eps_query = (double)lv.queries * pq;
        const int mu = lv.log_msg_cols + lv.log_num_interleaved;
        lv.eps_ood = ood_bits(lv.log_inv_rate, cfg.eta, mu, lv.ood_samples);
        if (i > 0) {
            while (lv.eps_ood + 1e-3 < (double)target_bits) {
                lv.ood_samples++;
                lv.eps_ood = ood_bits(lv.log_inv_rate, cfg.eta, mu, lv.ood_samples);
            }
        }
        cfg.levels.push_back(lv);
    }
    // Round 0 (ood_round_bits / ood_round_params, src/ligerito_flock.rs:1384-1413)
    {
        const LigLevel& l0 = cfg.levels[0];
        double rho = exp2(-(double)l0.log_inv_rate);
        double list = 1.0 / (2.0 * cfg.eta * sqrt(rho));
        double pairs = std::max(list * (list - 1.0) / 2.0, 1.0);
        double degree = std::max(exp2((double)log_n) - 1.0, 1.0);
        double bits = 128.0 - log2(pairs) - log2(degree);
        double deficit = (double)target_bits - bits;
        cfg.round0_grinding_bits = deficit <= 0.0 ? 0 : (int)ceil(deficit);
    }
    *out = cfg;
    return "";
}

}  // namespace bitz
