// Host-side protocol pieces of the prover: an incremental BLAKE3 hasher,
// the Fiat-Shamir transcripts, the Ligerito ladder configuration and the
// proof byte writer. Plain C++ (no device code).
#pragma once
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "modq.h"

namespace bitz {

// The state of an incremental BLAKE3 hasher as plain data, for continuing a
// transcript on the device (the device code mirrors Blake3::update and
// Blake3::finalize_xof on this struct).
struct B3State {
    uint32_t stack[54][8];
    uint32_t cv[8];       // current chunk's chaining value
    uint64_t counter;     // current chunk's index
    uint8_t buf[64];      // current block, zero-padded
    uint32_t buf_len;
    uint32_t blocks;      // whole blocks compressed in the current chunk
    int32_t stack_len;
    uint32_t pad;
};

// ---- incremental BLAKE3 (public specification), hash mode ----
class Blake3 {
   public:
    Blake3();
    void update(const void* data, size_t len);
    // Bytes absorbed so far.
    uint64_t position() const { return chunk_.counter * 1024 + chunk_len(); }
    // Absorb 2^k whole chunks, given the chaining value of their subtree,
    // at a position that is a multiple of 2^k chunks; more input must
    // follow. Equal to absorbing the chunks' bytes.
    void absorb_subtree(const uint32_t cv[8], int k);
    // XOF output of everything absorbed so far; the state is unchanged.
    void finalize_xof(uint8_t* out, size_t len) const;
    void export_state(B3State* s) const;

   private:
    struct Chunk {
        uint32_t cv[8];
        uint64_t counter;
        uint8_t buf[64];
        uint8_t buf_len;
        uint8_t blocks;
    };
    void chunk_reset(uint64_t counter);
    size_t chunk_len() const { return (size_t)chunk_.blocks * 64 + chunk_.buf_len; }
    void push_cv(uint32_t cv[8], uint64_t total_chunks);
    Chunk chunk_;
    uint32_t stack_[54][8];
    int stack_len_ = 0;
};

// The compression function: the SSE4.1 row form where the CPU has it, the
// portable one otherwise (identical outputs; both kept for the tests).
void blake3_compress_host(const uint32_t cv[8], const uint8_t block[64], uint64_t counter, uint32_t block_len,
                          uint32_t flags, uint32_t out[16]);
void blake3_compress_portable(const uint32_t cv[8], const uint8_t block[64], uint64_t counter, uint32_t block_len,
                              uint32_t flags, uint32_t out[16]);
void blake3_compress_sse41(const uint32_t cv[8], const uint8_t block[64], uint64_t counter, uint32_t block_len,
                           uint32_t flags, uint32_t out[16]);
bool blake3_using_sse41();
// Chaining values: a whole non-final chunk, a non-root parent node.
void blake3_chunk_cv(const uint8_t chunk[1024], uint64_t counter, uint32_t cv[8]);
void blake3_parent_cv(const uint32_t left[8], const uint32_t right[8], uint32_t cv[8]);
// The aligned power-of-two runs [c, c + 2^k) that tile chunks [c0, c0 + m),
// in order (each run starts at a multiple of its length): their k.
std::vector<int> blake3_subtree_runs(uint64_t c0, uint64_t m);
void blake3_hash_host(const void* data, size_t len, uint8_t out[32]);

// ---- Blake3Transcript (src/transcript/mod.rs:18-96, traits.rs:371-376) ----
class Transcript {
   public:
    // absorb_slice: 0x06 || len u64 || bytes || 0x07
    void absorb_slice(const void* data, size_t len);
    void absorb_slice(const std::string& s) { absorb_slice(s.data(), s.size()); }
    // absorb_slice of n messages of len bytes each (data = the n messages
    // back to back): the framed messages are absorbed in one update, which
    // takes the hasher's multi-threaded whole-chunk path when long.
    void absorb_slices(const void* data, size_t n, size_t len);
    void absorb_u64(uint64_t v) { absorb_slice(&v, 8); }
    // absorb_slice of len bytes whose whole chunks are hashed elsewhere:
    // slice_plan gives the bytes before the first chunk boundary inside the
    // message (lead), the whole chunks after them and the first one's index;
    // absorb_slice_runs then takes the lead bytes, the subtree chaining
    // values of the runs blake3_subtree_runs(c0, chunks) in order, and the
    // remaining len - lead - 1024 chunks bytes.
    struct SlicePlan {
        uint64_t lead, chunks, c0;
    };
    SlicePlan slice_plan(size_t len) const;
    void absorb_slice_runs(size_t len, const SlicePlan& plan, const uint8_t* lead_bytes, const uint32_t* run_cvs,
                           const uint8_t* tail_bytes);
    // get_challenge of 16 bytes: XOF bytes, then absorb 0x12 || bytes || 0x34
    void challenge(uint8_t* out, size_t len);
    // fill_sampling_bytes: XOF bytes, then absorb the sampling-read frame
    void sampling_bytes(uint8_t* out, size_t len);
    void digest(uint8_t out[32]) const { h_.finalize_xof(out, 32); }
    void export_state(B3State* s) const { h_.export_state(s); }
    // absorb_inner: bytes into the hash with no framing (a statement
    // frame's fields, src/ligerito_flock.rs:1083-1098)
    void absorb_inner(const void* data, size_t len) { put(data, len); }
    // Every byte the hash receives from now on is also appended to tape
    // (nullptr stops it); absorb_slice_runs is not taped.
    void set_tape(std::vector<uint8_t>* tape) { tape_ = tape; }

   private:
    void put(const void* data, size_t len) {
        h_.update(data, len);
        if (tape_) tape_->insert(tape_->end(), (const uint8_t*)data, (const uint8_t*)data + len);
    }
    Blake3 h_;
    std::vector<uint8_t>* tape_ = nullptr;
};

// Field element / integer wire forms
struct u128w {
    uint64_t lo, hi;
};

// ---- inner transcript: narg string + hint stream over a Transcript ----
class InnerTranscript {
   public:
    InnerTranscript(const std::string& session, const uint8_t* instance, size_t instance_len);
    void public_message(const void* data, size_t len) { t_.absorb_slice(data, len); }
    void public_message(const std::string& s) { t_.absorb_slice(s); }
    Transcript::SlicePlan public_plan(size_t len) const { return t_.slice_plan(len); }
    void public_message_runs(size_t len, const Transcript::SlicePlan& plan, const uint8_t* lead_bytes,
                             const uint32_t* run_cvs, const uint8_t* tail_bytes) {
        t_.absorb_slice_runs(len, plan, lead_bytes, run_cvs, tail_bytes);
    }
    void prover_message(const void* data, size_t len) {
        t_.absorb_slice(data, len);
        const uint8_t* p = (const uint8_t*)data;
        narg_.insert(narg_.end(), p, p + len);
    }
    // n prover messages of len bytes each, back to back in data
    void prover_messages(const void* data, size_t n, size_t len) {
        t_.absorb_slices(data, n, len);
        const uint8_t* p = (const uint8_t*)data;
        narg_.insert(narg_.end(), p, p + n * len);
    }
    // u32 length-prefixed byte string as one prover message
    void prover_message_bytes(const void* data, size_t len);
    void verifier_gf(uint8_t out[16]) { t_.challenge(out, 16); }
    void hint_bytes(const void* data, size_t len);
    const std::vector<uint8_t>& narg() const { return narg_; }
    void export_state(B3State* s) const { t_.export_state(s); }
    const std::vector<uint8_t>& hints() const { return hints_; }

   private:
    Transcript t_;
    std::vector<uint8_t> narg_, hints_;
};

// ---- Ligerito ladder (custom Johnson, src/ligerito_flock.rs:535-665) ----
struct LigLevel {
    int log_inv_rate, log_msg_cols, log_num_interleaved, k_recursive;
    int queries, grinding_bits, fold_grinding_bits, ood_samples;
    // the level's soundness diagnostics as the security config carries them
    // (expected_eps_pg_bits, expected_eps_query_bits, expected_eps_ood_bits)
    double eps_pg, eps_query, eps_ood;
};
struct LigConfig {
    int m, log_n, initial_k, yr_log_n, target_bits;
    double eta;
    std::vector<LigLevel> levels;
    int round0_grinding_bits;  // Round 0 (ood_round_params) at the target
};
// returns empty string on success, else the rejection reason
std::string custom_johnson_config(int m, int r0, int k0, int target_bits, LigConfig* out);

// ---- the outer transcript's statement, Round 0 and instance (outer.cpp) ----
// bincode (1.3, fixint) of the ladder's LigeritoSecurityConfig and its
// BLAKE3 digest, the ladder identity bound after the statement
void ligerito_bincode(const LigConfig& cfg, std::vector<uint8_t>* out);
void ligerito_digest(const LigConfig& cfg, uint8_t out[32]);
// min(113, 126 - t): the standalone prime's bit width
int standalone_q_bits(int t);
// the standalone statement frame, the session label and the ladder digest
void bind_standalone_statement(Transcript& tr, const uint8_t root[32], const LigConfig& cfg, int t, int s,
                               const uint8_t generator[16]);
// Round 0's header: its domain, the packed variable count, the grinding bits
void ood_round_header(Transcript& tr, int packed_vars, uint32_t bits);
// a typed grinding boundary: frame, domain, index, difficulty, then the
// 32-byte seed; the found (or checked) nonce is absorbed after it
void grinding_seed(Transcript& tr, const char* domain, uint64_t index, uint32_t bits, uint8_t seed[32]);
void absorb_grinding_nonce(Transcript& tr, uint64_t nonce);
// the claim frame (q, claimed value) of a standalone opening
void absorb_standalone_claim(Transcript& tr, u128 q, u128 claimed);
// the prime search over [lo, hi] with its policy bound first; false with
// *err set when the search is exhausted or the interval is invalid
bool sample_prime_context(Transcript& tr, u128 lo, u128 hi, uint32_t security_bits, u128* out, std::string* err);
bool sample_prime_with_bits(Transcript& tr, int bits, u128* out, std::string* err);
// sample_below: uniform below bound from whole u64 sampling reads, at most
// max_draws rejections; false when exhausted
bool sample_below(Transcript& tr, u128 bound, uint64_t max_draws, u128* out);
// a 256-bit draw reduced mod q (two 16-byte challenges, low first)
u128 sample_residue(Transcript& tr, u128 q);

}  // namespace bitz
