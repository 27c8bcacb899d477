// Host-side protocol pieces of the prover: an incremental BLAKE3 hasher,
// the Fiat-Shamir transcripts, the Ligerito ladder configuration and the
// proof byte writer. Plain C++ (no device code).
#pragma once
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

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

void blake3_compress_host(const uint32_t cv[8], const uint8_t block[64], uint64_t counter, uint32_t block_len,
                          uint32_t flags, uint32_t out[16]);
void blake3_hash_host(const void* data, size_t len, uint8_t out[32]);

// ---- Blake3Transcript (src/transcript/mod.rs:18-96, traits.rs:371-376) ----
class Transcript {
   public:
    // absorb_slice: 0x06 || len u64 || bytes || 0x07
    void absorb_slice(const void* data, size_t len);
    void absorb_slice(const std::string& s) { absorb_slice(s.data(), s.size()); }
    void absorb_u64(uint64_t v) { absorb_slice(&v, 8); }
    // get_challenge of 16 bytes: XOF bytes, then absorb 0x12 || bytes || 0x34
    void challenge(uint8_t* out, size_t len);
    // fill_sampling_bytes: XOF bytes, then absorb the sampling-read frame
    void sampling_bytes(uint8_t* out, size_t len);
    void digest(uint8_t out[32]) const { h_.finalize_xof(out, 32); }
    void export_state(B3State* s) const { h_.export_state(s); }

   private:
    Blake3 h_;
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
    void prover_message(const void* data, size_t len) {
        t_.absorb_slice(data, len);
        const uint8_t* p = (const uint8_t*)data;
        narg_.insert(narg_.end(), p, p + len);
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
};
struct LigConfig {
    int m, log_n, initial_k, yr_log_n, target_bits;
    double eta;
    std::vector<LigLevel> levels;
    int round0_grinding_bits;  // Round 0 (ood_round_params) at the target
};
// returns empty string on success, else the rejection reason
std::string custom_johnson_config(int m, int r0, int k0, int target_bits, LigConfig* out);

}  // namespace bitz
