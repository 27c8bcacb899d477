// Test harness: the library's outer-transcript pieces (host_proto.cpp,
// outer.cpp) built into a small shared library, so the CPU tests can read
// the exact byte stream the transcript hash receives and compare it with
// the reference's layout and with the oracle.
#include <vector>

#include "../bitz-poc_amd/csrc/host_proto.cpp"
#include "../bitz-poc_amd/csrc/outer.cpp"

namespace {
int64_t emit(const std::vector<uint8_t>& v, uint8_t* out, int64_t cap) {
    if ((int64_t)v.size() <= cap) memcpy(out, v.data(), v.size());
    return (int64_t)v.size();
}
}  // namespace

extern "C" {
// bincode of the custom:r:k ladder's security config; -1 if rejected
int64_t oh_ladder_bincode(int m, int r, int k, uint8_t* out, int64_t cap) {
    bitz::LigConfig cfg;
    if (!bitz::custom_johnson_config(m, r, k, 100, &cfg).empty()) return -1;
    std::vector<uint8_t> b;
    bitz::ligerito_bincode(cfg, &b);
    return emit(b, out, cap);
}

// the bytes the outer transcript hashes for the standalone statement and
// Round 0's header
int64_t oh_statement_stream(const uint8_t root[32], int t, int s, int r, int k, const uint8_t gen[16],
                            uint8_t* out, int64_t cap) {
    bitz::LigConfig cfg;
    if (!bitz::custom_johnson_config(t + s, r, k, 100, &cfg).empty()) return -1;
    std::vector<uint8_t> tape;
    bitz::Transcript tr;
    tr.set_tape(&tape);
    bitz::bind_standalone_statement(tr, root, cfg, t, s, gen);
    bitz::ood_round_header(tr, t + s - 7, (uint32_t)cfg.round0_grinding_bits);
    return emit(tape, out, cap);
}

// after absorbing `pre` as one slice: the prime of `bits` bits, and the
// stream the search itself absorbed (policy, sampling reads); -1 on failure
int64_t oh_sample_prime(const uint8_t* pre, uint64_t pre_len, int bits, uint8_t q[16], uint8_t* out, int64_t cap) {
    bitz::Transcript tr;
    tr.absorb_slice(pre, pre_len);
    std::vector<uint8_t> tape;
    tr.set_tape(&tape);
    bitz::u128 p;
    std::string err;
    if (!bitz::sample_prime_with_bits(tr, bits, &p, &err)) return -1;
    bitz::u128_put(p, q);
    return emit(tape, out, cap);
}

// after absorbing `pre` as one slice: a grinding boundary's seed and the
// stream it absorbed (frame, domain, index, bits, challenge)
int64_t oh_grinding_seed(const uint8_t* pre, uint64_t pre_len, const char* domain, uint64_t index, uint32_t bits,
                         uint8_t seed[32], uint8_t* out, int64_t cap) {
    bitz::Transcript tr;
    tr.absorb_slice(pre, pre_len);
    std::vector<uint8_t> tape;
    tr.set_tape(&tape);
    bitz::grinding_seed(tr, domain, index, bits, seed);
    return emit(tape, out, cap);
}

// the claim frame's stream
int64_t oh_claim_stream(const uint8_t q[16], const uint8_t claimed[16], uint8_t* out, int64_t cap) {
    std::vector<uint8_t> tape;
    bitz::Transcript tr;
    tr.set_tape(&tape);
    bitz::absorb_standalone_claim(tr, bitz::u128_from(q), bitz::u128_from(claimed));
    return emit(tape, out, cap);
}

// a residue draw after absorbing `pre`
void oh_sample_residue(const uint8_t* pre, uint64_t pre_len, const uint8_t q[16], uint8_t out[16]) {
    bitz::Transcript tr;
    tr.absorb_slice(pre, pre_len);
    bitz::u128_put(bitz::sample_residue(tr, bitz::u128_from(q)), out);
}
}
