// The pieces of the standalone opening that a composed proof reuses: Round 0
// on the outer transcript and the BitZ opening (and its verification) of a
// linear claim sum_{c,b} roww[b] colw[c] bit(c, b) = claimed (mod q) on an
// inner transcript that starts with a session label, a 32-byte fork tag and
// an optional lead message. Defined in prover.hip.
#pragma once
#include <stdint.h>

#include <vector>

#include "common.h"
#include "host_gf.h"
#include "host_proto.h"

namespace bitz {

struct Round0Claim {
    std::vector<hgf> point;  // the out-of-domain point (zeta, zeta^2, ...)
    hgf y;                   // the packed message's MLE there
    bool has_nonce;
    uint64_t nonce;
};

// prove_ood_round_packed (src/ligerito_flock.rs:1520-1549) for the committed
// message of h, on `outer` (header, grinding boundary, zeta, y absorbed)
int round0_prove(bitz_ctx* ctx, const bitz_commit_handle* h, Transcript& outer, Round0Claim* r0);
// BitZProver::prove of the claim; proof = BitZOpeningProof::to_bytes
int open_linear_claim(bitz_ctx* ctx, const bitz_commit_handle* h, const char* session, const uint8_t tag[32],
                      const uint8_t* lead, size_t lead_len, u128 q, const void* roww, const void* colw,
                      u128 claimed, const Round0Claim& r0, std::vector<uint8_t>* proof);
// BitZVerifier::verify of the claim, from the parsed opening (narg, hints)
int verify_linear_claim(bitz_ctx* ctx, const uint8_t root[32], int t, int s, int r0, int k0, const char* session,
                        const uint8_t tag[32], const uint8_t* lead, size_t lead_len, u128 q, const void* roww,
                        const void* colw, u128 claimed, const std::vector<hgf>& opt, hgf ood_y,
                        const uint8_t* narg, uint64_t narg_len, const uint8_t* hints, uint64_t hints_len);
bool round0_pow_valid(const uint8_t seed[32], uint64_t nonce, uint32_t bits);

}  // namespace bitz
