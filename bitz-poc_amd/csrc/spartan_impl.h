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

// turns a reduced factored claim (u1, u2 on a derived grid) into one weight
// per committed cell (wf)
using TransposeFn = int (*)(void* user, const void* u1, const void* u2, void* wf);

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
// the BitZ generator g (smallest_generator, src/pcs.rs:85-102)
hgf bitz_generator();

// BitZVerifier::verify_virtual (virt.rs:304-360) after its statement: the
// mirror of open_virtual_claim, reading the opening (narg, hints) of a
// commitment to 2^(tf+sf) cells (custom:1:4)
int verify_virtual_claim(bitz_ctx* ctx, const uint8_t root[32], int tf, int sf, const char* session,
                         const uint8_t tag[32], const std::vector<std::vector<uint8_t>>& lead, int th, int sh, u128 q,
                         const void* roww, const void* colw, u128 claimed, TransposeFn transpose, void* user,
                         const std::vector<hgf>& opt, hgf ood_y, const uint8_t* narg, uint64_t narg_len,
                         const uint8_t* hints, uint64_t hints_len);

// absorb_spartan_message (src/piop/spartan/mod.rs:349-358): the frame
// domain, the tag and the payload, each behind its u64 length
void spartan_frame(Transcript& tr, const char* tag, const void* payload, size_t len);
// squeeze_field (mod.rs:371-384): sample_below(q) over 256 draws, then
// absorbed as a field element; false when sampling is exhausted
bool spartan_squeeze(Transcript& tr, u128 q, u128* out);

// BitZProver::prove_virtual (src/bitz/virt.rs:214-301) after its statement:
// on an inner transcript that starts with the session, the fork tag and the
// `lead` public messages, the claim sum roww[b] colw[c] d(c, b) = claimed
// (mod q) over the derived grid `grid` (th x sh, column-major rows), whose
// folds and GKR run there; `transpose` turns the reduced factored claim
// (u1: 2^th, u2: 2^sh Gf128) into one weight per committed cell (wf:
// 2^(t+s) Gf128), opened as a single-column inner product on the committed
// bits of h. proof = BitZOpeningProof::to_bytes.
int open_virtual_claim(bitz_ctx* ctx, const bitz_commit_handle* h, const char* session, const uint8_t tag[32],
                       const std::vector<std::vector<uint8_t>>& lead, const void* grid, int th, int sh, u128 q,
                       const void* roww, const void* colw, u128 claimed, TransposeFn transpose, void* user,
                       const Round0Claim& r0, std::vector<uint8_t>* proof);

}  // namespace bitz
