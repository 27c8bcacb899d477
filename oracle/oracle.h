/*
 * CPU oracle for the BitZ prover hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain-C restatement of the
 * reference algorithms (albert-garreta/BitZ-PoC and its vendored flock-core /
 * field crates), written from their source read as text. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or as the timed CPU baseline. The product library
 * (bitz-poc_amd/) never links or calls it.
 *
 * Each function cites the reference file:line it restates.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   gf128_*        NIST GCM known-answer tags (525 cases) + field identities
 *   blake3_*       published BLAKE3 digests + incremental/XOF self-consistency
 *   ntt_*          brute-force novel-basis evaluation on tiny sizes, inverse
 *                  round trip, linearity, replicate-from-layer-r == full NTT
 *   merkle_*       hand-built trees from the pinned BLAKE3 primitives,
 *                  multi-proof verification, tamper rejection
 */
#ifndef BITZ_ORACLE_H
#define BITZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t lo, hi;
} gf128_t;

/* ---- GF(2^128) = F2[X]/(X^128 + X^7 + X^2 + X + 1), polynomial basis ---- */
gf128_t or_gf_mul(gf128_t a, gf128_t b);
gf128_t or_gf_square(gf128_t a);
gf128_t or_gf_inv(gf128_t a);
gf128_t or_gf_pow(gf128_t a, uint64_t e_lo, uint64_t e_hi);
void or_gf_mul_arrays(const gf128_t* a, const gf128_t* b, gf128_t* out, size_t n);
/* 1 if the carry-less multiply used PCLMULQDQ, 0 if the portable path */
int or_gf_uses_pclmul(void);
void or_gf_force_portable(int on);

/* ---- BLAKE3 (public specification) ---- */
void or_blake3_hash(const uint8_t* data, size_t len, uint8_t* out, size_t out_len);
/* non-root chaining value of one chunk-or-less input (merkle.rs:192-196) */
void or_blake3_leaf_cv(const uint8_t* data, size_t len, uint8_t out[32]);
/* non-root PARENT chaining value of two children (merkle.rs:198-202) */
void or_blake3_parent_cv(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]);
/* keyed-free compression, exposed for device parity tests */
void or_blake3_compress(const uint32_t cv[8], const uint8_t block[64], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out[16]);
/* incremental hashing: a caller-owned state of or_b3_state_size() bytes;
 * or_b3_finalize reads the root XOF of everything absorbed so far and leaves
 * the state unchanged */
size_t or_b3_state_size(void);
void or_b3_init(void* st);
void or_b3_update(void* st, const uint8_t* data, size_t len);
void or_b3_finalize(const void* st, uint8_t* out, size_t out_len);

/* ---- Additive NTT (LCH novel basis), additive_ntt_f128.rs ---- */
/* evals table for the standard basis {1, x, ..., x^(ell-1)}:
 * out has ell*(ell+1)/2 entries, row i (length ell-i) at offset
 * i*ell - i*(i-1)/2 */
void or_ntt_evals(int ell, gf128_t* out);
gf128_t or_ntt_twiddle(int ell, const gf128_t* evals, int layer, uint64_t block);
void or_ntt_forward_interleaved(gf128_t* data, int log_d, int log_lanes, int start_layer,
                                int ell);
void or_ntt_inverse_interleaved(gf128_t* data, int log_d, int log_lanes, int ell);

/* ---- Merkle tree over BLAKE3 (merkle.rs) ---- */
void or_merkle_tree(const uint8_t* data, uint64_t num_leaves, uint64_t leaf_size,
                    uint8_t* tree /* (2*num_leaves-1)*32 */);
/* positions need not be sorted/unique; returns number of sibling hashes */
size_t or_merkle_multi_proof(const uint8_t* tree, uint64_t num_leaves, const uint64_t* positions,
                             size_t n_pos, uint8_t* proof_out /* capacity n_pos*log2(N)*32 */);
int or_merkle_verify_multi(const uint8_t root[32], uint64_t num_leaves,
                           const uint64_t* sorted_positions, const uint8_t* leaf_hashes,
                           size_t n_pos, const uint8_t* proof, size_t n_proof);

/* ---- PCS commit (pcs/commit.rs:180-315) ---- */
/* packed: 2^log_msg_len elements; codeword: 2^(log_msg_len+log_inv_rate);
 * tree: (2*n_leaves-1)*32 bytes with n_leaves = 2^(log_msg_len-log_batch+log_inv_rate) */
int or_commit(const gf128_t* packed, int log_msg_len, int log_batch, int log_inv_rate,
              gf128_t* codeword, uint8_t* tree, uint8_t root[32]);

#ifdef __cplusplus
}
#endif
#endif
