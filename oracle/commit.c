/*
 * PCS commit: replicate -> interleaved additive NTT from layer r -> BLAKE3
 * Merkle tree, restating vendor/flock-mod/crates/flock-core/src/pcs/commit.rs
 * :180-315 (commit, commit_into, replicate_message_fill, finalize_commit).
 * The packed message is the committed bit rows read as Gf128 words
 * (src/ligerito_flock.rs:198-217): p_msg[c*2^(t-7)+i] = {row_c[2i], row_c[2i+1]}.
 * Test infrastructure only (see oracle.h).
 */
#include "oracle.h"

#include <string.h>

int or_commit(const gf128_t* packed, int log_msg_len, int log_batch, int log_inv_rate,
              gf128_t* codeword, uint8_t* tree, uint8_t root[32]) {
    /* PcsParams::validate (commit.rs:125-136) */
    if (log_batch >= log_msg_len || log_inv_rate < 1) return -1;
    int log_dim = log_msg_len - log_batch;
    int k_code = log_dim + log_inv_rate;
    size_t msg_len = (size_t)1 << log_msg_len;
    size_t cw_len = (size_t)1 << (log_msg_len + log_inv_rate);
    /* replicate_message_fill (commit.rs:239-258): 2^r copies of the message */
    for (size_t off = 0; off < cw_len; off += msg_len) memcpy(codeword + off, packed, msg_len * 16);
    /* finalize_commit (commit.rs:262-292) */
    or_ntt_forward_interleaved(codeword, k_code, log_batch, log_inv_rate, k_code);
    uint64_t n_leaves = (uint64_t)1 << k_code;
    or_merkle_tree((const uint8_t*)codeword, n_leaves, (uint64_t)16 << log_batch, tree);
    memcpy(root, tree + (2 * n_leaves - 2) * 32, 32);
    return 0;
}
