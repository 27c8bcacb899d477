/*
 * Flat BLAKE3 Merkle tree and Octopus multi-proof, restating
 * vendor/flock-mod/crates/flock-core/src/merkle.rs:14-22 (layout),
 * :192-202 (leaf/parent), :620-660 (merkle_tree_sequential),
 * :698-828 (multi-proof + verify). Test infrastructure only (see oracle.h).
 *
 * Layout for N = 2^k leaves: tree[0..N) = leaf CVs, then each level above,
 * root last; 2N-1 hashes of 32 bytes.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

void or_merkle_tree(const uint8_t* data, uint64_t num_leaves, uint64_t leaf_size,
                    uint8_t* tree) {
#pragma omp parallel for if (num_leaves > 256)
    for (uint64_t i = 0; i < num_leaves; i++) {
        or_blake3_leaf_cv(data + i * leaf_size, leaf_size, tree + i * 32);
    }
    uint64_t read_start = 0, read_len = num_leaves;
    while (read_len > 1) {
        uint64_t next = read_len >> 1;
#pragma omp parallel for if (next > 1024)
        for (uint64_t i = 0; i < next; i++) {
            or_blake3_parent_cv(tree + (read_start + 2 * i) * 32,
                                tree + (read_start + 2 * i + 1) * 32,
                                tree + (read_start + read_len + i) * 32);
        }
        read_start += read_len;
        read_len = next;
    }
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

/* merkle_multi_proof (merkle.rs:698-745) */
size_t or_merkle_multi_proof(const uint8_t* tree, uint64_t num_leaves, const uint64_t* positions,
                             size_t n_pos, uint8_t* proof_out) {
    if (n_pos == 0 || num_leaves == 1) return 0;
    uint64_t* active = (uint64_t*)malloc(n_pos * sizeof(uint64_t));
    memcpy(active, positions, n_pos * sizeof(uint64_t));
    qsort(active, n_pos, sizeof(uint64_t), cmp_u64);
    size_t n = 0;
    for (size_t i = 0; i < n_pos; i++)
        if (n == 0 || active[n - 1] != active[i]) active[n++] = active[i];
    size_t out = 0;
    uint64_t level_start = 0, level_len = num_leaves;
    while (level_len > 1) {
        size_t m = 0;
        size_t i = 0;
        while (i < n) {
            uint64_t p = active[i];
            int sib = i + 1 < n && active[i + 1] == (p ^ 1);
            if (sib) {
                i += 2;
            } else {
                memcpy(proof_out + out * 32, tree + (level_start + (p ^ 1)) * 32, 32);
                out++;
                i += 1;
            }
            active[m++] = p >> 1;
        }
        n = m;
        level_start += level_len;
        level_len >>= 1;
    }
    free(active);
    return out;
}

/* verify_merkle_multi_proof (merkle.rs:754-828) */
int or_merkle_verify_multi(const uint8_t root[32], uint64_t num_leaves,
                           const uint64_t* pos, const uint8_t* leaf_hashes, size_t n_pos,
                           const uint8_t* proof, size_t n_proof) {
    if (num_leaves == 0 || (num_leaves & (num_leaves - 1))) return 0;
    if (n_pos == 0) return n_proof == 0;
    for (size_t i = 0; i < n_pos; i++) {
        if (pos[i] >= num_leaves) return 0;
        if (i > 0 && pos[i - 1] >= pos[i]) return 0;
    }
    if (num_leaves == 1) return n_proof == 0 && memcmp(leaf_hashes, root, 32) == 0;
    uint64_t* ap = (uint64_t*)malloc(n_pos * sizeof(uint64_t));
    uint8_t* ah = (uint8_t*)malloc(n_pos * 32);
    memcpy(ap, pos, n_pos * sizeof(uint64_t));
    memcpy(ah, leaf_hashes, n_pos * 32);
    size_t n = n_pos, used = 0;
    uint64_t level_len = num_leaves;
    int ok = 1;
    while (level_len > 1 && ok) {
        size_t m = 0, i = 0;
        while (i < n) {
            uint64_t p = ap[i];
            uint8_t l[32], r[32], h[32];
            int sib = i + 1 < n && ap[i + 1] == (p ^ 1);
            if (sib) {
                memcpy(l, ah + i * 32, 32);
                memcpy(r, ah + (i + 1) * 32, 32);
                i += 2;
            } else {
                if (used >= n_proof) {
                    ok = 0;
                    break;
                }
                const uint8_t* s = proof + used * 32;
                used++;
                if ((p & 1) == 0) {
                    memcpy(l, ah + i * 32, 32);
                    memcpy(r, s, 32);
                } else {
                    memcpy(l, s, 32);
                    memcpy(r, ah + i * 32, 32);
                }
                i += 1;
            }
            or_blake3_parent_cv(l, r, h);
            ap[m] = p >> 1;
            memcpy(ah + m * 32, h, 32);
            m++;
        }
        n = m;
        level_len >>= 1;
    }
    if (ok) ok = used == n_proof && n == 1 && memcmp(ah, root, 32) == 0;
    free(ap);
    free(ah);
    return ok;
}
