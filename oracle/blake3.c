/*
 * BLAKE3, restated from the public BLAKE3 specification (the reference uses
 * the `blake3` crate 1.8.7: merkle.rs:192-202, src/transcript/mod.rs,
 * src/utils/blake3x4.rs). Test infrastructure only (see oracle.h).
 *
 * Merkle leaves are BLAKE3 *non-root* chaining values of the leaf bytes
 * (hazmat finalize_non_root, merkle.rs:192-196); internal nodes are non-root
 * PARENT compressions of the two child CVs (merkle.rs:198-202).
 */
#include "oracle.h"

#include <string.h>

static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};

enum { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };
enum { BLOCK_LEN = 64, CHUNK_LEN = 1024 };

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static inline void g(uint32_t* s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}

void or_blake3_compress(const uint32_t cv[8], const uint8_t block[64], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out[16]) {
    uint32_t m[16], s[16];
    for (int i = 0; i < 16; i++) {
        m[i] = (uint32_t)block[4 * i] | ((uint32_t)block[4 * i + 1] << 8) |
               ((uint32_t)block[4 * i + 2] << 16) | ((uint32_t)block[4 * i + 3] << 24);
    }
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    s[8] = IV[0];
    s[9] = IV[1];
    s[10] = IV[2];
    s[11] = IV[3];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    for (int r = 0; r < 7; r++) {
        g(s, 0, 4, 8, 12, m[0], m[1]);
        g(s, 1, 5, 9, 13, m[2], m[3]);
        g(s, 2, 6, 10, 14, m[4], m[5]);
        g(s, 3, 7, 11, 15, m[6], m[7]);
        g(s, 0, 5, 10, 15, m[8], m[9]);
        g(s, 1, 6, 11, 12, m[10], m[11]);
        g(s, 2, 7, 8, 13, m[12], m[13]);
        g(s, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) {
            uint32_t t[16];
            for (int i = 0; i < 16; i++) t[i] = m[PERM[i]];
            memcpy(m, t, sizeof m);
        }
    }
    for (int i = 0; i < 8; i++) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

/* The "output" of a node before its final compression: enough to produce
 * either a chaining value (non-root) or root XOF bytes. */
typedef struct {
    uint32_t cv[8];
    uint8_t block[64];
    uint32_t block_len;
    uint64_t counter;
    uint32_t flags;
} node_out;

static void words_to_bytes(const uint32_t* w, int n, uint8_t* out) {
    for (int i = 0; i < n; i++) {
        out[4 * i] = (uint8_t)w[i];
        out[4 * i + 1] = (uint8_t)(w[i] >> 8);
        out[4 * i + 2] = (uint8_t)(w[i] >> 16);
        out[4 * i + 3] = (uint8_t)(w[i] >> 24);
    }
}

static void node_cv(const node_out* o, uint32_t cv[8]) {
    uint32_t out[16];
    or_blake3_compress(o->cv, o->block, o->counter, o->block_len, o->flags, out);
    memcpy(cv, out, 32);
}

/* chunk state: all blocks but the last compressed, last kept as node_out */
static void chunk_output(const uint8_t* data, size_t len, uint64_t chunk_counter, node_out* o) {
    uint32_t cv[8];
    memcpy(cv, IV, sizeof cv);
    size_t nblocks = len == 0 ? 1 : (len + BLOCK_LEN - 1) / BLOCK_LEN;
    for (size_t b = 0; b + 1 < nblocks; b++) {
        uint32_t flags = b == 0 ? CHUNK_START : 0;
        uint32_t out[16];
        or_blake3_compress(cv, data + b * BLOCK_LEN, chunk_counter, BLOCK_LEN, flags, out);
        memcpy(cv, out, 32);
    }
    size_t last_off = (nblocks - 1) * BLOCK_LEN;
    size_t last_len = len - last_off;
    memcpy(o->cv, cv, sizeof cv);
    memset(o->block, 0, 64);
    if (last_len) memcpy(o->block, data + last_off, last_len);
    o->block_len = (uint32_t)last_len;
    o->counter = chunk_counter;
    o->flags = CHUNK_END | (nblocks == 1 ? CHUNK_START : 0);
}

static void parent_output(const uint32_t l[8], const uint32_t r[8], node_out* o) {
    memcpy(o->cv, IV, 32);
    words_to_bytes(l, 8, o->block);
    words_to_bytes(r, 8, o->block + 32);
    o->block_len = BLOCK_LEN;
    o->counter = 0;
    o->flags = PARENT;
}

/* Output node of the subtree over data[0..len), chunk counters from
 * `first_chunk`. The left subtree holds the largest power-of-two number of
 * chunks strictly less than the total (BLAKE3 spec, section 2.1). */
static void subtree_output(const uint8_t* data, size_t len, uint64_t first_chunk, node_out* o) {
    if (len <= CHUNK_LEN) {
        chunk_output(data, len, first_chunk, o);
        return;
    }
    size_t chunks = (len + CHUNK_LEN - 1) / CHUNK_LEN;
    size_t left_chunks = 1;
    while (left_chunks * 2 < chunks) left_chunks *= 2;
    size_t left_len = left_chunks * CHUNK_LEN;
    node_out lo, ro;
    subtree_output(data, left_len, first_chunk, &lo);
    subtree_output(data + left_len, len - left_len, first_chunk + left_chunks, &ro);
    uint32_t lcv[8], rcv[8];
    node_cv(&lo, lcv);
    node_cv(&ro, rcv);
    parent_output(lcv, rcv, o);
}

void or_blake3_hash(const uint8_t* data, size_t len, uint8_t* out, size_t out_len) {
    node_out o;
    subtree_output(data, len, 0, &o);
    uint64_t ctr = 0;
    size_t done = 0;
    while (done < out_len) {
        uint32_t w[16];
        uint8_t bytes[64];
        or_blake3_compress(o.cv, o.block, ctr, o.block_len, o.flags | ROOT, w);
        words_to_bytes(w, 16, bytes);
        size_t take = out_len - done < 64 ? out_len - done : 64;
        memcpy(out + done, bytes, take);
        done += take;
        ctr++;
    }
}

void or_blake3_leaf_cv(const uint8_t* data, size_t len, uint8_t out[32]) {
    node_out o;
    subtree_output(data, len, 0, &o);
    uint32_t cv[8];
    node_cv(&o, cv);
    words_to_bytes(cv, 8, out);
}

void or_blake3_parent_cv(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]) {
    uint8_t block[64];
    memcpy(block, left, 32);
    memcpy(block + 32, right, 32);
    uint32_t w[16];
    or_blake3_compress(IV, block, 0, BLOCK_LEN, PARENT, w);
    words_to_bytes(w, 8, out);
}

/* Incremental hashing (BLAKE3 specification, section 5.1): the chunk being
 * filled keeps its last block uncompressed (it may be the chunk's final
 * block), and completed chunks' chaining values sit on a stack merged into
 * parents whenever the number of completed chunks has a trailing zero bit.
 * Finalizing reads the state without changing it, so a transcript can keep
 * absorbing after a challenge. */
typedef struct {
    uint32_t cv[8];
    uint64_t chunk;
    uint8_t buf[64];
    uint32_t buf_len;
    uint32_t blocks;
    uint32_t depth;
    uint32_t stack[54][8];
} b3_state;

size_t or_b3_state_size(void) { return sizeof(b3_state); }

void or_b3_init(void* st) {
    b3_state* h = (b3_state*)st;
    memset(h, 0, sizeof *h);
    memcpy(h->cv, IV, sizeof h->cv);
}

static void b3_push_chunk(b3_state* h, uint32_t cv[8]) {
    uint64_t done = h->chunk;
    while ((done & 1) == 0) {
        uint8_t block[64];
        uint32_t w[16];
        h->depth--;
        words_to_bytes(h->stack[h->depth], 8, block);
        words_to_bytes(cv, 8, block + 32);
        or_blake3_compress(IV, block, 0, BLOCK_LEN, PARENT, w);
        memcpy(cv, w, 32);
        done >>= 1;
    }
    memcpy(h->stack[h->depth++], cv, 32);
}

void or_b3_update(void* st, const uint8_t* data, size_t len) {
    b3_state* h = (b3_state*)st;
    while (len) {
        if (h->blocks * BLOCK_LEN + h->buf_len == CHUNK_LEN) {
            /* the chunk is full and more input follows: close it */
            node_out o;
            memcpy(o.cv, h->cv, 32);
            memcpy(o.block, h->buf, 64);
            o.block_len = BLOCK_LEN;
            o.counter = h->chunk;
            o.flags = CHUNK_END | (h->blocks == 0 ? CHUNK_START : 0);
            uint32_t cv[8];
            node_cv(&o, cv);
            h->chunk++;
            b3_push_chunk(h, cv);
            memcpy(h->cv, IV, 32);
            h->blocks = 0;
            h->buf_len = 0;
        }
        if (h->buf_len == BLOCK_LEN) {
            uint32_t w[16];
            or_blake3_compress(h->cv, h->buf, h->chunk, BLOCK_LEN, h->blocks == 0 ? CHUNK_START : 0, w);
            memcpy(h->cv, w, 32);
            h->blocks++;
            h->buf_len = 0;
        }
        size_t take = BLOCK_LEN - h->buf_len;
        if (take > len) take = len;
        memcpy(h->buf + h->buf_len, data, take);
        h->buf_len += (uint32_t)take;
        data += take;
        len -= take;
    }
}

void or_b3_finalize(const void* st, uint8_t* out, size_t out_len) {
    const b3_state* h = (const b3_state*)st;
    node_out o;
    memcpy(o.cv, h->cv, 32);
    memset(o.block, 0, 64);
    memcpy(o.block, h->buf, h->buf_len);
    o.block_len = h->buf_len;
    o.counter = h->chunk;
    o.flags = CHUNK_END | (h->blocks == 0 ? CHUNK_START : 0);
    for (uint32_t i = h->depth; i-- > 0;) {
        uint32_t cv[8];
        node_cv(&o, cv);
        parent_output(h->stack[i], cv, &o);
    }
    uint64_t ctr = 0;
    size_t done = 0;
    while (done < out_len) {
        uint32_t w[16];
        uint8_t bytes[64];
        or_blake3_compress(o.cv, o.block, ctr, o.block_len, o.flags | ROOT, w);
        words_to_bytes(w, 16, bytes);
        size_t take = out_len - done < 64 ? out_len - done : 64;
        memcpy(out + done, bytes, take);
        done += take;
        ctr++;
    }
}
