// The transcript of a proof continued on the device: BLAKE3 hashing of a
// B3State (host_proto.h) exactly as Blake3::update / Blake3::finalize_xof do
// on the host, and the Blake3Transcript framings (src/transcript/mod.rs:56-96):
// absorb_slice = 0x06 || len u64 || bytes || 0x07, a 16-byte challenge = XOF
// bytes, then absorb 0x12 || bytes || 0x34.
//
// Wave-uniform code: one wavefront (a 64-thread block) calls these functions
// together on a state and byte buffers in LDS. Control flow and the hash
// values are the same in every lane; byte moves are spread over the lanes
// (one LDS instruction per message) and each barrier is a single-wave one.
#pragma once
#include "blake3_dev.h"
#include "host_proto.h"

namespace b3s {

__device__ __forceinline__ void wave_sync() { __syncthreads(); }  // one wave per block

__device__ __forceinline__ void load_words(const uint8_t* b, uint32_t* w) {
    const uint32_t* p = (const uint32_t*)b;  // 4-byte aligned LDS buffers
#pragma unroll
    for (int i = 0; i < 16; i++) w[i] = p[i];
}

__device__ inline void push_cv(bitz::B3State* st, const uint32_t* cv, uint64_t total) {
    uint32_t cur[8];
    for (int i = 0; i < 8; i++) cur[i] = cv[i];
    int len = st->stack_len;
    while ((total & 1) == 0) {
        uint32_t block[16], k[8];
        for (int i = 0; i < 8; i++) {
            block[i] = st->stack[len - 1][i];
            block[8 + i] = cur[i];
        }
        b3::iv_init(k);
        b3::compress(k, block, 0, 64, b3::PARENT);
        for (int i = 0; i < 8; i++) cur[i] = k[i];
        len--;
        total >>= 1;
    }
    wave_sync();
    if (threadIdx.x == 0) {
        for (int i = 0; i < 8; i++) st->stack[len][i] = cur[i];
        st->stack_len = len + 1;
    }
    wave_sync();
}

__device__ inline void zero_buf(bitz::B3State* st) {
    if (threadIdx.x < 16) ((uint32_t*)st->buf)[threadIdx.x] = 0;
}

// absorb src[0, len) (LDS); len may exceed 64
__device__ inline void update(bitz::B3State* st, const uint8_t* src, uint32_t len) {
    const uint32_t lane = threadIdx.x;
    uint32_t pos = 0;
    while (pos < len) {
        uint32_t buf_len = st->buf_len, blocks = st->blocks;
        uint32_t clen = blocks * 64 + buf_len;
        if (clen == 1024) {  // finish the chunk: more input follows
            uint32_t w[16], cv[8];
            load_words(st->buf, w);
            for (int i = 0; i < 8; i++) cv[i] = st->cv[i];
            b3::compress(cv, w, st->counter, buf_len, b3::CHUNK_END | (blocks == 0 ? b3::CHUNK_START : 0u));
            uint64_t total = st->counter + 1;
            push_cv(st, cv, total);
            if (lane < 8) st->cv[lane] = b3::IV[lane];
            zero_buf(st);
            if (lane == 0) {
                st->counter = total;
                st->buf_len = 0;
                st->blocks = 0;
            }
            wave_sync();
            buf_len = 0;
            blocks = 0;
            clen = 0;
        }
        if (buf_len == 64) {
            uint32_t w[16], cv[8];
            load_words(st->buf, w);
            for (int i = 0; i < 8; i++) cv[i] = st->cv[i];
            b3::compress(cv, w, st->counter, 64, blocks == 0 ? b3::CHUNK_START : 0u);
            wave_sync();
            if (lane == 0)
                for (int i = 0; i < 8; i++) st->cv[i] = cv[i];
            zero_buf(st);
            if (lane == 0) {
                st->blocks = blocks + 1;
                st->buf_len = 0;
            }
            wave_sync();
            blocks++;
            buf_len = 0;
            clen = blocks * 64;
        }
        uint32_t take = 64 - buf_len;
        if (1024 - clen < take) take = 1024 - clen;
        if (len - pos < take) take = len - pos;
        if (lane < take) st->buf[buf_len + lane] = src[pos + lane];
        wave_sync();
        if (lane == 0) st->buf_len = buf_len + take;
        wave_sync();
        pos += take;
    }
}

// first 16 XOF bytes of everything absorbed so far, as 4 LE words (every lane)
__device__ inline void finalize16(const bitz::B3State* st, uint32_t* out4) {
    uint32_t cv[8], block[16];
    for (int i = 0; i < 8; i++) cv[i] = st->cv[i];
    load_words(st->buf, block);
    uint32_t block_len = st->buf_len;
    uint64_t counter = st->counter;
    uint32_t flags = b3::CHUNK_END | (st->blocks == 0 ? b3::CHUNK_START : 0u);
    for (int i = st->stack_len - 1; i >= 0; i--) {
        b3::compress(cv, block, counter, block_len, flags);
        for (int k = 0; k < 8; k++) {
            block[k] = st->stack[i][k];
            block[8 + k] = cv[k];
        }
        b3::iv_init(cv);
        block_len = 64;
        counter = 0;
        flags = b3::PARENT;
    }
    b3::compress(cv, block, 0, block_len, flags | b3::ROOT);
    for (int i = 0; i < 4; i++) out4[i] = cv[i];
}

// absorb_slice of data[0, len) (LDS, len <= 256) through the staging buffer
// stage (>= len + 10 bytes, 4-byte aligned, LDS)
__device__ inline void absorb_slice(bitz::B3State* st, const uint8_t* data, uint32_t len, uint8_t* stage) {
    const uint32_t lane = threadIdx.x;
    const uint64_t l = len;
    if (lane < 9) stage[lane] = lane == 0 ? 0x06 : (uint8_t)(l >> (8 * (lane - 1)));
    for (uint32_t i = lane; i < len; i += 64) stage[9 + i] = data[i];
    if (lane == 0) stage[9 + len] = 0x07;
    wave_sync();
    update(st, stage, len + 10);
}

// 16-byte challenge into out (LDS, 16 bytes), absorbed back as 0x12 || out || 0x34
__device__ inline void challenge16(bitz::B3State* st, uint8_t* out, uint8_t* stage) {
    const uint32_t lane = threadIdx.x;
    uint32_t w[4];
    finalize16(st, w);
    wave_sync();
    if (lane == 0) {
        for (int i = 0; i < 4; i++) {
            ((uint32_t*)out)[i] = w[i];
            for (int k = 0; k < 4; k++) stage[1 + 4 * i + k] = (uint8_t)(w[i] >> (8 * k));
        }
        stage[0] = 0x12;
        stage[17] = 0x34;
    }
    wave_sync();
    update(st, stage, 18);
}

}  // namespace b3s
