/*
 * Additive NTT over GF(2^128) in the LCH novel polynomial basis, restating
 * vendor/flock-mod/crates/flock-core/src/ntt/additive_ntt_f128.rs.
 * Test infrastructure only (see oracle.h).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

static size_t row_off(int ell, int i) { return (size_t)i * ell - (size_t)i * (i - 1) / 2; }

/* generate_evals_from_subspace (additive_ntt_f128.rs:67-103) for the standard
 * basis beta_i = x^i (AdditiveNttF128::standard, :136-140):
 * evals[0] = basis; evals[i][k] = evals[i-1][k+1]*(evals[i-1][k+1]+evals[i-1][0]);
 * then every row is divided by its 0-th entry. */
void or_ntt_evals(int ell, gf128_t* out) {
    if (ell <= 0) return;
    for (int k = 0; k < ell; k++) {
        gf128_t b = {k < 64 ? 1ull << k : 0, k >= 64 ? 1ull << (k - 64) : 0};
        out[k] = b;
    }
    for (int i = 1; i < ell; i++) {
        gf128_t* prev = out + row_off(ell, i - 1);
        gf128_t* cur = out + row_off(ell, i);
        int prev_len = ell - (i - 1);
        for (int k = 1; k < prev_len; k++) {
            gf128_t s = {prev[k].lo ^ prev[0].lo, prev[k].hi ^ prev[0].hi};
            cur[k - 1] = or_gf_mul(prev[k], s);
        }
    }
    for (int i = 0; i < ell; i++) {
        gf128_t* row = out + row_off(ell, i);
        gf128_t inv = or_gf_inv(row[0]);
        for (int k = 0; k < ell - i; k++) row[k] = or_gf_mul(row[k], inv);
    }
}

/* twiddle(layer, block) = span_get(evals[ell-layer-1][1..], block)
 * (additive_ntt_f128.rs:105-116, 153-156) */
gf128_t or_ntt_twiddle(int ell, const gf128_t* evals, int layer, uint64_t block) {
    const gf128_t* v = evals + row_off(ell, ell - layer - 1);
    gf128_t acc = {0, 0};
    for (int j = 0; j < layer; j++) {
        if ((block >> j) & 1) {
            acc.lo ^= v[1 + j].lo;
            acc.hi ^= v[1 + j].hi;
        }
    }
    return acc;
}

/* forward_transform_interleaved_scalar_from_layer (:243-278):
 * data[pos*lanes + lane]; per layer, per block, u += v*tw; v += u. */
void or_ntt_forward_interleaved(gf128_t* data, int log_d, int log_lanes, int start_layer,
                                int ell) {
    size_t lanes = (size_t)1 << log_lanes;
    gf128_t* evals = (gf128_t*)malloc(sizeof(gf128_t) * (row_off(ell, ell) + 1));
    or_ntt_evals(ell, evals);
    gf128_t* tws = (gf128_t*)malloc(sizeof(gf128_t) << (log_d > 0 ? log_d - 1 : 0));
    for (int layer = start_layer; layer < log_d; layer++) {
        size_t num_blocks = (size_t)1 << layer;
        size_t half = (size_t)1 << (log_d - layer - 1);
#pragma omp parallel for if (num_blocks > 64)
        for (size_t block = 0; block < num_blocks; block++) tws[block] = or_ntt_twiddle(ell, evals, layer, block);
        /* every (block, row) pair is one butterfly column of its own */
#pragma omp parallel for if (num_blocks * half * lanes > 4096)
        for (size_t br = 0; br < num_blocks * half; br++) {
            size_t block = br / half, row = br % half;
            gf128_t tw = tws[block];
            size_t start = (block << (log_d - layer)) * lanes;
            {
                gf128_t* top = data + start + row * lanes;
                gf128_t* bot = top + half * lanes;
                for (size_t lane = 0; lane < lanes; lane++) {
                    gf128_t v = bot[lane];
                    gf128_t p = or_gf_mul(v, tw);
                    gf128_t u = {top[lane].lo ^ p.lo, top[lane].hi ^ p.hi};
                    top[lane] = u;
                    bot[lane].lo = v.lo ^ u.lo;
                    bot[lane].hi = v.hi ^ u.hi;
                }
            }
        }
    }
    free(tws);
    free(evals);
}

/* inverse_transform (:762-782), interleaved: layers in reverse,
 * v += u; u += v*tw. */
void or_ntt_inverse_interleaved(gf128_t* data, int log_d, int log_lanes, int ell) {
    size_t lanes = (size_t)1 << log_lanes;
    gf128_t* evals = (gf128_t*)malloc(sizeof(gf128_t) * (row_off(ell, ell) + 1));
    or_ntt_evals(ell, evals);
    for (int layer = log_d - 1; layer >= 0; layer--) {
        size_t num_blocks = (size_t)1 << layer;
        size_t half = (size_t)1 << (log_d - layer - 1);
        for (size_t block = 0; block < num_blocks; block++) {
            gf128_t tw = or_ntt_twiddle(ell, evals, layer, block);
            size_t start = (block << (log_d - layer)) * lanes;
            for (size_t row = 0; row < half; row++) {
                gf128_t* top = data + start + row * lanes;
                gf128_t* bot = top + half * lanes;
                for (size_t lane = 0; lane < lanes; lane++) {
                    gf128_t u = top[lane];
                    gf128_t v = {bot[lane].lo ^ u.lo, bot[lane].hi ^ u.hi};
                    bot[lane] = v;
                    gf128_t p = or_gf_mul(v, tw);
                    top[lane].lo = u.lo ^ p.lo;
                    top[lane].hi = u.hi ^ p.hi;
                }
            }
        }
    }
    free(evals);
}
