/*
 * Opening-side hot-path functions of the BitZ PCS, restated on the CPU.
 * Test infrastructure only (see oracle.h). Each function cites the
 * reference function it restates.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"
#include "oracle_open.h"

static inline gf128_t gadd(gf128_t a, gf128_t b) {
    gf128_t o = {a.lo ^ b.lo, a.hi ^ b.hi};
    return o;
}

/* Loops over more than PAR elements run on the OpenMP threads (the CPU
 * baseline); sums in characteristic two are XORs, so the result does not
 * depend on the thread count. */
#define PAR 4096
#pragma omp declare reduction(gx : gf128_t : omp_out = gadd(omp_out, omp_in)) initializer(omp_priv = (gf128_t){0, 0})

void or_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}
int or_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* build_eq_x_r_vec (src/poly/utils.rs:46-110): eq[x] = prod_i (x_i ? r_i : 1 + r_i),
 * index bit i <-> r[i]. */
void or_eq_table(const gf128_t* r, int n, gf128_t* out) {
    out[0].lo = 1;
    out[0].hi = 0;
    for (int i = 0; i < n; i++) {
        uint64_t len = (uint64_t)1 << i;
#pragma omp parallel for if (len > PAR)
        for (uint64_t x = 0; x < len; x++) {
            gf128_t p = out[x];
            gf128_t hi = or_gf_mul(p, r[i]);
            out[x + len] = hi;
            out[x] = gadd(p, hi);
        }
    }
}

/* mle_eval (src/ligerito.rs:91) / forest_tests.rs:27-41: LSB-first folds
 * v'[i] = v[2i] + r (v[2i+1] - v[2i]). */
gf128_t or_mle_eval(const gf128_t* v, const gf128_t* pt, int n) {
    uint64_t len = (uint64_t)1 << n;
    gf128_t* w = (gf128_t*)malloc(len * sizeof(gf128_t));
    gf128_t* x = (gf128_t*)malloc((len / 2 + 1) * sizeof(gf128_t));
    memcpy(w, v, len * sizeof(gf128_t));
    for (int k = 0; k < n; k++) {
        len >>= 1;
#pragma omp parallel for if (len > PAR)
        for (uint64_t i = 0; i < len; i++) {
            gf128_t lo = w[2 * i], hi = w[2 * i + 1];
            x[i] = gadd(lo, or_gf_mul(pt[k], gadd(hi, lo)));
        }
        gf128_t* tmp = w;
        w = x;
        x = tmp;
    }
    gf128_t out = w[0];
    free(w);
    free(x);
    return out;
}

/* fold_column (src/bitz/fold.rs:58-68): eta_c = sum over set bits b of
 * column c of w_b, exact u128 with wrapping. rows: column-major, 2^(t-6)
 * words per column; w and out: (lo, hi) u64 pairs. */
void or_fold_columns(const uint64_t* rows, int t, int s, const uint64_t* w, uint64_t* out) {
    uint64_t words = (uint64_t)1 << (t - 6), cols = (uint64_t)1 << s;
#pragma omp parallel for if (cols * words > PAR)
    for (uint64_t c = 0; c < cols; c++) {
        unsigned __int128 total = 0;
        for (uint64_t wi = 0; wi < words; wi++) {
            uint64_t x = rows[c * words + wi];
            while (x) {
                int b = __builtin_ctzll(x);
                uint64_t idx = wi * 64 + b;
                total += ((unsigned __int128)w[2 * idx + 1] << 64) | w[2 * idx];
                x &= x - 1;
            }
        }
        out[2 * c] = (uint64_t)total;
        out[2 * c + 1] = (uint64_t)(total >> 64);
    }
}

/* column_images / row_images (fold.rs:173-182): g^e for each u128 exponent
 * (FixedBasePow computes the same field element). */
void or_pow_many(gf128_t g, const uint64_t* e, size_t n, gf128_t* out) {
#pragma omp parallel for if (n > 64)
    for (size_t i = 0; i < n; i++) out[i] = or_gf_pow(g, e[2 * i], e[2 * i + 1]);
}

/* Leaves of the batched grand product (src/bitz/forest_tests.rs:66-73,
 * reduce.rs:1-4): leaf[row * 2^s + c] = bit(c, row) ? images[row] : 1. */
void or_forest_leaves(const uint64_t* rows, int t, int s, const gf128_t* images, gf128_t* leaves) {
    uint64_t words = (uint64_t)1 << (t - 6), cols = (uint64_t)1 << s, nrows = (uint64_t)1 << t;
#pragma omp parallel for if (nrows * cols > PAR)
    for (uint64_t row = 0; row < nrows; row++)
        for (uint64_t c = 0; c < cols; c++) {
            uint64_t bit = (rows[c * words + row / 64] >> (row % 64)) & 1;
            gf128_t one = {1, 0};
            leaves[row * cols + c] = bit ? images[row] : one;
        }
}

/* GrandProductCircuit::batched_eval (gkr.rs:93-108), one level:
 * out[i] = in[i] * in[i + n/2]. */
void or_product_level(const gf128_t* in, uint64_t n, gf128_t* out) {
    uint64_t mid = n / 2;
#pragma omp parallel for if (mid > PAR)
    for (uint64_t i = 0; i < mid; i++) out[i] = or_gf_mul(in[i], in[i + mid]);
}

/* kernels::round_sums (src/bitz/kernels.rs:80-110): over the four halves of
 * (l, r) split at h: sum eq*E_end*O_end and sum eq*(E_hi-E_lo)(O_hi-O_lo),
 * E_end = E_hi when send_one else E_lo. */
void or_gkr_round_sums(const gf128_t* l, const gf128_t* r, uint64_t h, const gf128_t* eq,
                       int send_one, gf128_t out[2]) {
    gf128_t s_end = {0, 0}, s_inf = {0, 0};
#pragma omp parallel for if (h > PAR) reduction(gx : s_end, s_inf)
    for (uint64_t x = 0; x < h; x++) {
        gf128_t el = l[x], eh = l[x + h], ol = r[x], oh = r[x + h];
        gf128_t e_end = send_one ? eh : el, o_end = send_one ? oh : ol;
        s_end = gadd(s_end, or_gf_mul(eq[x], or_gf_mul(e_end, o_end)));
        s_inf = gadd(s_inf, or_gf_mul(eq[x], or_gf_mul(gadd(eh, el), gadd(oh, ol))));
    }
    out[0] = s_end;
    out[1] = s_inf;
}

/* fold_halves (gkr.rs:170-177): v[i] = v[i] + rho (v[i+h] - v[i]). */
void or_fold_halves(gf128_t* v, uint64_t h, gf128_t rho) {
#pragma omp parallel for if (h > PAR)
    for (uint64_t i = 0; i < h; i++) v[i] = gadd(v[i], or_gf_mul(rho, gadd(v[i + h], v[i])));
}

/* the same fold into out[0..h) (v unchanged) */
void or_fold_halves_out(const gf128_t* v, uint64_t h, gf128_t rho, gf128_t* out) {
#pragma omp parallel for if (h > PAR)
    for (uint64_t i = 0; i < h; i++) out[i] = gadd(v[i], or_gf_mul(rho, gadd(v[i + h], v[i])));
}

/* xi_combined_rows (src/ligerito.rs:538-561): m[b] = sum_c w2[c] bit(c, b). */
void or_xi_combined_rows(const uint64_t* rows, int t, int s, const gf128_t* w2, gf128_t* out) {
    uint64_t words = (uint64_t)1 << (t - 6), cols = (uint64_t)1 << s;
    memset(out, 0, sizeof(gf128_t) << t);
    /* each word index owns 64 rows of out */
#pragma omp parallel for if (cols * words > PAR)
    for (uint64_t wi = 0; wi < words; wi++)
        for (uint64_t c = 0; c < cols; c++) {
            uint64_t x = rows[c * words + wi];
            while (x) {
                int b = __builtin_ctzll(x);
                out[wi * 64 + b] = gadd(out[wi * 64 + b], w2[c]);
                x &= x - 1;
            }
        }
}

/* fold_rows_eq (src/bitz/sumcheck.rs:84-131): out[c] = sum_b eq[b] bit(c, b). */
void or_fold_rows_eq(const uint64_t* rows, int t, int s, const gf128_t* eq, gf128_t* out) {
    uint64_t words = (uint64_t)1 << (t - 6), cols = (uint64_t)1 << s;
#pragma omp parallel for if (cols * words > PAR)
    for (uint64_t c = 0; c < cols; c++) {
        gf128_t acc = {0, 0};
        for (uint64_t wi = 0; wi < words; wi++) {
            uint64_t x = rows[c * words + wi];
            while (x) {
                int b = __builtin_ctzll(x);
                acc = gadd(acc, eq[wi * 64 + b]);
                x &= x - 1;
            }
        }
        out[c] = acc;
    }
}

/* round_polynomial (src/bitz/sumcheck.rs:404-447): pairs (2k, 2k+1):
 * [a0, a0 + a1 + q, q] with a0 = sum m0 w0, a1 = sum m1 w1,
 * q = sum (m0 + m1)(w0 + w1). */
void or_round_polynomial(const gf128_t* m, const gf128_t* w, uint64_t n, gf128_t out[3]) {
    gf128_t a0 = {0, 0}, a1 = {0, 0}, q = {0, 0};
#pragma omp parallel for if (n / 2 > PAR) reduction(gx : a0, a1, q)
    for (uint64_t k = 0; k < n / 2; k++) {
        gf128_t m0 = m[2 * k], m1 = m[2 * k + 1], w0 = w[2 * k], w1 = w[2 * k + 1];
        a0 = gadd(a0, or_gf_mul(m0, w0));
        a1 = gadd(a1, or_gf_mul(m1, w1));
        q = gadd(q, or_gf_mul(gadd(m0, m1), gadd(w0, w1)));
    }
    out[0] = a0;
    out[1] = gadd(gadd(a0, a1), q);
    out[2] = q;
}

/* sumcheck.rs fold (:454-465) and ligerito fold_and_msg_lsb (:2634-2680):
 * v'[i] = v[2i] + r (v[2i] + v[2i+1]). In place, n -> n/2. */
void or_fold_lsb(gf128_t* v, uint64_t n, gf128_t r) {
    for (uint64_t i = 0; i < n / 2; i++) {
        gf128_t v0 = v[2 * i], v1 = v[2 * i + 1];
        v[i] = gadd(v0, or_gf_mul(r, gadd(v0, v1)));
    }
}

/* the same fold into out[0..n/2) (v unchanged) */
void or_fold_lsb_out(const gf128_t* v, uint64_t n, gf128_t r, gf128_t* out) {
#pragma omp parallel for if (n / 2 > PAR)
    for (uint64_t i = 0; i < n / 2; i++) {
        gf128_t v0 = v[2 * i], v1 = v[2 * i + 1];
        out[i] = gadd(v0, or_gf_mul(r, gadd(v0, v1)));
    }
}

/* fold_1b_rows_naive (vendor/flock-mod/.../pcs/ring_switch.rs:1332-1381):
 * out[i] = sum_x bit_i(p[x]) suffix[x], i in [0, 128). */
void or_fold_1b_rows(const gf128_t* p, const gf128_t* suffix, uint64_t n, gf128_t out[128]) {
    memset(out, 0, 128 * sizeof(gf128_t));
#pragma omp parallel if (n > PAR)
    {
        gf128_t part[128];
        memset(part, 0, sizeof part);
#pragma omp for
        for (uint64_t x = 0; x < n; x++) {
            uint64_t lo = p[x].lo, hi = p[x].hi;
            while (lo) {
                int b = __builtin_ctzll(lo);
                part[b] = gadd(part[b], suffix[x]);
                lo &= lo - 1;
            }
            while (hi) {
                int b = __builtin_ctzll(hi);
                part[64 + b] = gadd(part[64 + b], suffix[x]);
                hi &= hi - 1;
            }
        }
#pragma omp critical
        for (int b = 0; b < 128; b++) out[b] = gadd(out[b], part[b]);
    }
}

/* fold_b128_elems (ring_switch.rs:1486-1560): out[x] = sum over set bits i
 * of suffix[x] of w[i]. */
void or_fold_b128_elems(const gf128_t* suffix, uint64_t n, const gf128_t w[128], gf128_t* out) {
#pragma omp parallel for if (n > PAR)
    for (uint64_t x = 0; x < n; x++) {
        gf128_t acc = {0, 0};
        for (int i = 0; i < 128; i++) {
            uint64_t bit = i < 64 ? (suffix[x].lo >> i) & 1 : (suffix[x].hi >> (i - 64)) & 1;
            if (bit) acc = gadd(acc, w[i]);
        }
        out[x] = acc;
    }
}

/* tensor_algebra_transpose (ring_switch.rs:1415-1440) then inner product
 * (batch_claims, src/bitz/pcs.rs:916-920). */
gf128_t or_batch_claims(const gf128_t claims[128], const gf128_t w[128]) {
    gf128_t acc = {0, 0};
    for (int b = 0; b < 128; b++) {
        gf128_t u = {0, 0};
        for (int i = 0; i < 128; i++) {
            uint64_t bit = b < 64 ? (claims[i].lo >> b) & 1 : (claims[i].hi >> (b - 64)) & 1;
            if (!bit) continue;
            if (i < 64)
                u.lo |= 1ull << i;
            else
                u.hi |= 1ull << (i - 64);
        }
        acc = gadd(acc, or_gf_mul(u, w[b]));
    }
    return acc;
}

/* round_msg_lsb (pcs/ligerito.rs:2499-2545): u0 = sum f0 b0,
 * u2 = sum (f0 + f1)(b0 + b1) over pairs (2j, 2j+1). */
void or_lig_round_msg(const gf128_t* f, const gf128_t* b, uint64_t n, gf128_t out[2]) {
    gf128_t u0 = {0, 0}, u2 = {0, 0};
#pragma omp parallel for if (n / 2 > PAR) reduction(gx : u0, u2)
    for (uint64_t j = 0; j < n / 2; j++) {
        gf128_t f0 = f[2 * j], f1 = f[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
        u0 = gadd(u0, or_gf_mul(f0, b0));
        u2 = gadd(u2, or_gf_mul(gadd(f0, f1), gadd(b0, b1)));
    }
    out[0] = u0;
    out[1] = u2;
}

/* SumcheckProver::glue (ligerito.rs:3125-3146): acc[j] += alpha b_new[j]. */
void or_lig_glue(gf128_t* acc, const gf128_t* bnew, uint64_t n, gf128_t alpha) {
#pragma omp parallel for if (n > PAR)
    for (uint64_t j = 0; j < n; j++) acc[j] = gadd(acc[j], or_gf_mul(alpha, bnew[j]));
}

/* eval_sk_at_vks (ligerito.rs:1720-1745): s_k(v_k) for k = 0..log_n with
 * s_{i+1}(x) = s_i(x)^2 + s_i(v_i) s_i(x), v_i = X^i. */
void or_eval_sk_at_vks(int log_n, gf128_t* out) {
    if (log_n < 0) return;
    for (int k = 0; k <= log_n; k++) out[k].lo = out[k].hi = 0;
    out[0].lo = 1;
    if (log_n == 0) return;
    gf128_t* layer = (gf128_t*)malloc(sizeof(gf128_t) * log_n);
    for (int i = 1; i <= log_n; i++) {
        layer[i - 1].lo = i < 64 ? 1ull << i : 0;
        layer[i - 1].hi = i >= 64 ? 1ull << (i - 64) : 0;
    }
    int cur = log_n;
    for (int i = 0; i < log_n; i++) {
        for (int j = 0; j < cur; j++) {
            gf128_t s = layer[j];
            gf128_t v = gadd(or_gf_square(s), or_gf_mul(out[i], s));
            if (j == 0)
                out[i + 1] = v;
            else
                layer[j - 1] = v;
        }
        cur--;
    }
    free(layer);
}

/* induce_sumcheck_poly's basis half (ligerito.rs:1968-2070,
 * evaluate_scaled_basis_inplace :1760-1795): out[j] = sum_i aw[i] *
 * prod_{k in bits(j)} What_k(q_i), What_k(x) = s_k(x) / s_k(v_k). */
void or_induce_basis(int log_n, const gf128_t* sks_vks, const uint64_t* queries,
                     const gf128_t* alpha_w, size_t nq, gf128_t* out) {
    uint64_t n = (uint64_t)1 << log_n;
    memset(out, 0, n * sizeof(gf128_t));
    gf128_t* inv = (gf128_t*)malloc(sizeof(gf128_t) * (log_n + 1));
    for (int k = 0; k <= log_n; k++) inv[k] = or_gf_inv(sks_vks[k]);
    gf128_t* sx = (gf128_t*)malloc(sizeof(gf128_t) * (log_n + 1));
    gf128_t* basis = (gf128_t*)malloc(n * sizeof(gf128_t));
    for (size_t i = 0; i < nq; i++) {
        gf128_t x = {queries[i], 0};
        if (log_n > 0) {
            sx[0] = x;
            for (int k = 1; k < log_n; k++)
                sx[k] = gadd(or_gf_square(sx[k - 1]), or_gf_mul(sks_vks[k - 1], sx[k - 1]));
            for (int k = 0; k < log_n; k++) sx[k] = or_gf_mul(sx[k], inv[k]);
        }
        basis[0] = alpha_w[i];
        for (int k = 0; k < log_n; k++) {
            uint64_t cur = (uint64_t)1 << k;
#pragma omp parallel for if (cur > PAR)
            for (uint64_t j = 0; j < cur; j++) basis[j + cur] = or_gf_mul(sx[k], basis[j]);
        }
#pragma omp parallel for if (n > PAR)
        for (uint64_t j = 0; j < n; j++) out[j] = gadd(out[j], basis[j]);
    }
    free(inv);
    free(sx);
    free(basis);
}

/* Proof of work (src/bitz/pcs.rs:1252-1298): smallest nonce with
 * leading_zero_bits(BLAKE3(prefix16 || seed16 || nonce_le)) >= bits. */
static uint32_t lz_bits(const uint8_t* d) {
    uint32_t total = 0;
    for (int i = 0; i < 32; i++) {
        uint32_t z = d[i] ? (uint32_t)__builtin_clz((unsigned)d[i]) - 24 : 8;
        total += z;
        if (z != 8) break;
    }
    return total;
}

int or_pow_valid(const uint8_t prefix[16], const uint8_t seed[16], uint64_t nonce, uint32_t bits) {
    if (bits == 0) return nonce == 0;
    uint8_t buf[40], h[32];
    memcpy(buf, prefix, 16);
    memcpy(buf + 16, seed, 16);
    for (int i = 0; i < 8; i++) buf[32 + i] = (uint8_t)(nonce >> (8 * i));
    or_blake3_hash(buf, 40, h, 32);
    return lz_bits(h) >= bits;
}

uint64_t or_pow_find(const uint8_t prefix[16], const uint8_t seed[16], uint32_t bits) {
    if (bits == 0) return 0;
    for (uint64_t nonce = 0;; nonce++)
        if (or_pow_valid(prefix, seed, nonce, bits)) return nonce;
}
