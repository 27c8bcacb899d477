/* Opening-side oracle functions (test infrastructure only; see oracle.h). */
#ifndef BITZ_ORACLE_OPEN_H
#define BITZ_ORACLE_OPEN_H
#include "oracle.h"
#ifdef __cplusplus
extern "C" {
#endif
void or_eq_table(const gf128_t* r, int n, gf128_t* out);
gf128_t or_mle_eval(const gf128_t* v, const gf128_t* pt, int n);
void or_fold_columns(const uint64_t* rows, int t, int s, const uint64_t* w, uint64_t* out);
void or_pow_many(gf128_t g, const uint64_t* e, size_t n, gf128_t* out);
void or_forest_leaves(const uint64_t* rows, int t, int s, const gf128_t* images, gf128_t* leaves);
void or_product_level(const gf128_t* in, uint64_t n, gf128_t* out);
void or_gkr_round_sums(const gf128_t* l, const gf128_t* r, uint64_t h, const gf128_t* eq,
                       int send_one, gf128_t out[2]);
void or_fold_halves(gf128_t* v, uint64_t h, gf128_t rho);
void or_fold_halves_out(const gf128_t* v, uint64_t h, gf128_t rho, gf128_t* out);
void or_xi_combined_rows(const uint64_t* rows, int t, int s, const gf128_t* w2, gf128_t* out);
void or_fold_rows_eq(const uint64_t* rows, int t, int s, const gf128_t* eq, gf128_t* out);
void or_round_polynomial(const gf128_t* m, const gf128_t* w, uint64_t n, gf128_t out[3]);
void or_fold_lsb(gf128_t* v, uint64_t n, gf128_t r);
void or_fold_lsb_out(const gf128_t* v, uint64_t n, gf128_t r, gf128_t* out);
void or_fold_1b_rows(const gf128_t* p, const gf128_t* suffix, uint64_t n, gf128_t out[128]);
void or_fold_b128_elems(const gf128_t* suffix, uint64_t n, const gf128_t w[128], gf128_t* out);
gf128_t or_batch_claims(const gf128_t claims[128], const gf128_t w[128]);
void or_lig_round_msg(const gf128_t* f, const gf128_t* b, uint64_t n, gf128_t out[2]);
void or_lig_glue(gf128_t* acc, const gf128_t* bnew, uint64_t n, gf128_t alpha);
void or_eval_sk_at_vks(int log_n, gf128_t* out);
void or_induce_basis(int log_n, const gf128_t* sks_vks, const uint64_t* queries,
                     const gf128_t* alpha_w, size_t nq, gf128_t* out);
/* OpenMP threads of the loops above (the CPU baseline); results do not
 * depend on the count */
void or_set_threads(int n);
int or_max_threads(void);
int or_pow_valid(const uint8_t prefix[16], const uint8_t seed[16], uint64_t nonce, uint32_t bits);
uint64_t or_pow_find(const uint8_t prefix[16], const uint8_t seed[16], uint32_t bits);
#ifdef __cplusplus
}
#endif
#endif
