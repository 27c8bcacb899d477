// C ABI over the opening kernels (device pointers; see include/bitz_hip.h).
#include "host_proto.h"
#include "open_impl.h"

namespace {
int prelude(bitz_ctx* ctx) {
    BITZ_REQUIRE(ctx != nullptr, BITZ_ERR_ARG, "null context");
    BITZ_HIP(hipSetDevice(ctx->device));
    return BITZ_OK;
}
#define PRELUDE()                                \
    if (int _rc = prelude(ctx)) return _rc;      \
    bitz::PoolScope _pool_scope(&ctx->pool)
#define CALL(expr)                  \
    do {                            \
        int _rc = (expr);           \
        if (_rc) return _rc;        \
    } while (0)

// copy a host vector of field elements into the scratch "misc" buffer
int stage(bitz_ctx* ctx, const uint8_t* host, size_t bytes, void** dev) {
    if (!ctx->scratch.misc.alloc(bytes ? bytes : 16)) return BITZ_ERR_ALLOC;
    if (bytes) {
        BITZ_HIP(hipMemcpyAsync(ctx->scratch.misc.ptr, host, bytes, hipMemcpyHostToDevice, ctx->stream));
        BITZ_HIP(hipStreamSynchronize(ctx->stream));  // the host buffer may not outlive the call
    }
    *dev = ctx->scratch.misc.ptr;
    return BITZ_OK;
}
}  // namespace

extern "C" int bitz_eq_table_dev(bitz_ctx* ctx, const uint8_t* r_host, int n, void* out) {
    PRELUDE();
    BITZ_REQUIRE(out && (n == 0 || r_host), BITZ_ERR_ARG, "bitz_eq_table_dev: null argument");
    BITZ_REQUIRE(n >= 0 && n <= 28, BITZ_ERR_SHAPE, "bitz_eq_table_dev: n must be in [0, 28]");
    std::vector<bitz::hgf> r(n > 0 ? n : 1);
    for (int i = 0; i < n; i++) {
        memcpy(&r[i].lo, r_host + 16 * i, 8);
        memcpy(&r[i].hi, r_host + 16 * i + 8, 8);
    }
    return bitz::eq_args(ctx, r.data(), n, out);
}

extern "C" int bitz_fold_columns_dev(bitz_ctx* ctx, const void* rows, int t, int s, const void* w, void* out) {
    PRELUDE();
    BITZ_REQUIRE(rows && w && out, BITZ_ERR_ARG, "bitz_fold_columns_dev: null argument");
    BITZ_REQUIRE(t >= 7 && t <= 35 && s >= 0 && s <= 35 - t, BITZ_ERR_SHAPE, "bitz_fold_columns_dev: bad shape");
    return bitz::fold_columns(ctx, (const uint64_t*)rows, t, s, w, out);
}

extern "C" int bitz_pow_images_dev(bitz_ctx* ctx, const uint8_t g[16], const void* exps, uint64_t n, void* out) {
    PRELUDE();
    BITZ_REQUIRE(g && (n == 0 || (exps && out)), BITZ_ERR_ARG, "bitz_pow_images_dev: null argument");
    return bitz::pow_images(ctx, g, exps, n, out);
}

extern "C" int bitz_pack_columns_dev(bitz_ctx* ctx, const void* rows, int t, int s, void* packed) {
    PRELUDE();
    BITZ_REQUIRE(rows && packed, BITZ_ERR_ARG, "bitz_pack_columns_dev: null argument");
    BITZ_REQUIRE(t >= 6 && s >= 0 && t + s <= 35, BITZ_ERR_SHAPE, "bitz_pack_columns_dev: bad shape");
    return bitz::pack_columns(ctx, (const uint64_t*)rows, t, s, (uint64_t*)packed);
}

extern "C" int bitz_gkr_round_sums_dev(bitz_ctx* ctx, const void* l, const void* r, uint64_t h, const void* eq_c,
                                       int s_c, const void* eq_y, int send_one, uint8_t out_host[32]) {
    PRELUDE();
    BITZ_REQUIRE(l && r && eq_c && eq_y && out_host, BITZ_ERR_ARG, "bitz_gkr_round_sums_dev: null argument");
    BITZ_REQUIRE(h > 0 && s_c >= 0 && (1ull << s_c) <= h && h % (1ull << s_c) == 0, BITZ_ERR_SHAPE,
                 "bitz_gkr_round_sums_dev: eq_c must tile the half");
    CALL(bitz::gkr_round_sums(ctx, l, r, h, eq_c, s_c, eq_y, send_one, nullptr));
    return bitz::read_result(ctx, 2, out_host);
}

extern "C" int bitz_gkr_fold_round_sums_dev(bitz_ctx* ctx, void* l, void* r, uint64_t h, const uint8_t rho[16],
                                            const void* eq_c, int s_c, const void* eq_y, int send_one,
                                            uint8_t out_host[32]) {
    PRELUDE();
    BITZ_REQUIRE(l && r && rho && eq_y && out_host, BITZ_ERR_ARG, "bitz_gkr_fold_round_sums_dev: null argument");
    BITZ_REQUIRE(h > 0 && s_c >= 0 && (1ull << s_c) <= h && h % (1ull << s_c) == 0, BITZ_ERR_SHAPE,
                 "bitz_gkr_fold_round_sums_dev: eq_c must tile the half");
    CALL(bitz::gkr_fold_round_sums(ctx, l, r, h, rho, eq_c, s_c, eq_y, send_one, nullptr));
    return bitz::read_result(ctx, 2, out_host);
}

extern "C" int bitz_gkr_leaf_round1_sums_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                                             const uint8_t rho0[16], const void* eq_c, const void* eq_y,
                                             int send_one, uint8_t out_host[32]) {
    PRELUDE();
    BITZ_REQUIRE(packed && A && rho0 && eq_c && eq_y && out_host, BITZ_ERR_ARG,
                 "bitz_gkr_leaf_round1_sums_dev: null argument");
    BITZ_REQUIRE(t >= 3 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_leaf_round1_sums_dev: bad shape");
    CALL(bitz::gkr_leaf_round1_sums(ctx, (const uint64_t*)packed, A, t, s, rho0, eq_c, eq_y, send_one, nullptr));
    return bitz::read_result(ctx, 2, out_host);
}

extern "C" int bitz_gkr_premul_round_sums_dev(bitz_ctx* ctx, const void* l, void* r, uint64_t h, const void* eq_c,
                                              int s, const void* eq_y, int send_one, uint8_t out_host[32]) {
    PRELUDE();
    BITZ_REQUIRE(l && r && eq_c && eq_y && out_host, BITZ_ERR_ARG, "bitz_gkr_premul_round_sums_dev: null argument");
    BITZ_REQUIRE(s >= 0 && s <= 40 && h > 0 && (1ull << s) <= h && h % (1ull << s) == 0, BITZ_ERR_SHAPE,
                 "bitz_gkr_premul_round_sums_dev: eq_c must tile the half");
    CALL(bitz::gkr_premul_round_sums(ctx, l, r, h, eq_c, s, eq_y, send_one, nullptr));
    return bitz::read_result(ctx, 2, out_host);
}

extern "C" int bitz_gkr_round_sums_rowonly_dev(bitz_ctx* ctx, const void* l, const void* r, uint64_t h, int s,
                                               const void* eq_y, int send_one, uint8_t out_host[32]) {
    PRELUDE();
    BITZ_REQUIRE(l && r && eq_y && out_host, BITZ_ERR_ARG, "bitz_gkr_round_sums_rowonly_dev: null argument");
    BITZ_REQUIRE(s >= 0 && s <= 40 && h > 0 && (1ull << s) <= h && h % (1ull << s) == 0, BITZ_ERR_SHAPE,
                 "bitz_gkr_round_sums_rowonly_dev: rows of 2^s must tile the half");
    CALL(bitz::gkr_round_sums_rowonly(ctx, l, r, h, s, eq_y, send_one, nullptr));
    return bitz::read_result(ctx, 2, out_host);
}

extern "C" int bitz_tensor_table_dev(bitz_ctx* ctx, const void* a, const void* b, int n, void* out) {
    PRELUDE();
    BITZ_REQUIRE(a && b && out, BITZ_ERR_ARG, "bitz_tensor_table_dev: null argument");
    return bitz::tensor_table(ctx, a, b, n, out);
}

extern "C" int bitz_gkr_leaf_fold2_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                                       const uint8_t rho0[16], const uint8_t rho1[16], const void* eq_c, void* l_out,
                                       void* r_out) {
    PRELUDE();
    BITZ_REQUIRE(packed && A && rho0 && rho1 && l_out && r_out, BITZ_ERR_ARG, "bitz_gkr_leaf_fold2_dev: null argument");
    BITZ_REQUIRE(t >= 3 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_leaf_fold2_dev: bad shape");
    return bitz::gkr_leaf_fold2(ctx, (const uint64_t*)packed, A, t, s, rho0, rho1, eq_c, l_out, r_out);
}

extern "C" int bitz_gkr_leaf_round_sums_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                                            const void* eq_c, int s_c, const void* eq_y, int send_one,
                                            uint8_t out_host[32]) {
    PRELUDE();
    BITZ_REQUIRE(packed && A && eq_c && eq_y && out_host, BITZ_ERR_ARG, "bitz_gkr_leaf_round_sums_dev: null argument");
    BITZ_REQUIRE(t >= 2 && s_c <= t + s - 2, BITZ_ERR_SHAPE, "bitz_gkr_leaf_round_sums_dev: bad shape");
    CALL(bitz::gkr_leaf_round_sums(ctx, (const uint64_t*)packed, A, t, s, eq_c, s_c, eq_y, send_one, nullptr));
    return bitz::read_result(ctx, 2, out_host);
}

extern "C" int bitz_fold_halves_dev(bitz_ctx* ctx, void* v, uint64_t h, const uint8_t rho[16]) {
    PRELUDE();
    BITZ_REQUIRE(v && rho, BITZ_ERR_ARG, "bitz_fold_halves_dev: null argument");
    return bitz::fold_halves(ctx, v, h, rho);
}

extern "C" int bitz_gkr_leaf_fold_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                                      const uint8_t rho[16], void* l_out, void* r_out) {
    PRELUDE();
    BITZ_REQUIRE(packed && A && rho && l_out && r_out, BITZ_ERR_ARG, "bitz_gkr_leaf_fold_dev: null argument");
    return bitz::gkr_leaf_fold(ctx, (const uint64_t*)packed, A, t, s, rho, l_out, r_out);
}

extern "C" int bitz_gkr_level2_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s, void* out) {
    PRELUDE();
    BITZ_REQUIRE(packed && A && out, BITZ_ERR_ARG, "bitz_gkr_level2_dev: null argument");
    BITZ_REQUIRE(t >= 2 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_level2_dev: bad shape");
    return bitz::gkr_level2(ctx, (const uint64_t*)packed, A, t, s, out);
}

extern "C" int bitz_gkr_level3_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s, void* out) {
    PRELUDE();
    BITZ_REQUIRE(packed && A && out, BITZ_ERR_ARG, "bitz_gkr_level3_dev: null argument");
    BITZ_REQUIRE(t >= 3 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_level3_dev: bad shape");
    return bitz::gkr_level3(ctx, (const uint64_t*)packed, A, t, s, out);
}

extern "C" int bitz_gkr_level1_round0_sums_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                                               const void* eq_c, const void* eq_y, int send_one,
                                               uint8_t out_host[32]) {
    PRELUDE();
    BITZ_REQUIRE(packed && A && eq_c && eq_y && out_host, BITZ_ERR_ARG,
                 "bitz_gkr_level1_round0_sums_dev: null argument");
    BITZ_REQUIRE(t >= 3 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_level1_round0_sums_dev: bad shape");
    CALL(bitz::gkr_level1_round0_sums(ctx, (const uint64_t*)packed, A, t, s, eq_c, eq_y, send_one, nullptr));
    return bitz::read_result(ctx, 2, out_host);
}

extern "C" int bitz_gkr_level1_fold_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                                        const uint8_t rho[16], const void* eq_c, void* l_out, void* r_out) {
    PRELUDE();
    BITZ_REQUIRE(packed && A && rho && l_out && r_out, BITZ_ERR_ARG, "bitz_gkr_level1_fold_dev: null argument");
    BITZ_REQUIRE(t >= 3 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_level1_fold_dev: bad shape");
    return bitz::gkr_level1_fold(ctx, (const uint64_t*)packed, A, t, s, rho, eq_c, l_out, r_out);
}

extern "C" int bitz_gkr_nibble_rows_dev(bitz_ctx* ctx, const void* packed, int t, int s, void* nr) {
    PRELUDE();
    BITZ_REQUIRE(packed && nr, BITZ_ERR_ARG, "bitz_gkr_nibble_rows_dev: null argument");
    BITZ_REQUIRE(t >= 4 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_nibble_rows_dev: bad shape");
    return bitz::gkr_nibble_rows(ctx, (const uint64_t*)packed, t, s, nr);
}

static void chal_from_bytes(const uint8_t* b, int n, bitz::hgf* out) {
    for (int i = 0; i < n; i++) {
        memcpy(&out[i].lo, b + 16 * i, 8);
        memcpy(&out[i].hi, b + 16 * i + 8, 8);
    }
}

extern "C" int bitz_gkr_nibble_round_sums_dev(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s,
                                              const uint8_t* chal, const void* eq_c, const void* eq_y, int send_one,
                                              uint8_t out_host[32]) {
    PRELUDE();
    BITZ_REQUIRE(level >= 0 && level <= 2, BITZ_ERR_ARG, "bitz_gkr_nibble_round_sums_dev: level must be 0, 1 or 2");
    BITZ_REQUIRE(nr && A && eq_c && eq_y && out_host && (level == 2 || chal), BITZ_ERR_ARG,
                 "bitz_gkr_nibble_round_sums_dev: null argument");
    BITZ_REQUIRE(t >= 4 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_nibble_round_sums_dev: bad shape");
    bitz::hgf ch[3] = {};
    chal_from_bytes(chal, 2 - level, ch);
    CALL(bitz::gkr_nibble_round_sums(ctx, level, nr, A, t, s, ch, nullptr, eq_c, eq_y, send_one, nullptr));
    return bitz::read_result(ctx, 2, out_host);
}

extern "C" int bitz_gkr_nibble_fold_dev(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s,
                                        const uint8_t* chal, const void* eq_hi, void* l_out, void* r_out) {
    PRELUDE();
    BITZ_REQUIRE(level >= 0 && level <= 2, BITZ_ERR_ARG, "bitz_gkr_nibble_fold_dev: level must be 0, 1 or 2");
    BITZ_REQUIRE(nr && A && chal && l_out && r_out, BITZ_ERR_ARG, "bitz_gkr_nibble_fold_dev: null argument");
    BITZ_REQUIRE(t >= 4 && s >= 0 && t + s <= 40 && (!eq_hi || s > 8), BITZ_ERR_SHAPE,
                 "bitz_gkr_nibble_fold_dev: bad shape (eq_hi needs s > 8)");
    bitz::hgf ch[3] = {};
    chal_from_bytes(chal, 3 - level, ch);
    return bitz::gkr_nibble_fold(ctx, level, nr, A, t, s, ch, nullptr, eq_hi, l_out, r_out);
}

extern "C" int bitz_gkr_jit_round_sums_dev(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s,
                                           const uint8_t* chal, const void* eq_c, const void* eq_y, int send_one,
                                           uint8_t out_host[32]) {
    PRELUDE();
    BITZ_REQUIRE(level >= 0 && level <= 3, BITZ_ERR_ARG, "bitz_gkr_jit_round_sums_dev: level must be 0..3");
    BITZ_REQUIRE(nr && A && eq_c && eq_y && out_host && (level == 3 || chal), BITZ_ERR_ARG,
                 "bitz_gkr_jit_round_sums_dev: null argument");
    BITZ_REQUIRE(t >= 5 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_jit_round_sums_dev: bad shape");
    bitz::hgf ch[4] = {};
    if (level < 3) chal_from_bytes(chal, 3 - level, ch);
    CALL(bitz::gkr_jit_tables(ctx, level, A, t, ch, nullptr));
    CALL(bitz::gkr_jit_round_sums(ctx, level, nr, t, s, eq_c, eq_y, send_one, nullptr));
    return bitz::read_result(ctx, 2, out_host);
}

extern "C" int bitz_gkr_jit_fold_dev(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s,
                                     const uint8_t* chal, const void* eq_hi, void* l_out, void* r_out) {
    PRELUDE();
    BITZ_REQUIRE(level >= 0 && level <= 3, BITZ_ERR_ARG, "bitz_gkr_jit_fold_dev: level must be 0..3");
    BITZ_REQUIRE(nr && A && chal && l_out && r_out, BITZ_ERR_ARG, "bitz_gkr_jit_fold_dev: null argument");
    BITZ_REQUIRE(t >= 5 && s >= 0 && t + s <= 40 && (!eq_hi || s > 8), BITZ_ERR_SHAPE,
                 "bitz_gkr_jit_fold_dev: bad shape (eq_hi needs s > 8)");
    bitz::hgf ch[4] = {};
    chal_from_bytes(chal, 4 - level, ch);
    CALL(bitz::gkr_jit_tables(ctx, level, A, t, ch, nullptr));
    return bitz::gkr_jit_fold(ctx, level, nr, t, s, ch, nullptr, eq_hi, l_out, r_out);
}

extern "C" int bitz_gkr_level4_dev(bitz_ctx* ctx, const void* nr, const void* A, int t, int s, void* out) {
    PRELUDE();
    BITZ_REQUIRE(nr && A && out, BITZ_ERR_ARG, "bitz_gkr_level4_dev: null argument");
    BITZ_REQUIRE(t >= 5 && s >= 0 && t + s <= 40, BITZ_ERR_SHAPE, "bitz_gkr_level4_dev: bad shape");
    return bitz::gkr_level4(ctx, nr, A, t, s, out);
}

extern "C" int bitz_gkr_level1_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s, void* out) {
    PRELUDE();
    BITZ_REQUIRE(packed && A && out, BITZ_ERR_ARG, "bitz_gkr_level1_dev: null argument");
    return bitz::gkr_level1(ctx, (const uint64_t*)packed, A, t, s, out);
}

extern "C" int bitz_product_level_dev(bitz_ctx* ctx, const void* in, uint64_t n, void* out) {
    PRELUDE();
    BITZ_REQUIRE(in && out && n >= 2 && n % 2 == 0, BITZ_ERR_ARG, "bitz_product_level_dev: bad argument");
    return bitz::product_level(ctx, in, n, out);
}

extern "C" int bitz_xi_combined_dev(bitz_ctx* ctx, const void* packed, int t, int s, const void* w2, void* out) {
    PRELUDE();
    BITZ_REQUIRE(packed && w2 && out, BITZ_ERR_ARG, "bitz_xi_combined_dev: null argument");
    return bitz::xi_combined(ctx, (const uint64_t*)packed, t, s, w2, out);
}

extern "C" int bitz_fold_rows_eq_dev(bitz_ctx* ctx, const void* rows, int t, int s, const void* eq, void* out) {
    PRELUDE();
    BITZ_REQUIRE(rows && eq && out, BITZ_ERR_ARG, "bitz_fold_rows_eq_dev: null argument");
    BITZ_REQUIRE(t >= 6, BITZ_ERR_SHAPE, "bitz_fold_rows_eq_dev: t must be >= 6");
    return bitz::fold_rows_eq(ctx, (const uint64_t*)rows, t, s, eq, out);
}

extern "C" int bitz_round_poly_dev(bitz_ctx* ctx, const void* m, const void* w, uint64_t n, uint8_t out_host[48]) {
    PRELUDE();
    BITZ_REQUIRE(m && w && out_host && n >= 2 && n % 2 == 0, BITZ_ERR_ARG, "bitz_round_poly_dev: bad argument");
    CALL(bitz::round_poly(ctx, m, w, n, nullptr));
    return bitz::read_result(ctx, 3, out_host);
}

extern "C" int bitz_fold_lsb_dev(bitz_ctx* ctx, const void* v, uint64_t n, const uint8_t r[16], void* out) {
    PRELUDE();
    BITZ_REQUIRE(v && r && out && n >= 2 && n % 2 == 0, BITZ_ERR_ARG, "bitz_fold_lsb_dev: bad argument");
    return bitz::fold_lsb(ctx, v, n, r, out);
}

extern "C" int bitz_fold_1b_rows_dev(bitz_ctx* ctx, const void* p, const void* suffix, uint64_t n, void* out) {
    PRELUDE();
    BITZ_REQUIRE(p && suffix && out, BITZ_ERR_ARG, "bitz_fold_1b_rows_dev: null argument");
    return bitz::fold_1b_rows(ctx, p, suffix, n, out);
}

extern "C" int bitz_fold_b128_dev(bitz_ctx* ctx, const void* suffix, uint64_t n, const void* w, void* out) {
    PRELUDE();
    BITZ_REQUIRE(suffix && w && out, BITZ_ERR_ARG, "bitz_fold_b128_dev: null argument");
    return bitz::fold_b128(ctx, suffix, n, w, out);
}

extern "C" int bitz_axpy_dev(bitz_ctx* ctx, void* acc, const void* e, uint64_t n, const uint8_t eta[16]) {
    PRELUDE();
    BITZ_REQUIRE(acc && e && eta, BITZ_ERR_ARG, "bitz_axpy_dev: null argument");
    return bitz::axpy(ctx, acc, e, n, eta);
}

extern "C" int bitz_lig_msg_dev(bitz_ctx* ctx, const void* f, const void* b, uint64_t n, int with_eval,
                                uint8_t out_host[48]) {
    PRELUDE();
    BITZ_REQUIRE(f && b && out_host && n >= 2 && n % 2 == 0, BITZ_ERR_ARG, "bitz_lig_msg_dev: bad argument");
    CALL(bitz::lig_msg(ctx, f, b, n, with_eval, nullptr));
    return bitz::read_result(ctx, with_eval ? 3 : 2, out_host);
}

extern "C" int bitz_lig_fold_dev(bitz_ctx* ctx, const void* f, const void* b, uint64_t n, const uint8_t r[16],
                                 void* fo, void* bo) {
    PRELUDE();
    BITZ_REQUIRE(f && b && r && fo && bo && n >= 2 && n % 2 == 0, BITZ_ERR_ARG, "bitz_lig_fold_dev: bad argument");
    return bitz::lig_fold(ctx, f, b, n, r, fo, bo);
}

extern "C" int bitz_induce_basis_dev(bitz_ctx* ctx, const uint64_t* queries_host, uint32_t nq, int log_n,
                                     const void* alpha_w, void* out) {
    PRELUDE();
    BITZ_REQUIRE(out && (nq == 0 || (queries_host && alpha_w)), BITZ_ERR_ARG, "bitz_induce_basis_dev: null argument");
    BITZ_REQUIRE(log_n >= 0 && log_n <= 30, BITZ_ERR_SHAPE, "bitz_induce_basis_dev: log_n out of range");
    // s_k(v_k) and inverses on the host (eval_sk_at_vks, pcs/ligerito.rs:1720-1745)
    std::vector<bitz::hgf> sks(log_n + 1, bitz::hgf{0, 0}), inv(log_n + 1);
    sks[0] = bitz::hgf{1, 0};
    if (log_n > 0) {
        std::vector<bitz::hgf> layer(log_n);
        for (int i = 1; i <= log_n; i++) layer[i - 1] = bitz::hgf{i < 64 ? 1ull << i : 0, i >= 64 ? 1ull << (i - 64) : 0};
        int cur = log_n;
        for (int i = 0; i < log_n; i++) {
            for (int j = 0; j < cur; j++) {
                bitz::hgf s = layer[j];
                bitz::hgf v = bitz::hgf_add(bitz::hgf_mul(s, s), bitz::hgf_mul(sks[i], s));
                if (j == 0) sks[i + 1] = v; else layer[j - 1] = v;
            }
            cur--;
        }
    }
    for (int k = 0; k <= log_n; k++) inv[k] = bitz::hgf_inv(sks[k]);
    size_t tb = (size_t)(log_n + 1) * 16;
    std::vector<uint8_t> host(2 * tb + (size_t)nq * 8);
    memcpy(host.data(), sks.data(), tb);
    memcpy(host.data() + tb, inv.data(), tb);
    if (nq) memcpy(host.data() + 2 * tb, queries_host, (size_t)nq * 8);
    void* dev;
    CALL(stage(ctx, host.data(), host.size(), &dev));
    const uint8_t* d = (const uint8_t*)dev;
    int rc = bitz::induce_basis(ctx, (const uint64_t*)(d + 2 * tb), nq, log_n, d, d + tb, alpha_w, out);
    if (rc) return rc;
    BITZ_HIP(hipStreamSynchronize(ctx->stream));  // host staging vector dies here
    return BITZ_OK;
}

extern "C" int bitz_pow_find_dev(bitz_ctx* ctx, const uint8_t prefix[16], const uint8_t seed[16], uint32_t bits,
                                 uint64_t* nonce) {
    PRELUDE();
    BITZ_REQUIRE(prefix && seed && nonce, BITZ_ERR_ARG, "bitz_pow_find_dev: null argument");
    return bitz::pow_find(ctx, prefix, seed, bits, nonce);
}

extern "C" int bitz_transcript_challenges_dev(bitz_ctx* ctx, const uint8_t* prefix, uint64_t prefix_len,
                                              const uint8_t* msgs, uint32_t n, uint32_t msg_len, uint8_t* out_host) {
    PRELUDE();
    BITZ_REQUIRE((prefix || prefix_len == 0) && (msgs || n == 0) && (out_host || n == 0), BITZ_ERR_ARG,
                 "bitz_transcript_challenges_dev: null argument");
    bitz::Transcript tr;
    tr.absorb_slice(prefix, prefix_len);
    bitz::B3State st;
    tr.export_state(&st);
    bitz::DevBuf dst, dmsg, dout;
    BITZ_REQUIRE(dst.alloc(sizeof(st)) && dmsg.alloc((size_t)n * msg_len + 1) && dout.alloc((size_t)n * 16 + 1),
                 BITZ_ERR_ALLOC, "bitz_transcript_challenges_dev: allocation failed");
    BITZ_HIP(hipMemcpyAsync(dst.ptr, &st, sizeof(st), hipMemcpyHostToDevice, ctx->stream));
    if (n) BITZ_HIP(hipMemcpyAsync(dmsg.ptr, msgs, (size_t)n * msg_len, hipMemcpyHostToDevice, ctx->stream));
    CALL(bitz::transcript_challenges(ctx, dst.ptr, dmsg.ptr, n, msg_len, dout.ptr));
    if (n) BITZ_HIP(hipMemcpyAsync(out_host, dout.ptr, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream));
    BITZ_HIP(hipStreamSynchronize(ctx->stream));
    return BITZ_OK;
}
