// Internal (C++) entry points of the opening kernels, shared by the C ABI
// wrappers and the prover orchestration. All pointers are device pointers
// unless named *_host; work is enqueued on ctx->stream.
#pragma once
#include "common.h"

namespace bitz {

// Sums produced by the reduction kernels land in ctx->scratch.result (or in
// out_dev when non-null); read_result copies K field elements to the host.
int read_result(bitz_ctx* ctx, int K, uint8_t* host_out);
// bytes of host data to device memory dev through the pinned upload ring
// (asynchronous: host may be reused as soon as this returns)
int upload_async(bitz_ctx* ctx, const void* host, size_t bytes, void* dev);

int eq_table(bitz_ctx* ctx, const void* r_dev, int n, void* out);
// Tables over host coordinates passed as a kernel argument (no upload, no
// synchronisation): out[x] = prod_i (bit i of x ? b[i] : a[i]), and the eq
// table (a = 1 + r, b = r), for n <= kMaxTensorN.
constexpr int kMaxTensorN = 32;
struct TensorArgs {
    uint4 a[kMaxTensorN];
    uint4 b[kMaxTensorN];
};
int tensor_args(bitz_ctx* ctx, const hgf* a, const hgf* b, int n, void* out);
// Up to kTensorJobs such tables of at most kTensorJobN coordinates each in
// one launch (their blocks side by side): table j over a[j], b[j] (n[j]
// coordinates) into out[j].
constexpr int kTensorJobs = 4;
constexpr int kTensorJobN = 14;
int tensor_args_multi(bitz_ctx* ctx, int njobs, const hgf* const* a, const hgf* const* b, const int* n,
                      void* const* out);
// BLAKE3 chaining values of m whole chunks of a transcript stream (chunk i at
// data + 1024 i, counter c0 + i; data 4-byte aligned) merged into the
// subtrees of the runs ks (blake3_subtree_runs(c0, m)): run r's root lands
// at roots + 32 r. cvs: scratch of 32 m bytes.
int blake3_runs(bitz_ctx* ctx, const void* data, uint64_t m, uint64_t c0, const std::vector<int>& ks, void* cvs,
                void* roots);
int eq_args(bitz_ctx* ctx, const hgf* r, int n, void* out);
int fold_columns(bitz_ctx* ctx, const uint64_t* rows, int t, int s, const void* w, void* out);
int pow_images(bitz_ctx* ctx, const uint8_t g[16], const void* exps, uint64_t n, void* out);
int pack_columns(bitz_ctx* ctx, const uint64_t* rows, int t, int s, uint64_t* packed);
int gkr_round_sums(bitz_ctx* ctx, const void* l, const void* r, uint64_t h, const void* eq_c, int s_c,
                   const void* eq_y, int send_one, void* out_dev);
// the previous round's fold by rho (halves of 2h entries -> h) fused with
// this round's sums over the folded halves. eq_lo (2^lo_bits entries, with
// eq_c null): the right half carries the column weights of the index bits
// lo_bits..s-1 only, and the sums take eq_lo[x & (2^lo_bits - 1)] as well
// (lo_bits <= 8)
int gkr_fold_round_sums(bitz_ctx* ctx, void* l, void* r, uint64_t h, const uint8_t rho[16], const void* eq_c,
                        int s_c, const void* eq_y, int send_one, void* out_dev, const void* rho_dev = nullptr,
                        const void* eq_lo = nullptr, int lo_bits = 0);
// (with s_c == s from K-bit row patterns: off the row words nr (t >= 4; built
// into a scratch buffer when nr is null), the gather kernel below t = 4; the
// same for gkr_leaf_round1_sums and gkr_level1_round0_sums)
int gkr_leaf_round_sums(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const void* eq_c,
                        int s_c, const void* eq_y, int send_one, void* out_dev, const void* nr = nullptr);
int fold_halves(bitz_ctx* ctx, void* v, uint64_t h, const uint8_t rho[16]);
// both halves l and r (2h entries each) folded by rho in one launch
// (rho_dev: the challenge in device memory, from the device transcript; rho unused then)
int fold_pair(bitz_ctx* ctx, void* l, void* r, uint64_t h, const uint8_t rho[16], const void* rho_dev = nullptr);
int gkr_leaf_fold(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const uint8_t rho[16],
                  void* l_out, void* r_out);
// round 1 of the leaf layer (after round 0 folded by rho0) from 8-bit row
// patterns; weights eq_y[y] * eq_c[c] over all s column coordinates
int gkr_leaf_round1_sums(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const uint8_t rho0[16],
                         const void* eq_c, const void* eq_y, int send_one, void* out_dev,
                         const void* rho0_dev = nullptr, const void* nr = nullptr);
// the leaf layer folded by rho0 then rho1 into halves of 2^(t+s-3) entries;
// with eq_c the right half is written multiplied by eq_c[c] (c = low s bits),
// with eq_hi (s > 8, eq_c null) by eq_hi[c >> 8]
int gkr_leaf_fold2(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const uint8_t rho0[16],
                   const uint8_t rho1[16], const void* eq_c, void* l_out, void* r_out, const void* rho0_dev = nullptr,
                   const void* rho1_dev = nullptr, const void* eq_hi = nullptr);
// round sums whose right half gets eq_c[x & (2^s-1)] multiplied in (in place);
// weights eq_y[x >> s] only
int gkr_premul_round_sums(bitz_ctx* ctx, const void* l, void* r, uint64_t h, const void* eq_c, int s,
                          const void* eq_y, int send_one, void* out_dev);
// round sums over halves whose right operand already carries eq_c: weights
// eq_y[x >> s]; with eq_lo the right operand carries the weights of column
// bits lo_bits..s-1 only and eq_lo[x & (2^lo_bits - 1)] weighs the slots too
int gkr_round_sums_rowonly(bitz_ctx* ctx, const void* l, const void* r, uint64_t h, int s, const void* eq_y,
                           int send_one, void* out_dev, const void* eq_lo = nullptr, int lo_bits = 0);
// v[x] *= eq_hi[(x & (2^s - 1)) >> 8] for x < n (n a multiple of 2^s, s > 8)
int scale_cols_hi(bitz_ctx* ctx, void* v, uint64_t n, int s, const void* eq_hi);
// out[x] = prod_i (bit i of x ? b[i] : a[i]), x < 2^n, n <= 14
int tensor_table(bitz_ctx* ctx, const void* a, const void* b, int n, void* out);
// the layer over level 1 (never materialised) from 8-bit row patterns:
// round-0 sums (weights eq_y[y] (2^(t-3)) * eq_c[c]) and the round-0 fold
// into halves of 2^(t+s-3) entries (right half times eq_c when given)
int gkr_level1_round0_sums(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const void* eq_c,
                           const void* eq_y, int send_one, void* out_dev, const void* nr = nullptr);
int gkr_level1_fold(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, const uint8_t rho[16],
                    const void* eq_c, void* l_out, void* r_out, const void* rho_dev = nullptr,
                    const void* eq_hi = nullptr);
// levels 2 and 3 of the product tree (2^(t+s-2), 2^(t+s-3) entries) from the bits
int gkr_level2(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, void* out);
int gkr_level3(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, void* out);
int gkr_level1(bitz_ctx* ctx, const uint64_t* packed, const void* A, int t, int s, void* out);
// Low GKR levels off 16-bit row words (forest.hip): nr[(y << s) | c] holds
// the bits of column c at rows y + w 2^(t-4), w < 16 (even w in the low
// byte, odd w in the high byte). The level-l layer (l = 0, 1, 2), after its
// first 2 - l rounds (challenges chal[0..2-l)), takes round 2 - l's sums and
// the fold by chal[2 - l] into halves of 2^(t+s-4) entries (right half
// times eq_hi[c >> 8] when given, s > 8). Challenges from chal_host, or from
// the device array chal_dev when non-null.
int gkr_nibble_rows(bitz_ctx* ctx, const uint64_t* packed, int t, int s, void* nr);
int gkr_nibble_round_sums(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s, const hgf* chal_host,
                          const void* chal_dev, const void* eq_c, const void* eq_y, int send_one, void* out_dev);
int gkr_nibble_fold(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s, const hgf* chal_host,
                    const void* chal_dev, const void* eq_hi, void* l_out, void* r_out);
// The table round after a low layer's row-word round (forest.hip;
// Forest::jit_round_sums / jit_fold_round_with, src/bitz/forest.rs:990-1176):
// gkr_jit_tables builds the per-row tables F of level 0-2 from the layer's
// challenges chal[0..3-level) (the row-word round's fold tables), or level
// 3's leaf product tables (no challenges), in ctx->scratch.nib_tab;
// gkr_jit_round_sums takes round 3 - level's sums (weights eq_y[y'],
// y' < 2^(t-5), times eq_c[c]); gkr_jit_fold pre-scales the tables by
// chal[3 - level] and folds into halves of 2^(t+s-5) entries (right half
// times eq_hi[c >> 8] when given, s > 8). gkr_level4 writes level 4
// (2^(t+s-4) entries) from level 3's tables.
int gkr_jit_tables(bitz_ctx* ctx, int level, const void* A, int t, const hgf* chal_host, const void* chal_dev);
int gkr_jit_round_sums(bitz_ctx* ctx, int level, const void* nr, int t, int s, const void* eq_c, const void* eq_y,
                       int send_one, void* out_dev);
int gkr_jit_fold(bitz_ctx* ctx, int level, const void* nr, int t, int s, const hgf* chal_host, const void* chal_dev,
                 const void* eq_hi, void* l_out, void* r_out);
int gkr_level4(bitz_ctx* ctx, const void* nr, const void* A, int t, int s, void* out);
// Reduces nblocks x K block partials (ctx->scratch.partials) to out_dev, or
// to the mailbox for the host when out_dev is null (deferred when the
// context defers sums to the device transcript).
int sums_finish(bitz_ctx* ctx, unsigned nblocks, int K, void* out_dev);
// l[0..n) and r[0..n) (device) into pinned host memory by one kernel that
// announces them through the mailbox flag; *host_out points at l's copy,
// r's follows (valid until the next call)
int download_pair_mapped(bitz_ctx* ctx, const void* l, const void* r, uint64_t n, void** host_out);
int product_level(bitz_ctx* ctx, const void* in, uint64_t n, void* out);
int xi_combined(bitz_ctx* ctx, const uint64_t* packed, int t, int s, const void* w2, void* out);
int fold_rows_eq(bitz_ctx* ctx, const uint64_t* rows, int t, int s, const void* eq, void* out);
int round_poly(bitz_ctx* ctx, const void* m, const void* w, uint64_t n, void* out_dev);
int fold_lsb(bitz_ctx* ctx, const void* v, uint64_t n, const uint8_t r[16], void* out);
int fold_1b_rows(bitz_ctx* ctx, const void* p, const void* suffix, uint64_t n, void* out_dev);
// The ring-switch bit fold with its block partials in a caller-given device
// buffer of fold_1b_rows_chunks(n) * 128 * 16 bytes (at most
// kFold1bMaxChunks chunks); allocates nothing.
constexpr unsigned kFold1bMaxChunks = 512;
unsigned fold_1b_rows_chunks(uint64_t n);
int fold_1b_rows_into(bitz_ctx* ctx, const void* p, const void* suffix, uint64_t n, void* out_dev, void* partials);
// Resolves the no-allocation entry point's device array on `device` (current).
int noalloc_setup(int device);
int reduce_columns(bitz_ctx* ctx, const uint4* partials, uint32_t nblocks, uint32_t width, uint4* out);
int fold_b128(bitz_ctx* ctx, const void* suffix, uint64_t n, const void* w, void* out);
int axpy(bitz_ctx* ctx, void* acc, const void* e, uint64_t n, const uint8_t eta[16]);
int lig_msg(bitz_ctx* ctx, const void* f, const void* b, uint64_t n, int with_eval, void* out_dev);
int lig_fold(bitz_ctx* ctx, const void* f, const void* b, uint64_t n, const uint8_t r[16], void* fo, void* bo);
int induce_basis(bitz_ctx* ctx, const uint64_t* queries_dev, uint32_t nq, int log_n, const void* sks,
                 const void* inv_sks, const void* aw, void* out);
int pow_find(bitz_ctx* ctx, const uint8_t prefix[16], const uint8_t seed[16], uint32_t bits, uint64_t* nonce);
int gather_rows(bitz_ctx* ctx, const void* cw, const uint64_t* q_dev, uint32_t nq, uint32_t lanes, void* out);
int gather_nodes(bitz_ctx* ctx, const void* tree, const uint64_t* idx_dev, uint32_t n, void* out);

int dot(bitz_ctx* ctx, const void* a, const void* b, uint64_t n, void* out_dev);
int mul_elem(bitz_ctx* ctx, const void* a, const void* b, uint64_t n, void* out, int plus_one);
int mul_inplace(bitz_ctx* ctx, void* v, const void* b, uint64_t n);
int scale(bitz_ctx* ctx, const void* a, uint64_t n, const uint8_t c[16], void* out);
// out[x] = bit x of a flat bit stream (2^k bits as Gf128 0 or 1)
int bits_to_gf(bitz_ctx* ctx, const uint64_t* bits, uint64_t n, void* out);
// Inner-product sumcheck round J (0..2) over a single column of bits: the
// round's table has n entries (the column after J folds); for J > 0 the
// weights w (2n entries) are first folded by r into w_out (n entries). The
// 3 * 2^J XOR accumulators acc[s 2^J + c] go to the mailbox (read_result):
// round sum s is sum_c e_c acc[s 2^J + c], e the eq table of the J earlier
// challenges (s: m0 w0, m1 w1, (m0 + m1)(w0 + w1)).
int bits_round(bitz_ctx* ctx, int J, const uint64_t* bits, const void* w, uint64_t n, const uint8_t r[16],
               void* w_out);
// The table after three folds, A_3[k] = sum_{c < 8} e8[c] bit(8k + c) (e8:
// the eq table of the three challenges, 8 elements), and the weights folded
// by r from 2 n_out to n_out entries.
int bits_materialize(bitz_ctx* ctx, const uint64_t* bits, const void* w, uint64_t n_out, const uint8_t r[16],
                     const uint8_t e8[128], void* a_out, void* w_out);

// commit pipeline (commit.hip)
// One GKR round's Fiat-Shamir step on the device, after a sums launch whose
// block partials (nb blocks x 2) are in the scratch partials: the sums are
// reduced, scaled by factor * corr[round], absorbed as the round message
// (msgs[2 round..]), and the challenge is squeezed into rho[round]; factor
// then takes eq_factor(rho, internal[round]).
int gkr_fs_round(bitz_ctx* ctx, unsigned nb, void* state, void* factor, const void* internal, const void* corr,
                 void* rho, void* msgs, uint32_t round);
// Device transcript check: n x (absorb_slice(msg_i), 16-byte challenge) on a
// device copy of `state`; challenges to out (device, n x 16 B).
int transcript_challenges(bitz_ctx* ctx, void* state, const void* msgs, uint32_t n, uint32_t msg_len, void* out);
// the device twiddle table of an NTT domain of dimension ell (cached in ctx)
int ensure_twiddles(bitz_ctx* ctx, int ell, const uint4** out);
int commit_device(bitz_ctx* ctx, const void* packed, int log_msg_len, int log_batch, int log_inv_rate,
                  void* codeword, void* tree);

}  // namespace bitz
