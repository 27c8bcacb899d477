/*
 * bitz_hip.h -- C ABI of the MI355X (gfx950) BitZ prover hot path.
 *
 * Plain pointers and sizes only. Every entry point returns an int status:
 * BITZ_OK (0) or a negative BITZ_ERR_* code; bitz_last_error() returns a
 * thread-local message for the last failure. No exception or panic crosses
 * this boundary. A bitz_ctx binds one device and one HIP stream; contexts and
 * handles are not thread-safe (one thread per context).
 *
 * Device-pointer entry points ("_dev") take device memory the caller owns
 * (hipMalloc, or a torch tensor's data_ptr()) and enqueue work on the
 * context's stream; they return after enqueueing unless documented as
 * synchronizing. Gf128 buffers are arrays of 16-byte elements in the
 * reference's wire form (lo_le || hi_le, src/bitz/codec.rs:4-21).
 *
 * Reference interfaces each entry point replaces (albert-garreta/BitZ-PoC):
 *   bitz_ntt_forward_dev  flock_core AdditiveNttF128::forward_transform_interleaved_from_layer
 *                         (vendor/flock-mod/crates/flock-core/src/ntt/additive_ntt_f128.rs:205)
 *   bitz_merkle_tree_dev  flock_core merkle::merkle_tree (merkle.rs:550) under HashKind::Blake3
 *   bitz_commit_dev       flock_core pcs::commit::commit (pcs/commit.rs:180), Blake3 Merkle
 *   bitz_commit           bitz::Pcs::commit (src/bitz/pcs.rs:270-284) ->
 *                         ligerito_flock::commit_rs_ligerito_rows (src/ligerito_flock.rs:270)
 */
#ifndef BITZ_HIP_H
#define BITZ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BITZ_OK 0
#define BITZ_ERR_SHAPE (-1)  /* geometry rejected (reference: assert/ShapeError) */
#define BITZ_ERR_ALLOC (-2)  /* device allocation failed */
#define BITZ_ERR_DEVICE (-3) /* HIP runtime error */
#define BITZ_ERR_ARG (-4)    /* null or inconsistent argument */
#define BITZ_ERR_VERIFY (-5) /* a proof rejected by the verifier (reason in bitz_last_error) */

typedef struct bitz_ctx bitz_ctx;
typedef struct bitz_commit_handle bitz_commit_handle;

/* ---- context ---- */
int bitz_ctx_create(int device, bitz_ctx** out);
void bitz_ctx_destroy(bitz_ctx* ctx);
/* Enqueue all further work on `hip_stream` (e.g. torch.cuda.current_stream().cuda_stream);
 * NULL selects the device's default (null) stream, which torch uses unless told otherwise.
 * A new context starts on a stream of its own. */
int bitz_ctx_set_stream(bitz_ctx* ctx, void* hip_stream);
void* bitz_ctx_stream(bitz_ctx* ctx);
int bitz_ctx_synchronize(bitz_ctx* ctx);
/* Device memory of a context is served by a caching pool (blocks freed by
 * one call are reused by the next, in stream order; cf. flock's codeword
 * scratch pool, vendor/flock-mod/crates/flock-core/src/pcs/commit.rs:157-163).
 * bitz_ctx_trim drains the stream and returns every cached block to HIP;
 * bitz_ctx_cached_bytes reports the bytes currently cached. Changing the
 * stream with bitz_ctx_set_stream drains the old stream first. */
int bitz_ctx_trim(bitz_ctx* ctx);
/* GKR rounds take their Fiat-Shamir challenges from a device copy of the
 * transcript (enable 1): a layer's rounds are enqueued without host round
 * trips and the host replays the messages afterwards, checking every
 * challenge. 0 (the default) runs each round through the host transcript.
 * Proof bytes are the same either way. */
int bitz_ctx_set_device_transcript(bitz_ctx* ctx, int enable);
uint64_t bitz_ctx_cached_bytes(bitz_ctx* ctx);
/* Device buffers from the context's pool, for callers that have no HIP
 * binding of their own (the reference keeps its tables in Vec<Gf>; these move
 * them to and from the device): alloc, free (returns the block to the pool),
 * and synchronous upload / download on the context's stream. */
int bitz_dev_alloc(bitz_ctx* ctx, uint64_t bytes, void** out);
int bitz_dev_free(bitz_ctx* ctx, void* ptr);
int bitz_dev_upload(bitz_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int bitz_dev_download(bitz_ctx* ctx, void* dst, const void* src, uint64_t bytes);
const char* bitz_last_error(void);
/* Per-kernel timing: with enable != 0, instrumented launches are bracketed by
 * HIP events on the context's stream (stats reset). bitz_ctx_kernel_stats
 * writes newline-separated kernel names and, per kernel, 4 doubles:
 * launches, total ms, algorithmic bytes, algorithmic 32-bit integer ops;
 * returns the number of kernels. */
int bitz_ctx_set_timing(bitz_ctx* ctx, int enable);
/* Timing restricted to the launches of one kernel name (as reported by
 * bitz_ctx_kernel_stats); host phases are not recorded. Stats reset. The
 * name "phases" times no launch and records the host phases only (wall
 * clock between phase marks, without the events' cost). A measurement aid
 * of this library, with no counterpart in the reference. */
int bitz_ctx_set_timing_only(bitz_ctx* ctx, const char* kernel);
int bitz_ctx_kernel_stats(bitz_ctx* ctx, char* names, uint64_t names_cap, double* values, uint64_t values_cap);
/* Launch-shape choices of the context (none changes a result; the reference
 * has no counterpart, its CPU loops have no launch shapes): slots per thread
 * of the round-sum kernels (round_r_max, a power of two <= 32), the chunk
 * count below which that halves (round_min_chunks), blocks per round-sum
 * launch (round_block_cap), the largest sumcheck round size 2^host_tail_log
 * run on the host (4..14) and butterfly layers per global NTT pass
 * (ntt_radix_log, 1..4). set_tune validates and overrides; get_tune writes
 * the five values and 1 if they came from bitz_ctx_autotune. */
int bitz_ctx_set_tune(bitz_ctx* ctx, uint32_t round_r_max, uint32_t round_min_chunks, uint32_t round_block_cap,
                      uint32_t host_tail_log, uint32_t ntt_radix_log);
int bitz_ctx_get_tune(bitz_ctx* ctx, uint32_t out[6]);
/* Chooses the five values above for jobs of shape (t, s, log_batch,
 * log_inv_rate) by timing candidates on the device over synthetic data of
 * that shape (round kernels at the layer's round sizes, host against device
 * rounds, the commit at both NTT radices); tune_ms (optional) gets the
 * wall-clock time it took. Synchronizing; allocates about the job's
 * footprint while it runs. */
int bitz_ctx_autotune(bitz_ctx* ctx, uint32_t t, uint32_t s, uint32_t log_batch, uint32_t log_inv_rate,
                      double* tune_ms);
/* Library build tag (architecture and version). */
const char* bitz_version(void);

/* ---- arithmetic parity helpers (device pointers) ---- */
/* out[i] = a[i] * b[i] in GF(2^128) (variable x variable path) */
int bitz_gf_mul_dev(bitz_ctx* ctx, const void* a, const void* b, void* out, uint64_t n);
/* out[i] = a[i] * c in GF(2^128) through the LDS window table of c */
int bitz_gf_mul_const_dev(bitz_ctx* ctx, const void* a, const uint8_t c[16], void* out,
                          uint64_t n);
/* BLAKE3 non-root leaf chaining values of n equal leaves of leaf_size bytes
 * (leaf_size <= 1024): out = n * 32 bytes */
int bitz_blake3_leaves_dev(bitz_ctx* ctx, const void* data, uint64_t n, uint64_t leaf_size,
                           void* out);

/* ---- Reed-Solomon encode / Merkle / commit (device pointers) ---- */
/* In-place interleaved forward additive NTT over GF(2^128), LCH novel basis
 * with the standard subspace basis {1, X, ..., X^(ell-1)}:
 * data[pos * 2^log_lanes + lane], 2^log_d positions, layers start_layer..log_d-1.
 * Requires log_d <= ell <= 64, start_layer <= log_d, log_lanes <= 6. */
int bitz_ntt_forward_dev(bitz_ctx* ctx, void* data, int log_d, int log_lanes, int start_layer,
                         int ell);
/* Flat BLAKE3 Merkle tree (merkle.rs layout: leaves first, root last,
 * 2*num_leaves-1 hashes of 32 B). num_leaves a power of two, leaf_size <= 1024. */
int bitz_merkle_tree_dev(bitz_ctx* ctx, const void* data, uint64_t num_leaves,
                         uint64_t leaf_size, void* tree);
/* PCS commit of a packed message (2^log_msg_len Gf128): replicate into
 * 2^log_inv_rate copies, NTT from layer log_inv_rate, BLAKE3 tree over
 * positions (leaf = 2^log_batch Gf128). codeword: 2^(log_msg_len+log_inv_rate)
 * Gf128; tree: (2*2^k_code-1)*32 B with k_code = log_msg_len-log_batch+log_inv_rate.
 * root (host, 32 B) may be NULL; when given the call synchronizes the stream. */
int bitz_commit_dev(bitz_ctx* ctx, const void* packed, int log_msg_len, int log_batch,
                    int log_inv_rate, void* codeword, void* tree, uint8_t* root);

/* ---- opening-side kernels (device pointers) ----
 * Each replaces the reference function named beside it; field-element
 * buffers are Gf128 arrays, "rows" is the column-major bit matrix (column c
 * = 2^(t-6) u64 words), "packed" the row-major packing (packed[g*2^t+b] =
 * bit b of columns 64g..64g+63). Results written to out_host are copied to
 * host memory and the call synchronizes. */
/* build_eq_x_r_vec (src/poly/utils.rs:46): out[x] = prod_i (x_i ? r_i : 1+r_i), n <= 28 */
int bitz_eq_table_dev(bitz_ctx* ctx, const uint8_t* r_host, int n, void* out);
/* fold_columns (src/bitz/fold.rs:80-170): out[c] = sum_b w[b] bit(c,b), wrapping u128 */
int bitz_fold_columns_dev(bitz_ctx* ctx, const void* rows, int t, int s, const void* w, void* out);
/* column_images/row_images via FixedBasePow (fold.rs:173-182, preparation.rs:177-263):
 * out[i] = g^exps[i], exps are u128 */
int bitz_pow_images_dev(bitz_ctx* ctx, const uint8_t g[16], const void* exps, uint64_t n, void* out);
/* pack_columns_from_rows (src/ligerito.rs:655-690) */
int bitz_pack_columns_dev(bitz_ctx* ctx, const void* rows, int t, int s, void* packed);
/* kernels::round_sums (src/bitz/kernels.rs:80-110) over materialised halves l, r
 * (2h entries each), weights eq_y[x >> s_c] * eq_c[x & (2^s_c-1)];
 * out_host = [sum eq E_end O_end, sum eq (E_hi-E_lo)(O_hi-O_lo)] */
int bitz_gkr_round_sums_dev(bitz_ctx* ctx, const void* l, const void* r, uint64_t h, const void* eq_c,
                            int s_c, const void* eq_y, int send_one, uint8_t out_host[32]);
/* the same sums on the leaf layer, leaves read from the bits and row images
 * (Forest::bit_rounds, src/bitz/forest.rs:381-423) */
int bitz_gkr_leaf_round_sums_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                                 const void* eq_c, int s_c, const void* eq_y, int send_one,
                                 uint8_t out_host[32]);
/* round 1 of the leaf layer after round 0 folded by rho0, from 8-bit row
 * patterns (Forest::one_pass_bit_rounds, src/bitz/forest.rs:424-582); weights
 * eq_y[y] (2^(t-3) entries) * eq_c[c] (2^s entries) */
int bitz_gkr_leaf_round1_sums_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                                  const uint8_t rho0[16], const void* eq_c, const void* eq_y, int send_one,
                                  uint8_t out_host[32]);
/* the leaf layer folded by rho0 and then rho1 into halves of 2^(t+s-3) entries;
 * a non-null eq_c (2^s entries) is multiplied into the right half
 * (jit_fold_round_with carrying the weights, src/bitz/forest.rs:635-650) */
int bitz_gkr_leaf_fold2_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s, const uint8_t rho0[16],
                            const uint8_t rho1[16], const void* eq_c, void* l_out, void* r_out);
/* Round sums with the column weights carried by the right half (Left::Weigh /
 * Left::Weighted, src/bitz/kernels.rs:114-131; the reference carries them in
 * the left half, the sums are the same). premul: r[x], r[x+h] are replaced
 * in place by eq_c[x & (2^s-1)] * r[.] and the sums use weights eq_y[x >> s]
 * over the new values; rowonly: r already carries eq_c. */
int bitz_gkr_premul_round_sums_dev(bitz_ctx* ctx, const void* l, void* r, uint64_t h, const void* eq_c, int s,
                                   const void* eq_y, int send_one, uint8_t out_host[32]);
int bitz_gkr_round_sums_rowonly_dev(bitz_ctx* ctx, const void* l, const void* r, uint64_t h, int s,
                                    const void* eq_y, int send_one, uint8_t out_host[32]);
/* out[x] = prod_i (bit i of x ? b[i] : a[i]), x < 2^n (n <= 14): the inverse
 * column-weight table that removes carried weights (src/bitz/gkr.rs:56-80) */
int bitz_tensor_table_dev(bitz_ctx* ctx, const void* a, const void* b, int n, void* out);
/* fused_fold_round (src/bitz/kernels.rs:146-200): l and r hold the previous
 * round's halves (4h entries each); both are folded by rho in place into 2h
 * entries (quarters q0 + rho (q2 - q0), q1 + rho (q3 - q1)) and this round's
 * sums over the folded halves (h slots, weights as above) are returned; with
 * eq_c NULL the right half carries the column weights (weights eq_y[x >> s_c]) */
int bitz_gkr_fold_round_sums_dev(bitz_ctx* ctx, void* l, void* r, uint64_t h, const uint8_t rho[16],
                                 const void* eq_c, int s_c, const void* eq_y, int send_one, uint8_t out_host[32]);
/* fold_halves (src/bitz/gkr.rs:170-177): v[i] += rho (v[i+h] - v[i]), i < h */
int bitz_fold_halves_dev(bitz_ctx* ctx, void* v, uint64_t h, const uint8_t rho[16]);
/* the leaf layer folded once by rho into materialised halves (2^(t+s-2) each) */
int bitz_gkr_leaf_fold_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                           const uint8_t rho[16], void* l_out, void* r_out);
/* level 1 of the batched product tree from the bits (GrandProductCircuit::batched_eval, gkr.rs:93) */
int bitz_gkr_level1_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s, void* out);
/* levels 2 (2^(t+s-2) entries) and 3 (2^(t+s-3)) straight from the bits
 * (Forest::fold_table / product_level_into over bit patterns, src/bitz/forest.rs:741-898) */
int bitz_gkr_level2_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s, void* out);
int bitz_gkr_level3_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s, void* out);
/* the GKR layer whose halves are level 1, level 1 left implicit
 * (Forest::prove_jit_level, src/bitz/forest.rs:583-740): round-0 sums with
 * weights eq_y[y] (2^(t-3) entries) * eq_c[c], and the round-0 fold by rho into
 * halves of 2^(t+s-3) entries (right half times eq_c[c] when eq_c is given) */
int bitz_gkr_level1_round0_sums_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s,
                                    const void* eq_c, const void* eq_y, int send_one, uint8_t out_host[32]);
int bitz_gkr_level1_fold_dev(bitz_ctx* ctx, const void* packed, const void* A, int t, int s, const uint8_t rho[16],
                             const void* eq_c, void* l_out, void* r_out);
/* The low GKR levels off 16-bit row words (Forest::bit_rounds and the
 * table-driven rounds, src/bitz/forest.rs:1-31, 381-740): nr[(y << s) | c]
 * (u16, y < 2^(t-4)) holds the bits of column c at rows y + w 2^(t-4), w < 16,
 * even w in the low byte, odd w in the high byte. The level-l layer
 * (l = 0 leaves, 1, 2), after its first 2 - l rounds (challenges chal, 16 bytes
 * each), takes round 2 - l's sums (weights eq_y[y] (2^(t-4) entries) *
 * eq_c[c]) and the fold by one more challenge into halves of 2^(t+s-4)
 * entries (right half times eq_hi[c >> 8] when given, s > 8). */
int bitz_gkr_nibble_rows_dev(bitz_ctx* ctx, const void* packed, int t, int s, void* nr);
int bitz_gkr_nibble_round_sums_dev(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s,
                                   const uint8_t* chal, const void* eq_c, const void* eq_y, int send_one,
                                   uint8_t out_host[32]);
int bitz_gkr_nibble_fold_dev(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s,
                             const uint8_t* chal, const void* eq_hi, void* l_out, void* r_out);
/* The table round after the row-word round (Forest::prove_jit_level,
 * jit_round_sums, jit_fold_round_with, src/bitz/forest.rs:583-685, 990-1176).
 * The level-l layer (l = 0..2) after rounds 0..2-l (challenges chal[0..3-l)),
 * or the level-3 layer from its start (chal unused), has halves that are
 * per-row 256-entry tables of its slots' row words. bitz_gkr_jit_round_sums_dev
 * takes round 3 - l's sums (weights eq_y[y'] (2^(t-5) entries) * eq_c[c]);
 * bitz_gkr_jit_fold_dev folds by chal[3 - l] into halves of 2^(t+s-5) entries
 * (right half times eq_hi[c >> 8] when given, s > 8). bitz_gkr_level4_dev
 * writes level 4 of the product tree (2^(t+s-4) entries, level_up of level 3,
 * src/bitz/forest.rs:864-898) from the row words. */
int bitz_gkr_jit_round_sums_dev(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s,
                                const uint8_t* chal, const void* eq_c, const void* eq_y, int send_one,
                                uint8_t out_host[32]);
int bitz_gkr_jit_fold_dev(bitz_ctx* ctx, int level, const void* nr, const void* A, int t, int s,
                          const uint8_t* chal, const void* eq_hi, void* l_out, void* r_out);
int bitz_gkr_level4_dev(bitz_ctx* ctx, const void* nr, const void* A, int t, int s, void* out);
/* one product level: out[i] = in[i] * in[i + n/2] (gkr.rs:93-108) */
int bitz_product_level_dev(bitz_ctx* ctx, const void* in, uint64_t n, void* out);
/* xi_combined_rows_packed (src/ligerito.rs:563-630): out[b] = sum_c w2[c] bit(c,b) */
int bitz_xi_combined_dev(bitz_ctx* ctx, const void* packed, int t, int s, const void* w2, void* out);
/* fold_rows_eq (src/bitz/sumcheck.rs:84-131): out[c] = sum_b eq[b] bit(c,b) */
int bitz_fold_rows_eq_dev(bitz_ctx* ctx, const void* rows, int t, int s, const void* eq, void* out);
/* round_polynomial (src/bitz/sumcheck.rs:404-447): out_host = [a0, a0+a1+q, q] partial sums
 * reported as [sum m0 w0, sum m1 w1, sum (m0+m1)(w0+w1)] */
int bitz_round_poly_dev(bitz_ctx* ctx, const void* m, const void* w, uint64_t n, uint8_t out_host[48]);
/* sumcheck fold (sumcheck.rs:454-465): out[i] = v[2i] + r (v[2i] + v[2i+1]); out != v */
int bitz_fold_lsb_dev(bitz_ctx* ctx, const void* v, uint64_t n, const uint8_t r[16], void* out);
/* fold_1b_rows_naive (pcs/ring_switch.rs:1332-1381): out[i] = sum_x bit_i(p[x]) suffix[x], 128 outputs */
int bitz_fold_1b_rows_dev(bitz_ctx* ctx, const void* p, const void* suffix, uint64_t n, void* out);
/* No counterpart in the reference. bitz_fold_1b_rows_dev for callers that
   cannot allocate device memory (graph capture, serving loops, harnesses that
   forbid allocation): its block partials live in a fixed 1 MiB device array
   per GPU (512 chunks x 128 sums), so the call makes no allocation. Calls
   that overlap on one GPU share that array and race: serialise them. */
int bitz_fold_1b_rows_noalloc(bitz_ctx* ctx, const void* p, const void* suffix, uint64_t n, void* out);
/* fold_b128_elems (ring_switch.rs:1486-1560): out[x] = sum over set bits i of suffix[x] of w[i] */
int bitz_fold_b128_dev(bitz_ctx* ctx, const void* suffix, uint64_t n, const void* w, void* out);
/* acc[x] += eta * e[x] (SumcheckProver::glue, ligerito.rs:3125; add_ood_basis, src/ligerito_flock.rs:1590) */
int bitz_axpy_dev(bitz_ctx* ctx, void* acc, const void* e, uint64_t n, const uint8_t eta[16]);
/* round_msg_lsb / round_msg_and_eval_lsb (ligerito.rs:2499-2600): out_host = [u0, u2(, y)] */
int bitz_lig_msg_dev(bitz_ctx* ctx, const void* f, const void* b, uint64_t n, int with_eval,
                     uint8_t out_host[48]);
/* fold_and_msg_lsb's fold (ligerito.rs:2634-2680): f'[j] = f0 + r (f0+f1), b' likewise */
int bitz_lig_fold_dev(bitz_ctx* ctx, const void* f, const void* b, uint64_t n, const uint8_t r[16],
                      void* fo, void* bo);
/* induce_sumcheck_poly basis (ligerito.rs:1968-2070): out[j] = sum_i aw[i] X_j(q_i), 2^log_n outputs */
int bitz_induce_basis_dev(bitz_ctx* ctx, const uint64_t* queries_host, uint32_t nq, int log_n,
                          const void* alpha_w, void* out);
/* find_pow (src/bitz/pcs.rs:1252-1267): smallest nonce with >= bits leading zero bits of
 * BLAKE3(prefix16 || seed16 || nonce_le) */
int bitz_pow_find_dev(bitz_ctx* ctx, const uint8_t prefix[16], const uint8_t seed[16], uint32_t bits,
                      uint64_t* nonce);

/* ---- host-buffer commit with a device-resident handle ---- */
/* rows: 2^s columns, each 2^t bits as 2^(t-6) u64 words, column-major
 * (rows + c*2^(t-6)); the packed message is exactly these words read as
 * Gf128 (src/ligerito_flock.rs:198-217). Keeps rows/packed message, codeword
 * and tree on the device inside the handle. Synchronizing. */
int bitz_commit(bitz_ctx* ctx, const uint64_t* rows, uint32_t t, uint32_t s, uint32_t log_batch,
                uint32_t log_inv_rate, uint8_t root[32], bitz_commit_handle** out);
/* the same commitment from rows already in device memory (copied into the handle) */
int bitz_commit_device_rows(bitz_ctx* ctx, const void* rows, uint32_t t, uint32_t s, uint32_t log_batch,
                            uint32_t log_inv_rate, uint8_t root[32], bitz_commit_handle** out);
void bitz_commit_free(bitz_commit_handle* h);
/* Batched commitments in one launch. No counterpart in the reference (its
 * commitments run one at a time, commit_rs_flock_from_rows,
 * src/ligerito_flock.rs:198-296); each item's codeword and tree equal
 * bitz_commit_device_rows's for the same rows. The items' kernel sequences
 * (NTT passes, the LDS pass with the leaf hashes, the upper Merkle levels)
 * are captured into one HIP graph, items on up to 8 parallel branches, and
 * the graph is launched once; the instantiated graph is kept per item list
 * (pointers and shapes, the last 4 lists), so the same batch of resident
 * buffers replays it. packed: the device rows (2^(t+s-7) Gf128, the layout
 * bitz_commit takes); codeword: 16 << (t+s-7+log_inv_rate) bytes; tree:
 * (2 leaves - 1) * 32 bytes, leaves = 2^(t+s-7-log_batch+log_inv_rate).
 * Synchronous: roots are written when it returns. */
typedef struct bitz_commit_item {
    const void* packed;
    uint32_t t, s, log_batch, log_inv_rate;
    void* codeword;
    void* tree;
    uint8_t root[32]; /* out */
} bitz_commit_item;
int bitz_commit_batch(bitz_ctx* ctx, uint32_t n, bitz_commit_item* items);
/* device pointers held by the handle (valid until bitz_commit_free) */
const void* bitz_commit_packed(const bitz_commit_handle* h);
const void* bitz_commit_codeword(const bitz_commit_handle* h);
const void* bitz_commit_tree(const bitz_commit_handle* h);
/* copy one codeword position (2^log_batch Gf128) and the tree to the host */
int bitz_commit_read_tree(bitz_commit_handle* h, uint8_t* tree_out, uint64_t bytes);

/* ---- the BitZ standalone opening ----
 * prove_standalone (src/piop/spartan/protocol/bitz_opener.rs:672-707): the
 * outer Blake3Transcript binds the statement (root, Ligerito ladder, shape,
 * generator, prime width), runs Round 0, samples the prime q and the point,
 * then BitZProver::prove (src/bitz/mod.rs:142-199) opens
 * <eq(r1) (x) eq(r2), bits> = claimed mod q through the fold round, the
 * batched grand-product GKR, the inner-product sumcheck, the ring switch and
 * the recursive Ligerito opening (custom Johnson ladder at the commitment's
 * rate and initial_k, 100-bit target). Proof bytes follow
 * BitZOpeningProof::to_bytes (bitz_opener.rs:300-323). Synchronizing. */
/* the honest claimed value (standalone_evaluation, bitz_opener.rs:643-670) */
int bitz_standalone_evaluation(bitz_ctx* ctx, bitz_commit_handle* h, uint8_t claimed_out[16]);
/* the proof of `claimed`; *proof_len receives the size (also when capacity is too small) */
int bitz_prove_standalone(bitz_ctx* ctx, bitz_commit_handle* h, const uint8_t claimed[16], uint8_t* proof,
                          uint64_t capacity, uint64_t* proof_len);
/* custom Johnson ladder (src/ligerito_flock.rs:535-665): writes [m, log_n, initial_k, yr_log_n,
 * round0_grinding, levels, then per level: log_inv_rate, log_msg_cols, log_num_interleaved,
 * k_recursive, queries, grinding_bits, fold_grinding_bits, ood_samples]; returns the count */
int bitz_ligerito_config(uint32_t m, uint32_t log_inv_rate, uint32_t initial_k, uint32_t target_bits,
                         int32_t* out, uint32_t capacity);
/* verify_standalone (src/piop/spartan/protocol/bitz_opener.rs:709-747) ->
 * BitZVerifier::verify (src/bitz/mod.rs:218-255) and the recursive Ligerito
 * verifier (pcs/ligerito.rs:4585-4760) with the residual evaluated densely:
 * BITZ_OK when the proof (BitZOpeningProof::to_bytes layout) opens the
 * commitment `root` of shape (t, s, log_batch = initial_k, log_inv_rate) to
 * `claimed` (16-byte little-endian integer), BITZ_ERR_VERIFY with the reason
 * otherwise. Transcript checks run on the host, the dense tables (eq tables,
 * images, the ring-switch basis, the induced bases) on the context's device.
 * Synchronizing. */
int bitz_verify_standalone(bitz_ctx* ctx, const uint8_t root[32], uint32_t t, uint32_t s, uint32_t log_inv_rate,
                           uint32_t initial_k, const uint8_t claimed[16], const uint8_t* proof, uint64_t proof_len);

/* ---- batched and served commit+open jobs ----
 * A job is one commitment and its standalone opening (Pcs::commit +
 * standalone_evaluation / prove_standalone). The server runs jobs on
 * `workers` host threads, each with a context (pool + stream) of its own, so
 * jobs overlap on the device; it admits them in submission order when a worker
 * is free and the footprint of the jobs in flight plus the next one fits
 * hbm_budget bytes (a job larger than the budget runs alone). A job's proof
 * bytes are those of the same job run by itself. Job structs belong to the
 * caller and must stay valid until done. */
typedef struct bitz_job {
    const void* rows;        /* 2^s * 2^(t-6) u64 words, column-major (host, or device memory) */
    uint32_t t, s, log_batch, log_inv_rate;
    uint32_t rows_on_device; /* rows are device memory */
    uint32_t have_claim;     /* 0: claimed is computed (standalone_evaluation) inside the job */
    uint8_t claimed[16];     /* u128 LE, in or out */
    uint8_t* proof;          /* caller buffer of proof_capacity bytes */
    uint64_t proof_capacity;
    uint8_t root[32];        /* out */
    uint64_t proof_len;      /* out */
    int32_t status;          /* out: BITZ_OK or the job's error (message in error[]) */
    int32_t done;            /* out: set under the server's lock, read it with bitz_server_done */
    double submit_ms, start_ms, end_ms; /* server clock (bitz_server_clock_ms) */
    char error[128];
    /* kind BITZ_JOB_MUL_U64: the committed rows are the witness grid of
     * mul_n u64 multiplications (MulWitness::bitz_bit_rows,
     * src/piop/spartan/mul.rs:268-330) of the host arrays mul_x, mul_y,
     * packed on the device; rows is unused and t, s are set at submission
     * (MulLayout). Kind BITZ_JOB_MUL_U64_PROOF: the full u64 multiplication
     * proof of mul_x, mul_y (bitz_prove_mul_u64 with the products computed
     * on the device, custom:1:4; log_batch and log_inv_rate must be 4 and 1),
     * verified by bitz_verify_mul_u64 when verify is set; claimed receives
     * the first 16 bytes of the outer transcript state digest. Kind
     * BITZ_JOB_SHA256_CHAIN: the SHA-256 chain proof of 2^chain_log_n host
     * blocks chain_blocks from FIPS 180-4's H_0 (bitz_prove_sha256_chain),
     * verified by bitz_verify_sha256_chain when verify is set; chain_digest
     * receives H_N, claimed the first 16 bytes of the state digest, t and s
     * the product tensor's grid (the footprint's shape). */
    uint32_t kind;           /* BITZ_JOB_ROWS (0), BITZ_JOB_MUL_U64 (1), _MUL_U64_PROOF (2), _SHA256_CHAIN (3) */
    uint32_t verify;         /* nonzero: the worker verifies the proof (bitz_verify_standalone) */
    const uint64_t* mul_x;
    const uint64_t* mul_y;
    uint64_t mul_n;
    int32_t verified;        /* out: 1 when the proof was verified and accepted */
    int32_t reserved;
    const uint32_t* chain_blocks; /* 2^chain_log_n x 16 u32 words (host) */
    uint32_t chain_log_n;
    uint32_t chain_digest[8];     /* out */
    uint32_t reserved2;
} bitz_job;
#define BITZ_JOB_ROWS 0
#define BITZ_JOB_MUL_U64 1
#define BITZ_JOB_MUL_U64_PROOF 2
#define BITZ_JOB_SHA256_CHAIN 3
typedef struct bitz_server bitz_server;
/* device bytes one job holds while it runs (0 for an invalid shape) */
uint64_t bitz_job_footprint(uint32_t t, uint32_t s, uint32_t log_batch, uint32_t log_inv_rate);
int bitz_server_create(int device, uint32_t workers, uint64_t hbm_budget, bitz_server** out);
int bitz_server_submit(bitz_server* srv, bitz_job* job);
int bitz_server_done(bitz_server* srv, const bitz_job* job);
int bitz_server_wait_all(bitz_server* srv);
double bitz_server_clock_ms(bitz_server* srv);
void bitz_server_destroy(bitz_server* srv);
/* submits n jobs (largest first) to a fresh server and waits for all */
int bitz_batch_run(int device, uint32_t workers, uint64_t hbm_budget, bitz_job* jobs, uint32_t n);

/* ---- committed bit grid of the native multiplication relations ---- */
/* MulLayout::new + with_split_shift + bitz_params (src/piop/spartan/mul.rs:71-139)
 * for word_bits 32, 64 or 128: out = {gate_vars g (capacity 2^g = max(n, 256)
 * rounded up to a power of two), col_vars s, row_vars t = g - s + log2(4 word_bits),
 * committed bits n = t + s}. Errors: empty batch, domain too large, shift out of
 * range (MulError, mul.rs:48-59) -> BITZ_ERR_SHAPE. */
int bitz_mul_layout(uint32_t word_bits, uint64_t multiplications, int32_t split_shift, uint32_t out[4]);
/* MulWitness::bitz_bit_rows / write_bitz_bit_rows (mul.rs:268-330): the 2^s
 * columns of 2^t bits (2^(t-6) u64 words each, column c at rows + c 2^(t-6),
 * the layout bitz_commit takes) of the limbs [x | y | lo | hi] of each gate:
 * bit b of limb l at row bit ((l word_bits + b) << (g - s)) | (gate >> s) of
 * column gate & (2^s - 1) (bitz_cell, mul.rs:140-150); padding gates are
 * zero and every word is written. x, y (and lo, hi when given: supplied
 * product limbs, MulWitness::from_rows) are device arrays of n values of
 * word_bits bits (little-endian); with lo = hi = NULL the limbs are the
 * 2 word_bits-bit product x*y (MulRow::new). rows_words must be 2^(t+s-6)
 * (InvalidRowShape otherwise). Asynchronous on the context's stream. */
int bitz_mul_bit_rows_dev(bitz_ctx* ctx, uint32_t word_bits, uint64_t multiplications, int32_t split_shift,
                          const void* x, const void* y, const void* lo, const void* hi, void* rows,
                          uint64_t rows_words);

/* ---- the u64 multiplication proof (Spartan PIOP over F_q, discharged by BitZ) ---- */
/* protocol::commit + protocol::prove for MulLayout<u64> at the lambda = 100
 * profile with the custom:1:4 Johnson ladder
 * (src/piop/spartan/protocol/mod.rs:1117-1170 -> bitz_opener.rs:439-494,
 * prove_piop mod.rs:1410-1506): x, y are device arrays of n u64 operands;
 * the products lo, hi are computed on the device, or taken from lo, hi when
 * given (a supplied witness, MulWitness::from_rows; a wrong one yields a
 * proof the verifier rejects). Writes the commitment root, the outer
 * Blake3Transcript's state digest after the proof (the value the
 * reference's tests/transcript_state_pins.rs:163-172 pins) and the proof:
 * the Spartan messages as 16-byte field elements (4 per outer round, Az Bz
 * Cz, 3 per inner round; SpartanProof::payload_elements, mod.rs:940-957)
 * followed by BitZOpeningProof::to_bytes (bitz_opener.rs:297-323).
 * *proof_len is set even when capacity is too small (BITZ_ERR_ARG).
 * Shapes: committed grids of 2^20..2^35 bits with t <= 24. */
int bitz_prove_mul_u64(bitz_ctx* ctx, const void* x, const void* y, const void* lo, const void* hi, uint64_t n,
                       int32_t split_shift, uint8_t root[32], uint8_t state[32], uint8_t* proof, uint64_t capacity,
                       uint64_t* proof_len);
/* protocol::verify for MulLayout<u64> (mod.rs:1174-1207 -> bitz_opener.rs
 * verify): replays the outer transcript from (root, n, split_shift) and the
 * proof, checks every sumcheck round, the outer terminal
 * eq(tau, r_x)(Az Bz - Cz), the inner terminal against the public weight at
 * r_y, and the BitZ opening of the bitified claim. BITZ_OK when accepted,
 * BITZ_ERR_VERIFY otherwise; state (nullable) receives the outer transcript
 * digest. */
int bitz_verify_mul_u64(bitz_ctx* ctx, const uint8_t root[32], uint64_t n, int32_t split_shift,
                        const uint8_t* proof, uint64_t proof_len, uint8_t state[32]);

/* ---- the SHA-256 chain proof (linear PIOP over F_q, virtual BitZ opening) ---- */
/* generate_sha256_chain_witnesses + commit_sha256_chain_witness +
 * prove_sha256_chain at the lambda = 100 profile (custom:1:4)
 * (src/piop/spartan/sha256/chain.rs:635-845 -> protocol/linear.rs:371-582 ->
 * bitz_opener.rs:1597-1624, src/bitz/virt.rs:214-301): blocks are 2^log_n
 * host blocks of 16 u32 words (log_n in 7..16), chained from initial_state
 * (8 u32 words; NULL = FIPS 180-4's H_0): H_{i+1} = Compress(H_i, M_i). The
 * compression circuit is compiled on the host (once per initial state); the
 * witness, the commitment and the opening run on the device. Writes the
 * commitment root of the source bits, the final chaining value H_N, the
 * outer Blake3Transcript's state digest after the proof (the value
 * tests/transcript_state_pins.rs:471-497 pins for 2^7 blocks) and the proof
 * (BitZOpeningProof::to_bytes; the boundary nonces are zero at this
 * profile). *proof_len is set even when capacity is too small (BITZ_ERR_ARG). */
int bitz_prove_sha256_chain(bitz_ctx* ctx, const uint32_t* blocks, uint32_t log_n, const uint32_t* initial_state,
                            uint8_t root[32], uint32_t digest[8], uint8_t state[32], uint8_t* proof,
                            uint64_t capacity, uint64_t* proof_len);
/* verify_sha256_chain (chain.rs:850-872 -> linear.rs:586-705,
 * virt.rs:304-360): replays the outer transcript from the statement (blocks,
 * initial_state, digest), the root and the proof's Round-0 value, and checks
 * the virtual opening. BITZ_OK when accepted, BITZ_ERR_VERIFY otherwise;
 * state (nullable) receives the outer transcript digest. */
int bitz_verify_sha256_chain(bitz_ctx* ctx, const uint32_t* blocks, uint32_t log_n, const uint32_t* initial_state,
                             const uint32_t digest[8], const uint8_t root[32], const uint8_t* proof,
                             uint64_t proof_len, uint8_t state[32]);

/* The device transcript (Blake3Transcript, src/transcript/mod.rs:56-96) checked
 * against the host's: a host transcript absorbs `prefix` (absorb_slice), its
 * state goes to the device, which then runs n x (absorb_slice(msg_i of
 * msg_len bytes), 16-byte challenge) and writes the challenges to out_host. */
int bitz_transcript_challenges_dev(bitz_ctx* ctx, const uint8_t* prefix, uint64_t prefix_len, const uint8_t* msgs,
                                   uint32_t n, uint32_t msg_len, uint8_t* out_host);

/* BLAKE3 of a host buffer on the host (the transcript hash); _parts absorbs
 * data[0, split) and data[split, len) as two updates (incremental hashing) */
void bitz_host_blake3(const uint8_t* data, uint64_t len, uint8_t out[32]);
void bitz_host_blake3_parts(const uint8_t* data, uint64_t len, uint64_t split, uint8_t out[32]);

#ifdef __cplusplus
}
#endif
#endif
