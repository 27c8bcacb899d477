// Host GF(2^128) loops for the last rounds of a sumcheck, where the two
// halves hold at most a few thousand elements: at that size one round on the
// host (a few microseconds with carry-less multiply instructions) is cheaper
// than a kernel launch and a device-to-host round trip per round.
//
// Elements are {lo, hi} 64-bit word pairs (the 16-byte Gf128 encoding of
// src/bitz/codec.rs:4-21), modulus x^128 + x^7 + x^2 + x + 1. Plain C++ (no
// device code); PCLMULQDQ is used when the CPU has it, an integer carry-less
// multiply otherwise, with identical results.
#pragma once
#include <stdint.h>

namespace bitz {
namespace hostgf {

struct E {
    uint64_t lo, hi;
};

// a * b
E mul(E a, E b);

// eq table over n coordinates: out[x] = prod_i (bit i of x ? r[i] : 1 + r[i])
void eq_table(const E* r, int n, E* out);

// One round of the GKR layer sumcheck over h slots of l = [el | eh] and
// r = [ol | oh] (2h entries each), slot weights w[x]:
//   out[0] = sum_x w[x] * (send_one ? eh * oh : el * ol)
//   out[1] = sum_x w[x] * (eh + el) * (oh + ol)
// (Sums::slot, src/bitz/kernels.rs:299-310, summed over the round's slots).
void gkr_round_sums(const E* l, const E* r, uint64_t h, const E* w, int send_one, E out[2]);

// v[j] = v[j] + rho * (v[j] + v[j + h]) for j < h (fold,
// src/bitz/kernels.rs:326-328).
void fold(E* v, uint64_t h, E rho);

// v[j] = v[j] + v[j + h] for j < h: an eq table over n coordinates becomes
// the table over the first n - 1 (the last coordinate's two weights sum to 1).
void drop_top(E* v, uint64_t h);

// One round of the inner-product sumcheck over pairs (m[2k], m[2k+1]),
// (w[2k], w[2k+1]): out = (sum m0 w0, sum m1 w1, sum (m0 + m1)(w0 + w1))
// (src/bitz/sumcheck.rs:146-198).
void inner_round(const E* m, const E* w, uint64_t pairs, E out[3]);

// v[i] = v[2i] + r (v[2i] + v[2i+1]) for i < half, in place (the LSB-first
// fold of the inner-product sumcheck, src/bitz/sumcheck.rs:372-400).
void fold_lsb(E* v, uint64_t half, E r);

// true when the carry-less multiply instructions are in use (PCLMULQDQ;
// VPCLMULQDQ on 512-bit registers, four elements at a time, where present)
bool using_clmul();
bool using_vpclmul();

}  // namespace hostgf
}  // namespace bitz
