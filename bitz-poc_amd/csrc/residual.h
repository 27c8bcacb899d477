// Succinct evaluation of the Ligerito verifier's basis at the residual
// (src/bitz/pcs.rs:588-610 evaluate_basis, pcs/ligerito.rs:4362-4440 and
// 1852-1990, pcs/ring_switch.rs:2676-2790): each basis component of the
// recursive opening, folded least significant coordinate first by the
// challenges after its introduction, evaluated at every binary point y of
// the residual's yr_log coordinates, without building the component's
// table. Plain host C++ over the 16-byte element encoding of host_gf.h.
#pragma once
#include <stdint.h>

#include "host_gf.h"

namespace bitz {
namespace residual {

using hostgf::E;

// The ring-switch basis b(x) = sum_i bw[i] bit_i(eq(z, x)), x in {0,1}^nz,
// folded by r[0..nr) and evaluated at the binary points y < 2^(nz - nr):
// out[y] = sum_x eq((r, y), x) b(x).
// Evaluated in L (x) L (L = GF(2^128) over GF(2)): sum_x eq(z, x) (x) eq(r', x)
// is the product over coordinates of 1 (x) 1 + z_k (x) 1 + 1 (x) r'_k, whose
// first factor's coordinates are then weighted by bw; a binary y coordinate
// scales the first factor by eq(z_k, y_k) only.
void ring_switch(const E* z, int nz, const E* r, int nr, const E bw[128], E* out);

// The induced basis of one query set: out over x in {0,1}^log_n is
// sum_i aw[i] prod_{k in bits(x)} W_k(q_i), W_k(q) = s_k(q) / s_k(v_k) the
// normalised subspace polynomials (pcs/ligerito.rs:1760-1790); folded by
// r[0..log_n - yr_log) and evaluated at y < 2^yr_log. sks[k] = s_k(v_k),
// inv[k] its inverse, k < log_n.
void induced(int log_n, const uint64_t* queries, int nq, const E* aw, const E* sks, const E* inv, const E* r,
             int yr_log, E* out);

// scale * eq(z, x) over x in {0,1}^nz, folded by r[0..nz - yr_log) and
// evaluated at y < 2^yr_log (each folded coordinate contributes
// eq(z_b, r_b) = 1 + z_b + r_b).
void eq_point(const E* z, int nz, const E* r, int yr_log, E scale, E* out);

// 128 x 128 bit transpose: bit i of out[b] = bit b of in[i]
void transpose128(const E in[128], E out[128]);

}  // namespace residual
}  // namespace bitz
