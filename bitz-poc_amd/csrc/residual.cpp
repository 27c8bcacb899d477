// Succinct residual evaluation of the Ligerito verifier's basis components
// (declarations and their definitions in residual.h).
#include "residual.h"

#include <vector>

namespace bitz {
namespace residual {

namespace {

inline E add(E a, E b) { return E{a.lo ^ b.lo, a.hi ^ b.hi}; }
inline E one() { return E{1, 0}; }
inline bool bit(E v, int j) { return j < 64 ? (v.lo >> j) & 1 : (v.hi >> (j - 64)) & 1; }

// v * x modulo x^128 + x^7 + x^2 + x + 1
inline E times_x(E v) {
    const uint64_t carry = v.hi >> 63;
    E o{v.lo << 1, (v.hi << 1) | (v.lo >> 63)};
    o.lo ^= carry * 0x87ull;
    return o;
}

// in place: bit j of a[i] <-> bit i of a[j], i, j < 64. Each stage swaps the
// off-diagonal j x j blocks of every 2j x 2j block: row k's columns
// [c + j] with row k + j's columns [c] (c's bit j clear).
void transpose64(uint64_t a[64]) {
    static const uint64_t kMask[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                      0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
    int stage = 0;
    for (int j = 32; j != 0; j >>= 1, stage++) {
        const uint64_t m = kMask[stage];
        for (int k = 0; k < 64; k++) {
            if (k & j) continue;
            const uint64_t t = ((a[k] >> j) ^ a[k | j]) & m;
            a[k] ^= t << j;
            a[k | j] ^= t;
        }
    }
}

}  // namespace

void transpose128(const E in[128], E out[128]) {
    // blocks of the 128 x 128 bit matrix M[i][b] = bit b of in[i]:
    // A = rows 0..63 bits 0..63, B = rows 0..63 bits 64..127,
    // C = rows 64..127 bits 0..63, D = rows 64..127 bits 64..127
    uint64_t A[64], B[64], C[64], D[64];
    for (int i = 0; i < 64; i++) {
        A[i] = in[i].lo;
        B[i] = in[i].hi;
        C[i] = in[64 + i].lo;
        D[i] = in[64 + i].hi;
    }
    transpose64(A);
    transpose64(B);
    transpose64(C);
    transpose64(D);
    for (int b = 0; b < 64; b++) {
        out[b] = E{A[b], C[b]};
        out[64 + b] = E{B[b], D[b]};
    }
}

void ring_switch(const E* z, int nz, const E* r, int nr, const E bw[128], E* out) {
    // S = sum_i beta_i (x) s[i], beta_i = x^i the first factor's basis
    E s[128], v[128], w[128], zt[128];
    for (auto& e : s) e = E{0, 0};
    s[0] = one();
    for (int k = 0; k < nr; k++) {
        // (z (x) 1) S: in the transposed form S = sum_b v[b] (x) beta_b the
        // first factor is each v[b]
        transpose128(s, v);
        for (int b = 0; b < 128; b++) w[b] = hostgf::mul(z[k], v[b]);
        transpose128(w, zt);
        for (int i = 0; i < 128; i++) s[i] = add(add(s[i], hostgf::mul(r[k], s[i])), zt[i]);
    }
    // a binary suffix y scales the first factor by c_y = eq(z[nr..nz), y):
    // out[y] = sum_i s[i] phi(c_y beta_i), phi(u) = sum_j bw[j] bit_j(u).
    // c -> sum_i s[i] phi(c beta_i) is GF(2)-linear: psi[k] is its value at
    // beta_k, sum_i s[i] phi(x^(k + i)).
    E ph[255];
    E xd = one();
    for (int d = 0; d < 255; d++) {
        E acc{0, 0};
        for (int j = 0; j < 128; j++)
            if (bit(xd, j)) acc = add(acc, bw[j]);
        ph[d] = acc;
        xd = times_x(xd);
    }
    E psi[128];
    for (int k = 0; k < 128; k++) {
        E acc{0, 0};
        for (int i = 0; i < 128; i++)
            if (s[i].lo | s[i].hi) acc = add(acc, hostgf::mul(s[i], ph[k + i]));
        psi[k] = acc;
    }
    const int yl = nz - nr;
    std::vector<E> c(1ull << yl);
    hostgf::eq_table(z + nr, yl, c.data());
    for (uint64_t y = 0; y < c.size(); y++) {
        E acc{0, 0};
        for (int k = 0; k < 128; k++)
            if (bit(c[y], k)) acc = add(acc, psi[k]);
        out[y] = acc;
    }
}

void induced(int log_n, const uint64_t* queries, int nq, const E* aw, const E* sks, const E* inv, const E* r,
             int yr_log, E* out) {
    const int p = log_n - yr_log;
    const uint64_t ny = 1ull << yr_log;
    for (uint64_t y = 0; y < ny; y++) out[y] = E{0, 0};
    std::vector<E> wk(log_n), tab(ny);
    for (int i = 0; i < nq; i++) {
        // W_k(q): s_0(q) = q, s_k = s_{k-1}^2 + s_{k-1}(v_{k-1}) s_{k-1}
        E sv{queries[i], 0};
        for (int k = 0; k < log_n; k++) {
            if (k > 0) sv = add(hostgf::mul(sv, sv), hostgf::mul(sks[k - 1], sv));
            wk[k] = hostgf::mul(sv, inv[k]);
        }
        // folded coordinates: 1 + r_k (1 + W_k); binary ones: W_k or 1
        E pre = aw[i];
        for (int k = 0; k < p; k++) pre = hostgf::mul(pre, add(one(), hostgf::mul(r[k], add(one(), wk[k]))));
        tab[0] = pre;
        for (int j = 0; j < yr_log; j++)
            for (uint64_t y = 0; y < (1ull << j); y++) tab[y + (1ull << j)] = hostgf::mul(tab[y], wk[p + j]);
        for (uint64_t y = 0; y < ny; y++) out[y] = add(out[y], tab[y]);
    }
}

void eq_point(const E* z, int nz, const E* r, int yr_log, E scale, E* out) {
    const int p = nz - yr_log;
    E sc = scale;
    for (int b = 0; b < p; b++) sc = hostgf::mul(sc, add(add(one(), z[b]), r[b]));
    hostgf::eq_table(z + p, yr_log, out);
    for (uint64_t y = 0; y < (1ull << yr_log); y++) out[y] = hostgf::mul(out[y], sc);
}

}  // namespace residual
}  // namespace bitz
