// Test harness: the host GF(2^128) loops of libbitz_hip (host_gf.cpp), built
// into a small shared library with both multiply paths exported, so the CPU
// tests can check each against the oracle.
#include "../bitz-poc_amd/csrc/host_gf.cpp"

using bitz::hostgf::E;

extern "C" {
int hg_has_clmul(void) { return bitz::hostgf::using_clmul() ? 1 : 0; }
int hg_has_vpclmul(void) { return bitz::hostgf::using_vpclmul() ? 1 : 0; }

void hg_mul(int clmul, const uint64_t* a, const uint64_t* b, uint64_t* out, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) {
        E x{a[2 * i], a[2 * i + 1]}, y{b[2 * i], b[2 * i + 1]};
        E z = clmul == 2 ? bitz::hostgf::mul(x, y)
                         : clmul ? bitz::hostgf::mul_call<true>(x, y) : bitz::hostgf::mul_call<false>(x, y);
        out[2 * i] = z.lo;
        out[2 * i + 1] = z.hi;
    }
}

void hg_eq_table(int clmul, const uint64_t* r, int n, uint64_t* out) {
    if (clmul == 2)
        bitz::hostgf::eq_table((const E*)r, n, (E*)out);
    else if (clmul)
        bitz::hostgf::eq_table_t<true>((const E*)r, n, (E*)out);
    else
        bitz::hostgf::eq_table_t<false>((const E*)r, n, (E*)out);
}

void hg_round_sums(int clmul, const uint64_t* l, const uint64_t* r, uint64_t h, const uint64_t* w, int send_one,
                   uint64_t* out) {
    if (clmul == 2)
        bitz::hostgf::gkr_round_sums((const E*)l, (const E*)r, h, (const E*)w, send_one, (E*)out);
    else if (clmul)
        bitz::hostgf::round_sums_t<true>((const E*)l, (const E*)r, h, (const E*)w, send_one, (E*)out);
    else
        bitz::hostgf::round_sums_t<false>((const E*)l, (const E*)r, h, (const E*)w, send_one, (E*)out);
}

void hg_inner_round(int clmul, const uint64_t* m, const uint64_t* w, uint64_t pairs, uint64_t* out) {
    if (clmul == 2)
        bitz::hostgf::inner_round((const E*)m, (const E*)w, pairs, (E*)out);
    else if (clmul)
        bitz::hostgf::inner_round_t<true>((const E*)m, (const E*)w, pairs, (E*)out);
    else
        bitz::hostgf::inner_round_t<false>((const E*)m, (const E*)w, pairs, (E*)out);
}

void hg_fold_lsb(int clmul, uint64_t* v, uint64_t half, const uint64_t* r) {
    E x{r[0], r[1]};
    if (clmul == 2)
        bitz::hostgf::fold_lsb((E*)v, half, x);
    else if (clmul)
        bitz::hostgf::fold_lsb_t<true>((E*)v, half, x);
    else
        bitz::hostgf::fold_lsb_t<false>((E*)v, half, x);
}

void hg_drop_top(uint64_t* v, uint64_t h) { bitz::hostgf::drop_top((E*)v, h); }

void hg_fold(int clmul, uint64_t* v, uint64_t h, const uint64_t* rho) {
    E r{rho[0], rho[1]};
    if (clmul == 2)
        bitz::hostgf::fold((E*)v, h, r);
    else if (clmul)
        bitz::hostgf::fold_t<true>((E*)v, h, r);
    else
        bitz::hostgf::fold_t<false>((E*)v, h, r);
}
}
