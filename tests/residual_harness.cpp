// Test harness: the succinct residual evaluation of libbitz_hip
// (residual.cpp, with the host GF loops it multiplies with) built into a
// small shared library, so the CPU tests can check it against the oracle's
// dense tables folded coordinate by coordinate.
#include "../bitz-poc_amd/csrc/host_gf.cpp"
#include "../bitz-poc_amd/csrc/residual.cpp"

using bitz::hostgf::E;

extern "C" {
void rh_transpose128(const uint64_t* in, uint64_t* out) { bitz::residual::transpose128((const E*)in, (E*)out); }

void rh_ring_switch(const uint64_t* z, int nz, const uint64_t* r, int nr, const uint64_t* bw, uint64_t* out) {
    bitz::residual::ring_switch((const E*)z, nz, (const E*)r, nr, (const E*)bw, (E*)out);
}

void rh_induced(int log_n, const uint64_t* queries, int nq, const uint64_t* aw, const uint64_t* sks,
                const uint64_t* inv, const uint64_t* r, int yr_log, uint64_t* out) {
    bitz::residual::induced(log_n, queries, nq, (const E*)aw, (const E*)sks, (const E*)inv, (const E*)r, yr_log,
                            (E*)out);
}

void rh_eq_point(const uint64_t* z, int nz, const uint64_t* r, int yr_log, const uint64_t* scale, uint64_t* out) {
    bitz::residual::eq_point((const E*)z, nz, (const E*)r, yr_log, E{scale[0], scale[1]}, (E*)out);
}
}
