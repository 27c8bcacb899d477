// The outer (Blake3Transcript) side of a standalone opening: the typed
// statement and claim frames, the ladder digest, Round 0's header and
// grinding boundary, and the transcript-driven prime search. Each function
// follows the reference function it names; see host_proto.h.
#include <string.h>

#include <algorithm>

#include "host_proto.h"

namespace bitz {
namespace {

// The embedded slim template every custom Johnson ladder starts from
// (vendor/flock-mod/crates/flock-core/configs/ligerito/m{22..35}_slim.toml,
// the same header strings at every m); the hash is forced to BLAKE3
// (src/ligerito_flock/configuration.rs:146).
const char kAnalysisVersion[] = "johnson_ood_row_union_over_bchks25_thm_4_6";
const char kFieldName[] = "f128";
const char kHashName[] = "blake3";

void put_u64(std::vector<uint8_t>* o, uint64_t v) {
    for (int i = 0; i < 8; i++) o->push_back((uint8_t)(v >> (8 * i)));
}
void put_u32(std::vector<uint8_t>* o, uint32_t v) {
    for (int i = 0; i < 4; i++) o->push_back((uint8_t)(v >> (8 * i)));
}
void put_f64(std::vector<uint8_t>* o, double v) {
    uint64_t bits;
    memcpy(&bits, &v, 8);
    put_u64(o, bits);
}
void put_str(std::vector<uint8_t>* o, const char* s) {
    const size_t n = strlen(s);
    put_u64(o, n);
    o->insert(o->end(), (const uint8_t*)s, (const uint8_t*)s + n);
}

// StatementFrame (src/ligerito_flock.rs:1057-1160): a field is
// 0x06 || tag || kind || count u64 || payload || 0x07, absorbed unframed.
enum : uint8_t { kBytes = 1, kU8 = 2, kU64 = 3, kU128 = 4, kGf128 = 5 };
class Frame {
   public:
    Frame(Transcript& tr, const char* domain) : tr_(tr) {
        tr_.absorb_slice(std::string("bitz/ligerito-flock/statement-frame/v1"));
        tr_.absorb_slice(std::string(domain));
    }
    void field(uint8_t tag, uint8_t kind, uint64_t count, const void* payload, size_t len) {
        uint8_t head[11];
        head[0] = 0x06;
        head[1] = tag;
        head[2] = kind;
        memcpy(head + 3, &count, 8);
        tr_.absorb_inner(head, 11);
        tr_.absorb_inner(payload, len);
        const uint8_t end = 0x07;
        tr_.absorb_inner(&end, 1);
    }
    void usize(uint8_t tag, uint64_t v) { field(tag, kU64, 1, &v, 8); }
    void usizes(uint8_t tag, const std::vector<uint64_t>& v) { field(tag, kU64, v.size(), v.data(), 8 * v.size()); }
    void byte(uint8_t tag, uint8_t v) { field(tag, kU8, 1, &v, 1); }

   private:
    Transcript& tr_;
};

}  // namespace

// bincode 1.3 with its fixint little-endian defaults, of
// LigeritoSecurityConfig (vendor/flock-mod/crates/flock-core/src/pcs/ligerito.rs:541-651):
// usize as u64, strings and vectors behind a u64 length, unit enum variants
// as u32 indices, Option as a 0/1 byte before the value, fields in
// declaration order.
void ligerito_bincode(const LigConfig& cfg, std::vector<uint8_t>* out) {
    std::vector<uint8_t>& o = *out;
    o.clear();
    put_u64(&o, cfg.m);
    put_u64(&o, cfg.log_n);
    put_u64(&o, cfg.initial_k);
    put_u64(&o, cfg.target_bits);
    put_str(&o, kAnalysisVersion);
    put_str(&o, kFieldName);
    put_str(&o, kHashName);
    put_u32(&o, 0);  // GrindingStep::PostCommitPreQueries
    put_u64(&o, cfg.levels.size());
    for (const LigLevel& lv : cfg.levels) {
        put_u64(&o, lv.log_inv_rate);
        put_u64(&o, lv.log_msg_cols);
        put_u64(&o, lv.log_num_interleaved);
        put_u64(&o, lv.k_recursive);
        put_u32(&o, 1);  // SoundnessRegime::JohnsonOod
        o.push_back(1);  // eta: Some
        put_f64(&o, cfg.eta);
        o.push_back(0);  // proximity_loss: None
        put_u64(&o, lv.queries);
        put_u64(&o, lv.grinding_bits);
        put_u64(&o, lv.fold_grinding_bits);
        put_u64(&o, lv.ood_samples);
        put_u64(&o, cfg.target_bits);
        put_f64(&o, lv.eps_pg);
        put_f64(&o, lv.eps_query);
        o.push_back(1);  // expected_eps_ood_bits: Some
        put_f64(&o, lv.eps_ood);
    }
    put_u64(&o, cfg.yr_log_n);
}

// ResolvedLigerito::digest (src/ligerito_flock/configuration.rs:157-159)
void ligerito_digest(const LigConfig& cfg, uint8_t out[32]) {
    std::vector<uint8_t> b;
    ligerito_bincode(cfg, &b);
    blake3_hash_host(b.data(), b.size(), out);
}

int standalone_q_bits(int t) { return std::min(113, 126 - t); }

// bind_standalone_statement (src/piop/spartan/protocol/bitz_opener.rs:589-609)
// -> absorb_standalone_mod_q_statement (src/ligerito_flock.rs:1245-1276).
// The commitment's parameters are the committed ones: m = t + s, the
// level-0 rate, the batch 2^initial_k, profile Fast (code 0, the commit's
// default, :222-228) and BLAKE3 (code 1); the configuration is the
// verifier's (pcs/ligerito.rs:1441-1490).
void bind_standalone_statement(Transcript& tr, const uint8_t root[32], const LigConfig& cfg, int t, int s,
                               const uint8_t generator[16]) {
    Frame f(tr, "bitz/ligerito-flock/standalone-mod-q-statement/v1");
    f.field(0x01, kBytes, 32, root, 32);
    f.usize(0x02, (uint64_t)cfg.m);
    f.usize(0x03, (uint64_t)cfg.levels[0].log_inv_rate);
    f.usize(0x04, (uint64_t)cfg.initial_k);
    f.byte(0x05, 0);
    f.byte(0x06, 1);
    std::vector<uint64_t> rates, rec_cols, rec_ks, queries, grinding, fold, ood;
    for (size_t i = 0; i < cfg.levels.size(); i++) {
        const LigLevel& lv = cfg.levels[i];
        rates.push_back(lv.log_inv_rate);
        if (i > 0) {
            rec_cols.push_back(lv.log_msg_cols);
            rec_ks.push_back(lv.k_recursive);
        }
        queries.push_back(lv.queries);
        grinding.push_back(lv.grinding_bits);
        fold.push_back(lv.fold_grinding_bits);
        ood.push_back(lv.ood_samples);
    }
    f.usizes(0x08, rates);
    f.usize(0x09, cfg.levels.size() - 1);
    f.usize(0x0a, (uint64_t)cfg.levels[0].log_msg_cols);
    f.usize(0x0b, (uint64_t)cfg.initial_k);
    f.usize(0x0c, (uint64_t)cfg.initial_k);
    f.usizes(0x0d, rec_cols);
    f.usizes(0x0e, rec_ks);
    f.usizes(0x0f, queries);
    f.usizes(0x10, grinding);
    f.usizes(0x11, fold);
    f.usizes(0x12, ood);
    f.byte(0x13, 1);
    f.usize(0x20, (uint64_t)t);
    f.usize(0x21, (uint64_t)s);
    f.usize(0x22, 1);
    f.field(0x30, kGf128, 1, generator, 16);
    f.usize(0x31, (uint64_t)standalone_q_bits(t));
    // a Johnson ladder runs Round 0: 1 + its grinding bits
    f.usize(0x32, 1 + (uint64_t)cfg.round0_grinding_bits);
    tr.absorb_slice(std::string("bitz/bitz-opener/standalone/v1"));
    uint8_t digest[32];
    ligerito_digest(cfg, digest);
    tr.absorb_slice(digest, 32);
}

// absorb_ood_round_header (src/ligerito_flock.rs:1488-1496)
void ood_round_header(Transcript& tr, int packed_vars, uint32_t bits) {
    tr.absorb_slice(std::string("bitz/core/ood-round/v1"));
    tr.absorb_u64((uint64_t)packed_vars);
    tr.absorb_slice(&bits, 4);
}

// derive_grinding_seed_in_domain (src/piop/spartan/grinding.rs:145-160)
void grinding_seed(Transcript& tr, const char* domain, uint64_t index, uint32_t bits, uint8_t seed[32]) {
    tr.absorb_slice(std::string("bitz/spartan/fiat-shamir-grinding/v1"));
    tr.absorb_slice(std::string(domain));
    tr.absorb_u64(index);
    tr.absorb_slice(&bits, 4);
    tr.challenge(seed, 32);
}

// absorb_grinding_nonce (grinding.rs:272-275)
void absorb_grinding_nonce(Transcript& tr, uint64_t nonce) {
    tr.absorb_slice(std::string("bitz/spartan/fiat-shamir-grinding/nonce/v1"));
    tr.absorb_u64(nonce);
}

// absorb_standalone_mod_q_claim (src/ligerito_flock.rs:1278-1284)
void absorb_standalone_claim(Transcript& tr, u128 q, u128 claimed) {
    Frame f(tr, "bitz/ligerito-flock/standalone-mod-q-claim/v1");
    uint8_t b[32];
    u128_put(q, b);
    u128_put(claimed, b + 16);
    f.field(0x30, kU128, 2, b, 32);
}

namespace {

int bit_length(u128 v) {
    int n = 0;
    while (v) {
        n++;
        v >>= 1;
    }
    return n;
}

// strong_round (sampling.rs:247-263): base^odd, then up to twos - 1 squarings
bool strong_round(const ModN& m, u128 base, u128 odd, int twos) {
    const u128 n = m.n, one = m.to_mont(1), minus_one = m.to_mont(n - 1);
    u128 b = m.to_mont(base), x = one;
    for (int i = bit_length(odd) - 1; i >= 0; i--) {
        x = m.mont(x, x);
        if ((odd >> i) & 1) x = m.mont(x, b);
    }
    if (x == one || x == minus_one) return true;
    for (int i = 1; i < twos; i++) {
        x = m.mont(x, x);
        if (x == minus_one) return true;
    }
    return false;
}

}  // namespace

// sample_prime_context (src/prime_sampling.rs:66-111) and
// sample_prime_public (vendor/field/src/prime/sampling.rs:290-345). The
// policy's rounds and rejection budget are budgets() (:33-56).
// sample_below (vendor/field/src/prime/sampling.rs:110-138): whole
// little-endian u64 words from the sampling stream, masked to the bound's
// width, rejected at or above it
bool sample_below(Transcript& tr, u128 bound, uint64_t max_draws, u128* out) {
    if (bound == 0) return false;
    const int bits = bit_length(bound - 1);
    if (bits == 0) {
        *out = 0;
        return true;
    }
    const int words = (bits + 63) / 64;
    const u128 mask = bits == 128 ? ~(u128)0 : (((u128)1 << bits) - 1);
    for (uint64_t d = 0; d < max_draws; d++) {
        u128 c = 0;
        for (int w = 0; w < words; w++) {
            uint8_t b[8];
            tr.sampling_bytes(b, 8);
            uint64_t v;
            memcpy(&v, b, 8);
            c |= (u128)v << (64 * w);
        }
        c &= mask;
        if (c < bound) {
            *out = c;
            return true;
        }
    }
    return false;
}

bool sample_prime_context(Transcript& tr, u128 lo, u128 hi, uint32_t security_bits, u128* out, std::string* err) {
    if (lo > hi) {
        *err = "prime interval is empty";
        return false;
    }
    if (hi >= ((u128)1 << 126)) {  // the Montgomery arithmetic's range
        *err = "prime interval maximum must be below 2^126";
        return false;
    }
    const u128 first_odd = (lo > 3 ? lo : 3) | 1;
    if (first_odd > hi) {
        *err = "prime interval contains no odd candidate";
        return false;
    }
    const u128 last_odd = hi - ((hi & 1) == 0 ? 1 : 0);
    const uint64_t attempts = first_odd == last_odd ? 1 : 64ull * (uint64_t)bit_length(hi);
    const uint32_t security = std::max<uint32_t>(security_bits + 16, 128);
    const uint32_t cand_bits = (uint32_t)bit_length((u128)(attempts - 1));
    const uint32_t rounds = (security + cand_bits + 1) / 2;
    const uint64_t draws = security + cand_bits + (uint32_t)bit_length(rounds);
    tr.absorb_slice(std::string("bitz/shared-prime-sampling/v1"));
    uint8_t b[16];
    u128_put(lo, b);
    tr.absorb_slice(b, 16);
    u128_put(hi, b);
    tr.absorb_slice(b, 16);
    tr.absorb_slice(&security_bits, 4);
    tr.absorb_u64(attempts);
    tr.absorb_u64(draws);
    // odd candidates 2(first + i) + 1 over i < count
    const u128 first = (lo > 3 ? lo : 3) >> 1;
    const u128 last = (hi & 1) ? hi >> 1 : (hi >> 1) - 1;
    const u128 count = last - first + 1;
    static const unsigned kSmall[16] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53};
    for (uint64_t a = 0; a < attempts; a++) {
        u128 index;
        if (!sample_below(tr, count, draws, &index)) {
            *err = "prime search: rejection sampling exhausted";
            return false;
        }
        const u128 cand = ((first + index) << 1) | 1;
        int small = 0;  // 1: a small prime, -1: a small factor
        for (unsigned p : kSmall) {
            if (cand == p) {
                small = 1;
                break;
            }
            if (cand % p == 0) {
                small = -1;
                break;
            }
        }
        if (small == 1) {
            *out = cand;
            return true;
        }
        if (small == -1) continue;
        u128 odd = cand - 1;
        int twos = 0;
        while ((odd & 1) == 0) {
            odd >>= 1;
            twos++;
        }
        ModN m;
        m.set(cand);
        bool accepted = true;
        for (uint32_t r = 0; r < rounds; r++) {
            u128 base;
            if (!sample_below(tr, cand - 3, draws, &base)) {
                *err = "prime search: rejection sampling exhausted";
                return false;
            }
            if (!strong_round(m, base + 2, odd, twos)) {
                accepted = false;
                break;
            }
        }
        if (accepted) {
            *out = cand;
            return true;
        }
    }
    *err = "no prime found after the policy's candidates";
    return false;
}

// sample_prime_with_bits (src/prime_sampling.rs:125-137)
bool sample_prime_with_bits(Transcript& tr, int bits, u128* out, std::string* err) {
    if (bits < 32 || bits > 120) {
        *err = "prime bit width must be in [32, 120]";
        return false;
    }
    const u128 top = (u128)1 << (bits - 1);
    return sample_prime_context(tr, top, top | (top - 1), 128, out, err);
}

// transcript_uniform_mod (src/prime_sampling.rs:11-33)
u128 sample_residue(Transcript& tr, u128 q) {
    uint8_t b[16];
    tr.challenge(b, 16);
    const u128 lo = u128_from(b);
    tr.challenge(b, 16);
    const u128 hi = u128_from(b);
    const u128 two128 = mulmod((u128)1 << 64, (u128)1 << 64, q);
    return addmod(mulmod(hi % q, two128, q), lo % q, q);
}

}  // namespace bitz
