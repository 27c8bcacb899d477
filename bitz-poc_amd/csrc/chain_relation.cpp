// The SHA-256 chain relation, compiled on the host once per initial state.
//
// The compression circuit (crates/circuit/src/sha256.rs:213-478) is
// replayed on a recorder with the constraint generator's semantics
// (crates/circuit/src/constraints.rs:389-547, lib.rs:524-542): a lift of a
// Boolean combination appends one row of the F2 map M, a hint allocates
// fresh Boolean witness columns, an asserted-zero integer combination
// appends one row of C. The chain then splits M by source column
// (src/piop/spartan/sha256/chain.rs:268-343): block and hint columns stay
// local (hints shifted past the 256 state cells), state columns read the
// previous instance's output cells (prev) and, for instance 0, the parity
// of the initial state's bits (first, on the constant column); the terminal
// rows copy the last instance's output cells (last). Digests follow
// binary.rs:1233-1260 (a binary CSC map), chain.rs:345-370 (the relation)
// and binary.rs:648-720 (the chained map).
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "chain_impl.h"
#include "host_proto.h"

namespace bitz {
namespace chain {

const uint32_t kIV[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                         0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};

namespace {

const uint32_t* const kK = kSha256K;

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// ---- the recorder ----
struct Bit {  // F2 combination: constant + sorted Boolean witness columns
    bool c = false;
    std::vector<uint32_t> s;
};
Bit operator^(const Bit& a, const Bit& b) {
    Bit o;
    o.c = a.c != b.c;
    std::set_symmetric_difference(a.s.begin(), a.s.end(), b.s.begin(), b.s.end(), std::back_inserter(o.s));
    return o;
}
struct Lin {  // integer combination over lifted cells (zero-based)
    int64_t c = 0;
    std::map<uint32_t, int64_t> t;
    Lin& operator+=(const Lin& o) {
        c += o.c;
        for (auto& kv : o.t) {
            int64_t v = (t[kv.first] += kv.second);
            if (v == 0) t.erase(kv.first);
        }
        return *this;
    }
    Lin scaled(int64_t k) const {
        Lin o;
        o.c = c * k;
        for (auto& kv : t)
            if (kv.second != 0 && k != 0) o.t[kv.first] = kv.second * k;
        return o;
    }
};
using Word = std::vector<Bit>;
struct UInt {
    Word w;     // the Boolean bits (32)
    Lin full;   // the Z value
};

struct Recorder {
    uint32_t next_bool = kInputBits;
    std::vector<Bit> m;       // lifts
    std::vector<Lin> c;       // asserted-zero rows
    Word hint(int n) {
        Word w(n);
        for (int i = 0; i < n; i++) w[i].s = {next_bool + (uint32_t)i};
        next_bool += n;
        return w;
    }
    // bitz_unsigned: the full value and the low 32 bits' value
    void lift(const Word& bits, Lin* full, Lin* low) {
        for (size_t i = 0; i < bits.size(); i++) {
            const uint32_t idx = (uint32_t)m.size();
            m.push_back(bits[i]);
            full->t[idx] += (int64_t)1 << i;
            if (i < 32) low->t[idx] += (int64_t)1 << i;
        }
    }
    UInt uint(const Word& bits) {
        UInt u;
        Lin low;
        u.w = bits;
        lift(bits, &u.full, &low);
        return u;
    }
};

Word wxor(const Word& a, const Word& b) {
    Word o(32);
    for (int i = 0; i < 32; i++) o[i] = a[i] ^ b[i];
    return o;
}
Word wrotr(const Word& a, int n) {
    Word o(32);
    for (int i = 0; i < 32; i++) o[i] = a[(i + n) % 32];
    return o;
}
Word wshr(const Word& a, int n) {
    Word o(32);
    for (int i = 0; i < 32; i++)
        if (i + n < 32) o[i] = a[i + n];
    return o;
}

// sum_32 (sha256.rs:217-252): width-bit hint, lifted; sum(terms) - wide = 0
UInt sum32(Recorder& r, int width, const std::vector<const Lin*>& terms) {
    Word bits = r.hint(width);
    Lin full, low;
    r.lift(bits, &full, &low);
    Lin total;
    for (auto* t : terms) total += *t;
    total += full.scaled(-1);
    r.c.push_back(total);
    UInt o;
    o.w.assign(bits.begin(), bits.begin() + 32);
    o.full = low;
    return o;
}
// sum_doubled_32 (sha256.rs:256-302): 2 sum(ordinary) + sum(doubled) - 2 half = 0
UInt sum_doubled32(Recorder& r, const std::vector<const Lin*>& ordinary, const std::vector<const Lin*>& doubled) {
    Word bits = r.hint(35);
    Lin full, low;
    r.lift(bits, &full, &low);
    Lin total, dsum;
    for (auto* t : ordinary) total += *t;
    for (auto* d : doubled) dsum += *d;
    Lin row = total.scaled(2);
    row += dsum;
    row += full.scaled(-2);
    r.c.push_back(row);
    UInt o;
    o.w.assign(bits.begin(), bits.begin() + 32);
    o.full = low;
    return o;
}

// compress (sha256.rs:329-431) over inputs [block 512 | state 256]
void record(Recorder& r) {
    auto input_word = [](int base) {
        Word w(32);
        for (int i = 0; i < 32; i++) w[i].s = {(uint32_t)(base + i)};
        return w;
    };
    std::vector<UInt> state, sched;
    for (int w = 0; w < 8; w++) state.push_back(r.uint(input_word(512 + 32 * w)));
    for (int w = 0; w < 16; w++) sched.push_back(r.uint(input_word(32 * w)));
    for (int i = 16; i < 64; i++) {
        const Word& w15 = sched[i - 15].w;
        UInt s0 = r.uint(wxor(wxor(wrotr(w15, 7), wrotr(w15, 18)), wshr(w15, 3)));
        const Word& w2 = sched[i - 2].w;
        UInt s1 = r.uint(wxor(wxor(wrotr(w2, 17), wrotr(w2, 19)), wshr(w2, 10)));
        sched.push_back(sum32(r, 34, {&sched[i - 16].full, &s0.full, &sched[i - 7].full, &s1.full}));
    }
    UInt a = state[0], b = state[1], c = state[2], d = state[3], e = state[4], f = state[5], g = state[6],
         h = state[7];
    for (int i = 0; i < 64; i++) {
        UInt big1 = r.uint(wxor(wxor(wrotr(e.w, 6), wrotr(e.w, 11)), wrotr(e.w, 25)));
        UInt xy = r.uint(wxor(e.w, f.w));
        UInt xz = r.uint(wxor(e.w, g.w));
        Lin choice = f.full;  // 2 Ch = y + z - (x ^ y) + (x ^ z)
        choice += g.full;
        choice += xy.full.scaled(-1);
        choice += xz.full;
        UInt big0 = r.uint(wxor(wxor(wrotr(a.w, 2), wrotr(a.w, 13)), wrotr(a.w, 22)));
        UInt xyz = r.uint(wxor(wxor(a.w, b.w), c.w));
        Lin majority = a.full;  // 2 Maj = x + y + z - (x ^ y ^ z)
        majority += b.full;
        majority += c.full;
        majority += xyz.full.scaled(-1);
        Lin k;
        k.c = kK[i];
        UInt ne = sum_doubled32(r, {&d.full, &h.full, &big1.full, &k, &sched[i].full}, {&choice});
        UInt na = sum_doubled32(r, {&h.full, &big1.full, &k, &sched[i].full, &big0.full}, {&choice, &majority});
        h = g;
        g = f;
        f = e;
        e = ne;
        d = c;
        c = b;
        b = a;
        a = na;
    }
    const UInt* working[8] = {&a, &b, &c, &d, &e, &f, &g, &h};
    for (int i = 0; i < 8; i++) sum32(r, 33, {&state[i].full, &working[i]->full});
}

void put64(std::vector<uint8_t>* o, uint64_t v) {
    for (int i = 0; i < 8; i++) o->push_back((uint8_t)(v >> (8 * i)));
}

void build_csc(const std::vector<std::vector<uint32_t>>& rows, Csc* out) {
    std::vector<uint32_t> count(kChainFBar + 1, 0);
    for (auto& r : rows)
        for (uint32_t c : r) count[c + 1]++;
    out->offsets.assign(kChainFBar + 1, 0);
    for (int c = 0; c < kChainFBar; c++) out->offsets[c + 1] = out->offsets[c] + count[c + 1];
    out->rows.assign(out->offsets[kChainFBar], 0);
    std::vector<uint32_t> fill(out->offsets.begin(), out->offsets.end() - 1);
    for (uint32_t r = 0; r < rows.size(); r++)
        for (uint32_t c : rows[r]) out->rows[fill[c]++] = r;
    // compute_digest (binary.rs:1233-1260)
    std::vector<uint8_t> b;
    const char dom[] = "bitz/f2-cell-map/v1";
    b.insert(b.end(), dom, dom + sizeof(dom) - 1);
    put64(&b, kChainHBar);
    put64(&b, kChainFBar);
    put64(&b, out->rows.size());
    for (uint32_t v : out->offsets) put64(&b, v);
    for (uint32_t v : out->rows) put64(&b, v);
    blake3_hash_host(b.data(), b.size(), out->digest);
}

int chain_output_column(int word, int bit) { return kOutputFBase - 256 + 33 * word + bit; }

std::shared_ptr<Relation> compile(const std::array<uint32_t, 8>& iv) {
    Recorder r;
    record(r);
    if ((int)r.m.size() + 1 != kHBar || (int)r.c.size() != kConstraints || (int)r.next_bool != kInputBits + kHintBits)
        return nullptr;
    auto rel = std::make_shared<Relation>();
    rel->initial_state = iv;
    // M with the constant row first (constraints.rs:461-475), columns + 1
    rel->m_off.push_back(0);
    rel->m_cols.push_back(0);
    rel->m_off.push_back(1);
    for (auto& row : r.m) {
        if (row.c) rel->m_cols.push_back(0);
        for (uint32_t x : row.s) rel->m_cols.push_back(x + 1);
        rel->m_off.push_back((uint32_t)rel->m_cols.size());
    }
    // the chained maps
    std::vector<std::vector<uint32_t>> loc(kChainHBar), prv(kChainHBar), fst(kChainHBar), lst(kChainHBar);
    for (int hr = 0; hr < kHBar; hr++) {
        bool parity = false;
        for (uint32_t k = rel->m_off[hr]; k < rel->m_off[hr + 1]; k++) {
            const uint32_t f = rel->m_cols[k];
            if (f < 513) {
                loc[hr].push_back(f);
            } else if (f < 769) {
                const int sb = (int)f - 513, word = sb / 32, bit = sb % 32;
                prv[hr].push_back((uint32_t)chain_output_column(word, bit));
                parity ^= ((iv[word] >> bit) & 1u) != 0;
            } else {
                loc[hr].push_back(f - 256);
            }
        }
        if (parity) fst[hr].push_back(0);
    }
    for (int t = 0; t < kTerminal; t++) lst[kHBar + t].push_back((uint32_t)chain_output_column(t / 32, t % 32));
    build_csc(loc, &rel->local);
    build_csc(prv, &rel->prev);
    build_csc(fst, &rel->first);
    build_csc(lst, &rel->last);
    // C by derived column (constant at column 0, lifts at index + 1)
    std::vector<std::vector<std::pair<uint32_t, int64_t>>> cols(kChainHBar);
    for (uint32_t row = 0; row < r.c.size(); row++) {
        const Lin& z = r.c[row];
        if (z.c) cols[0].push_back({row, z.c});
        for (auto& kv : z.t) cols[kv.first + 1].push_back({row, kv.second});
    }
    rel->c_off.push_back(0);
    for (auto& col : cols) {
        for (auto& e : col) {
            rel->c_rows.push_back(e.first);
            rel->c_vals.push_back(e.second);
        }
        rel->c_off.push_back((uint32_t)rel->c_rows.size());
    }
    // chain_relation_digest (chain.rs:345-370)
    std::vector<uint8_t> b;
    const char dom[] = "bitz/sha256-chain/local-relation/v2";
    b.insert(b.end(), dom, dom + sizeof(dom) - 1);
    for (const Csc* m : {&rel->local, &rel->prev, &rel->first, &rel->last}) b.insert(b.end(), m->digest, m->digest + 32);
    for (uint32_t w : iv)
        for (int i = 0; i < 4; i++) b.push_back((uint8_t)(w >> (8 * i)));
    b.push_back('C');
    put64(&b, kConstraints);
    put64(&b, kChainHBar);
    put64(&b, rel->c_rows.size());
    for (int col = 0; col < kChainHBar; col++) {
        put64(&b, rel->c_off[col + 1] - rel->c_off[col]);
        for (uint32_t k = rel->c_off[col]; k < rel->c_off[col + 1]; k++) {
            put64(&b, rel->c_rows[k]);
            put64(&b, (uint64_t)rel->c_vals[k]);
        }
    }
    blake3_hash_host(b.data(), b.size(), rel->digest);
    return rel;
}

}  // namespace

std::shared_ptr<const Relation> relation(const std::array<uint32_t, 8>& iv) {
    static std::mutex mu;
    static std::vector<std::shared_ptr<Relation>> cache;
    std::lock_guard<std::mutex> lock(mu);
    for (auto& r : cache)
        if (r->initial_state == iv) return r;
    auto r = compile(iv);
    if (!r) return nullptr;
    if (cache.size() >= 4) cache.erase(cache.begin());
    cache.push_back(r);
    return r;
}

bool make_batch(int k, const std::array<uint32_t, 8>& iv, Batch* b) {
    if (k < 7 || k > 16) return false;
    b->k = k;
    b->n = 1ull << k;
    b->rel = relation(iv);
    if (!b->rel) return false;
    // packed_domain_vars and balanced_binary_params (constraints.rs:879-910)
    const uint64_t live = 1 + b->n * kFInstance;
    int vars = 0;
    while ((1ull << vars) < live) vars++;
    b->tf = std::max(7, (vars + 1) / 2);
    b->sf = vars - b->tf;
    b->th = std::min(k, 13);
    b->sh = kLocalBits + k - b->th;
    // derive_params at sha256_instance_facts (profile.rs:480-646,
    // prime.rs:35-47): a single derived prime of width min(113, 128 - k)
    const int width = std::min(113, 128 - k);
    b->pmin = (u128)1 << (width - 1);
    const u128 bit_max = ((u128)1 << width) - 1;
    const u128 no_wrap = 1 + (~(u128)0) / (((u128)1 << k) + 1);
    b->pmax = std::min(bit_max, no_wrap);
    const Relation& r = *b->rel;
    b->map_nnz = r.local.rows.size() * b->n + r.prev.rows.size() * (b->n - 1) + r.first.rows.size() +
                 r.last.rows.size();
    std::vector<uint8_t> m;
    const char dom[] = "bitz/chained-packed-source-map/v1";
    m.insert(m.end(), dom, dom + sizeof(dom) - 1);
    for (const Csc* c : {&r.local, &r.prev, &r.first, &r.last}) m.insert(m.end(), c->digest, c->digest + 32);
    const uint64_t words[6] = {b->n, 1ull << (b->th + b->sh), 1ull << (b->tf + b->sf), (uint64_t)kChainHBar * b->n,
                               live, b->map_nnz};
    for (uint64_t v : words)
        put64(&m, v);
    blake3_hash_host(m.data(), m.size(), b->map_digest);
    return true;
}

}  // namespace chain
}  // namespace bitz
