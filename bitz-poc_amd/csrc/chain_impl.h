// The SHA-256 chain relation on the host (csrc/chain_relation.cpp): the
// compression circuit compiled to its F2 lift map M and integer constraint
// rows C, the chained maps, their canonical digests, and the batch geometry.
// Shared by the device prover (csrc/chain.hip).
#pragma once
#include <stdint.h>

#include <array>
#include <memory>
#include <vector>

#include "modq.h"
#include "sha256_host.h"

namespace bitz {
namespace chain {

constexpr int kInputBits = 512 + 256;
constexpr int kHintBits = 48 * 34 + 64 * 70 + 8 * 33;  // 6376
constexpr int kHBar = 20457;                           // lifted cells, constant included
constexpr int kFBar = 1 + kInputBits + kHintBits;      // 7145: [1 | block | state | hints]
constexpr int kConstraints = 184;
constexpr int kFInstance = 512 + kHintBits;            // 6888 committed cells per instance
constexpr int kChainFBar = kFInstance + 1;
constexpr int kTerminal = 256;
constexpr int kChainHBar = kHBar + kTerminal;          // 20713
constexpr int kLocalBits = 15;                         // local stride 2^15
constexpr int kOutputFBase = 6881;                     // output word w, bit b: 6881 + 33 w + b
constexpr int kPublicBits = (16 + 8) * 32;
constexpr int kLocalVars = 8;                          // 184 constraints -> 2^8

// a binary CSC map: for each source column, its derived rows (increasing)
struct Csc {
    std::vector<uint32_t> offsets, rows;
    uint8_t digest[32];
};

struct Relation {
    std::array<uint32_t, 8> initial_state;
    // M over the unchained f_bar (CSR: derived row -> f_bar columns)
    std::vector<uint32_t> m_off, m_cols;
    // C by derived column: (constraint row, coefficient)
    std::vector<uint32_t> c_off, c_rows;
    std::vector<int64_t> c_vals;
    Csc local, prev, first, last;  // the chained maps (kChainHBar x kChainFBar)
    uint8_t digest[32];            // chain_relation_digest
};

// the relation for a chain from `initial_state` (built once per state, kept)
std::shared_ptr<const Relation> relation(const std::array<uint32_t, 8>& initial_state);

// FIPS 180-4 5.3.3: H_0 (compress: sha256_host.h)
extern const uint32_t kIV[8];

// prepare_chain_instances at lambda 100: N = 2^k instances
struct Batch {
    int k;
    uint64_t n;
    int tf, sf;  // the committed source grid
    int th, sh;  // the product tensor's grid
    u128 pmin, pmax;
    uint64_t map_nnz;
    uint8_t map_digest[32];
    std::shared_ptr<const Relation> rel;
};
// false when k is outside [7, 16]
bool make_batch(int k, const std::array<uint32_t, 8>& initial_state, Batch* b);

// derived public column of public slot p (block bits, then terminal bits)
inline int public_h_column(int slot) { return slot < 512 ? 1 + 256 + slot : kHBar + (slot - 512); }

}  // namespace chain
}  // namespace bitz
