"""CPU restatement of the SHA-256 chain proof (prove_sha256_chain,
src/piop/spartan/sha256/chain.rs:817-845): the compression relation compiled
from its circuit, the chained source map, the witness, the linear-batching
PIOP over a runtime prime and its virtual BitZ opening.

TEST INFRASTRUCTURE ONLY (see oracle.py): the checker of the device prover.

Parts, each restating the reference file it cites:
  * the SHA-256 compression circuit (crates/circuit/src/sha256.rs:213-478)
    as the constraint generator records it (crates/circuit/src/
    constraints.rs:389-547, lib.rs:524-542): every `bitz` lift appends one
    row of the F2 map M (a XOR of Boolean witness columns), every hint
    allocates fresh Boolean witness columns, every `assert_zero` appends one
    integer row of C over the lifted cells;
  * the chain relation (chain.rs:268-370): M split by source column into
    the local / prev / first / last maps, the terminal rows, the signed
    constraint matrix, and their canonical digests (binary.rs:1224-1260,
    648-720);
  * the witness (chain.rs:635-794): the native chain, every compression's
    hint bits, the packed source rows [1 | block_i | hints_i ...] and the
    local-major product tensor D[local, instance];
  * the linear PIOP (protocol/linear.rs:371-582) with the chain's frames,
    bindings, runtime prime (sha256/prime.rs:233-262) and rank-one product
    claim (chain.rs:1080-1251);
  * the virtual opening (src/bitz/virt.rs:153-301) on this build's inner
    transcript, as the standalone opening.
"""
from __future__ import annotations

import numpy as np

import bitz_protocol as P
import oracle as O
import spartan_u64 as S
from bitz_protocol import g16, u64

# FIPS 180-4 4.2.2 and 5.3.3
ROUND_K = [
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2,
]
IV = [0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19]

INPUT_BITS = 512 + 256
HINT_BITS = 48 * 34 + 64 * 70 + 8 * 33       # 6376
H_BAR = 20_457                                 # lifted cells of one compression, constant included
F_BAR = 1 + INPUT_BITS + HINT_BITS             # 7145: [1 | block | state | hints]
CONSTRAINTS = 184
CHAIN_F_INSTANCE = 512 + HINT_BITS             # 6888: the state cells are not committed
CHAIN_F_BAR = CHAIN_F_INSTANCE + 1
TERMINAL = 256
CHAIN_H_BAR = H_BAR + TERMINAL                 # 20713
LOCAL_STRIDE = 1 << 15
LOCAL_BITS = 15
STATE_F_BASE, STATE_F_END = 1 + 512, 1 + 512 + 256
OUTPUT_F_BASE, OUTPUT_STRIDE = 6_881, 33
PUBLIC_BITS = (16 + 8) * 32
LOCAL_VARS = 8                                 # 184 constraints -> 2^8 (proof.rs:1230-1232)

PROTOCOL_DOMAIN = b"bitz/spartan-sha256-chain/product-linear/v1"
ASSIGNMENT_DOMAIN = b"bitz/spartan-sha256-chain-assignment/runtime-prime/v1"
STATEMENT_DOMAIN = b"bitz/spartan-sha256-chain/public-statement/v1"
PUBLIC_IO_DOMAIN = b"bitz/spartan-sha256-chain/public-io-batch/v1"
CONSTANT_ONE_DOMAIN = b"bitz/spartan-sha256-chain/constant-one-batch/v1"
LOCAL_POINT_DOMAIN = b"bitz/spartan-sha256-chain/local-row-point/v1"
INSTANCE_POINT_DOMAIN = b"bitz/spartan-sha256-chain/instance-point/v1"
CLAIM_DOMAIN = b"bitz/spartan-sha256-chain-opening/v1"
CLAIM_TAG = b"sha256-chain-opening-claim"
RELATION_DOMAIN = b"bitz/sha256-chain/local-relation/v2"
PRIME_DOMAIN = b"bitz/spartan-sha256/runtime-prime/v1"
VIRTUAL_SESSION = b"bitz/bitz-opener/virtual/v1"
VIRTUAL_LABEL = b"bitz/virtual-statement/v2"
MAP_DOMAIN = b"bitz/f2-cell-map/v1"
CHAINED_MAP_DOMAIN = b"bitz/chained-packed-source-map/v1"
PROFILE_NAME = b"lambda100"
MASK32 = (1 << 32) - 1


# ------------------------------------------------------------ native SHA-256
def _rotr(x: int, n: int) -> int:
    return ((x >> n) | (x << (32 - n))) & MASK32


def schedule(block) -> list[int]:
    w = [int(v) for v in block]
    for i in range(16, 64):
        s0 = _rotr(w[i - 15], 7) ^ _rotr(w[i - 15], 18) ^ (w[i - 15] >> 3)
        s1 = _rotr(w[i - 2], 17) ^ _rotr(w[i - 2], 19) ^ (w[i - 2] >> 10)
        w.append((w[i - 16] + s0 + w[i - 7] + s1) & MASK32)
    return w


def compression_hints(state, block) -> tuple[list[tuple[int, int]], list[int]]:
    """One compression (FIPS 180-4 6.2.2) with the circuit's hint values in
    allocation order (sha256.rs:393-471): the 48 schedule sums (34 bits),
    per round the e sum then the a sum (35 bits: the ordinary terms plus
    Ch, plus Maj for a: half of the doubled sum), then the 8 output sums
    (33 bits). Returns ((value, width) list, output state)."""
    hints = []
    w = [int(v) for v in block]
    for i in range(16, 64):
        s0 = _rotr(w[i - 15], 7) ^ _rotr(w[i - 15], 18) ^ (w[i - 15] >> 3)
        s1 = _rotr(w[i - 2], 17) ^ _rotr(w[i - 2], 19) ^ (w[i - 2] >> 10)
        total = w[i - 16] + s0 + w[i - 7] + s1
        hints.append((total, 34))
        w.append(total & MASK32)
    a, b, c, d, e, f, g, h = (int(v) for v in state)
    for i in range(64):
        big1 = _rotr(e, 6) ^ _rotr(e, 11) ^ _rotr(e, 25)
        ch = (e & f) ^ (~e & g & MASK32)
        big0 = _rotr(a, 2) ^ _rotr(a, 13) ^ _rotr(a, 22)
        maj = (a & b) ^ (a & c) ^ (b & c)
        e_half = d + h + big1 + ROUND_K[i] + w[i] + ch
        a_half = h + big1 + ROUND_K[i] + w[i] + big0 + ch + maj
        hints.append((e_half, 35))
        hints.append((a_half, 35))
        h, g, f, e = g, f, e, e_half & MASK32
        d, c, b, a = c, b, a, a_half & MASK32
    out = []
    for x, y in zip(state, (a, b, c, d, e, f, g, h)):
        total = int(x) + y
        hints.append((total, 33))
        out.append(total & MASK32)
    return hints, out


def compress(state, block) -> list[int]:
    return compression_hints(state, block)[1]


# ------------------------------------------------------------ the circuit
class _Bool:
    """An F2 linear combination: a constant and a set of Boolean witness
    columns (zero-based, without the constant)."""
    __slots__ = ("c", "s")

    def __init__(self, c: bool = False, s=frozenset()):
        self.c, self.s = c, s

    def __xor__(self, o: "_Bool") -> "_Bool":
        return _Bool(self.c ^ o.c, self.s ^ o.s)


_ZERO_BIT = _Bool()


class _Z:
    """An integer linear combination over the lifted cells."""
    __slots__ = ("c", "t")

    def __init__(self, c: int = 0, t=None):
        self.c, self.t = c, dict(t or {})

    def __add__(self, o: "_Z") -> "_Z":
        out = _Z(self.c + o.c, self.t)
        for k, v in o.t.items():
            nv = out.t.get(k, 0) + v
            if nv:
                out.t[k] = nv
            else:
                out.t.pop(k, None)
        return out

    def __neg__(self) -> "_Z":
        return _Z(-self.c, {k: -v for k, v in self.t.items()})

    def __sub__(self, o: "_Z") -> "_Z":
        return self + (-o)

    def scale(self, k: int) -> "_Z":
        return _Z(self.c * k, {i: v * k for i, v in self.t.items() if v * k})


class _Gen:
    """The constraint generator (constraints.rs:389-547)."""

    def __init__(self, inputs: int):
        self.next_bool = inputs
        self.m_rows: list[_Bool] = []
        self.c_rows: list[_Z] = []

    def lift(self, b: _Bool) -> int:
        self.m_rows.append(b)
        return len(self.m_rows) - 1

    def hint(self, n: int) -> list[_Bool]:
        first = self.next_bool
        self.next_bool += n
        return [_Bool(False, frozenset([first + i])) for i in range(n)]

    def uint(self, bits: list[_Bool]):
        """bitz_unsigned (lib.rs:524-542): one lift per bit; returns
        (bits, full, low 32 bits)"""
        full, low = _Z(), _Z()
        for i, b in enumerate(bits):
            term = _Z(0, {self.lift(b): 1 << i})
            full = full + term
            if i < 32:
                low = low + term
        return bits, full, low

    def assert_zero(self, z: _Z) -> None:
        self.c_rows.append(z)


def _w_xor(a, b):
    return [x ^ y for x, y in zip(a, b)]


def _w_rotr(a, n):
    return [a[(i + n) % 32] for i in range(32)]


def _w_shr(a, n):
    return [a[i + n] if i + n < 32 else _ZERO_BIT for i in range(32)]


def _sum32(gen: _Gen, width: int, terms):
    """sum_32 (sha256.rs:217-252): the width-bit hint, its lift, and
    sum(terms) - wide = 0; the result is the low 32 bits"""
    bits, full, low = gen.uint(gen.hint(width))
    total = _Z()
    for t in terms:
        total = total + t[1]
    gen.assert_zero(total - full)
    return bits[:32], low, low


def _sum_doubled32(gen: _Gen, ordinary, doubled):
    """sum_doubled_32 (sha256.rs:256-302): 2 sum(ordinary) + sum(doubled)
    - 2 half = 0 over the 35-bit half"""
    bits, full, low = gen.uint(gen.hint(35))
    total = _Z()
    for t in ordinary:
        total = total + t[1]
    dsum = _Z()
    for d in doubled:
        dsum = dsum + d
    gen.assert_zero(total.scale(2) + dsum - full.scale(2))
    return bits[:32], low, low


def compression_relation():
    """Replays compress (sha256.rs:329-431) on the generator over the 768
    inputs [block 512 | state 256]: returns (M rows, C rows, output bits)."""
    gen = _Gen(INPUT_BITS)
    inp = [_Bool(False, frozenset([i])) for i in range(INPUT_BITS)]
    block = [inp[32 * w: 32 * w + 32] for w in range(16)]
    state_w = [inp[512 + 32 * w: 512 + 32 * w + 32] for w in range(8)]
    state = [gen.uint(w) for w in state_w]
    sched = [gen.uint(w) for w in block]
    for i in range(16, 64):
        w15 = sched[i - 15][0]
        s0 = gen.uint(_w_xor(_w_xor(_w_rotr(w15, 7), _w_rotr(w15, 18)), _w_shr(w15, 3)))
        w2 = sched[i - 2][0]
        s1 = gen.uint(_w_xor(_w_xor(_w_rotr(w2, 17), _w_rotr(w2, 19)), _w_shr(w2, 10)))
        sched.append(_sum32(gen, 34, [sched[i - 16], s0, sched[i - 7], s1]))
    a, b, c, d, e, f, g, h = state
    for i in range(64):
        big1 = gen.uint(_w_xor(_w_xor(_w_rotr(e[0], 6), _w_rotr(e[0], 11)), _w_rotr(e[0], 25)))
        xy = gen.uint(_w_xor(e[0], f[0]))
        xz = gen.uint(_w_xor(e[0], g[0]))
        choice = f[1] + g[1] - xy[1] + xz[1]
        big0 = gen.uint(_w_xor(_w_xor(_w_rotr(a[0], 2), _w_rotr(a[0], 13)), _w_rotr(a[0], 22)))
        xyz = gen.uint(_w_xor(_w_xor(a[0], b[0]), c[0]))
        majority = a[1] + b[1] + c[1] - xyz[1]
        k = (None, _Z(ROUND_K[i]), _Z(ROUND_K[i]))
        new_e = _sum_doubled32(gen, [d, h, big1, k, sched[i]], [choice])
        new_a = _sum_doubled32(gen, [h, big1, k, sched[i], big0], [choice, majority])
        h, g, f, e = g, f, e, new_e
        d, c, b, a = c, b, a, new_a
    working = [a, b, c, d, e, f, g, h]
    outs = [_sum32(gen, 33, [state[i], working[i]]) for i in range(8)]
    assert gen.next_bool == INPUT_BITS + HINT_BITS and len(gen.m_rows) + 1 == H_BAR
    assert len(gen.c_rows) == CONSTRAINTS
    m_rows = [[0]]  # the implicit constant row (constraints.rs:461-475)
    for r in gen.m_rows:
        m_rows.append(([0] if r.c else []) + sorted(x + 1 for x in r.s))
    c_rows = []
    for z in gen.c_rows:
        row = ([(0, z.c)] if z.c else []) + sorted((k + 1, v) for k, v in z.t.items())
        c_rows.append(row)
    out_bits = [bit for o in outs for bit in o[0]]
    return m_rows, c_rows, out_bits


_RELATION = None


def relation():
    global _RELATION
    if _RELATION is None:
        _RELATION = compression_relation()
    return _RELATION


# ------------------------------------------------------------ sparse maps
def csc_from_rows(n_cols: int, rows) -> tuple[list[int], list[int]]:
    """CscMatrix::try_from_rows (sparse.rs:717-733 then into_csc): column
    offsets and the rows of each column in increasing order."""
    cols = [[] for _ in range(n_cols)]
    for r, entries in enumerate(rows):
        for c in entries:
            cols[c].append(r)
    offsets, indices = [0], []
    for col in cols:
        indices.extend(col)
        offsets.append(len(indices))
    return offsets, indices


def map_digest(n_rows: int, n_cols: int, offsets, indices) -> bytes:
    """compute_digest of a PreparedVirtualMap (binary.rs:1233-1260)."""
    h = O.Blake3()
    h.update(MAP_DOMAIN + u64(n_rows) + u64(n_cols) + u64(len(indices)))
    h.update(np.asarray(offsets, dtype="<u8").tobytes())
    h.update(np.asarray(indices, dtype="<u8").tobytes())
    return h.digest(32)


def chain_output_f_column(word: int, bit: int) -> int:
    return OUTPUT_F_BASE - 256 + OUTPUT_STRIDE * word + bit


def chain_public_h_column(slot: int) -> int:
    return 1 + 256 + slot if slot < 512 else H_BAR + (slot - 512)


class ChainRelation:
    """build_chain_local_relation (chain.rs:268-343) and
    chain_relation_digest (:345-370)."""

    def __init__(self, initial_state=IV):
        self.initial_state = [int(v) for v in initial_state]
        m_rows, c_rows, _ = relation()
        local = [[] for _ in range(CHAIN_H_BAR)]
        prev = [[] for _ in range(CHAIN_H_BAR)]
        first = [[] for _ in range(CHAIN_H_BAR)]
        last = [[] for _ in range(CHAIN_H_BAR)]
        for h_row, row in enumerate(m_rows):
            parity = False
            for f in row:
                if f < STATE_F_BASE:
                    local[h_row].append(f)
                elif f < STATE_F_END:
                    sb = f - STATE_F_BASE
                    word, bit = sb // 32, sb % 32
                    prev[h_row].append(chain_output_f_column(word, bit))
                    parity ^= bool((self.initial_state[word] >> bit) & 1)
                else:
                    local[h_row].append(f - 256)
            if parity:
                first[h_row].append(0)
        for t in range(TERMINAL):
            last[H_BAR + t].append(chain_output_f_column(t // 32, t % 32))
        self.maps = {}
        for name, rows in (("local", local), ("prev", prev), ("first", first), ("last", last)):
            off, idx = csc_from_rows(CHAIN_F_BAR, rows)
            self.maps[name] = dict(rows=rows, offsets=off, indices=idx,
                                   digest=map_digest(CHAIN_H_BAR, CHAIN_F_BAR, off, idx))
        # the signed constraint matrix over the widened assignment, by column
        cols = [[] for _ in range(CHAIN_H_BAR)]
        for r, row in enumerate(c_rows):
            for c, v in row:
                cols[c].append((r, v))
        self.c_rows, self.c_cols = c_rows, cols
        nnz = sum(len(c) for c in cols)
        h = O.Blake3()
        h.update(RELATION_DOMAIN)
        for name in ("local", "prev", "first", "last"):
            h.update(self.maps[name]["digest"])
        h.update(b"".join(int(w).to_bytes(4, "little") for w in self.initial_state))
        h.update(b"C" + u64(CONSTRAINTS) + u64(CHAIN_H_BAR) + u64(nnz))
        parts = []
        for col in cols:
            parts.append(u64(len(col)))
            for r, v in col:
                parts.append(u64(r) + int(v).to_bytes(8, "little", signed=True))
        h.update(b"".join(parts))
        self.digest = h.digest(32)
        self.m_rows = m_rows


# ------------------------------------------------------------ prepared batch
def _ceil_log2(n: int) -> int:
    return max(n - 1, 0).bit_length()


class ChainBatch:
    """prepare_chain_instances (chain.rs:524-566) for N = 2^k chained
    compressions at the lambda = 100 profile, JOHNSON opener (custom:1:4)."""

    def __init__(self, k: int, initial_state=IV):
        if not 7 <= k <= 16:
            raise ValueError("the chain needs 2^7..2^16 compressions")
        self.k, self.n = k, 1 << k
        self.rel = ChainRelation(initial_state)
        src_vars = _ceil_log2(1 + self.n * CHAIN_F_INSTANCE)
        tf = max(P.LOG_PACKING, (src_vars + 1) // 2)  # balanced_binary_params (constraints.rs:900-910)
        self.f = (tf, src_vars - tf)
        th = min(k, 13)
        self.h = (th, LOCAL_BITS + k - th)
        mp = self.rel.maps
        nnz = (len(mp["local"]["indices"]) * self.n + len(mp["prev"]["indices"]) * (self.n - 1)
               + len(mp["first"]["indices"]) + len(mp["last"]["indices"]))
        hsh = O.Blake3()
        hsh.update(CHAINED_MAP_DOMAIN)
        for name in ("local", "prev", "first", "last"):
            hsh.update(mp[name]["digest"])
        for v in (self.n, 1 << sum(self.h), 1 << sum(self.f), CHAIN_H_BAR * self.n,
                  1 + CHAIN_F_INSTANCE * self.n, nnz):
            hsh.update(u64(v))
        self.map_digest = hsh.digest(32)
        self.map_nnz = nnz
        # derive_params (profile.rs:480-646) at sha256_instance_facts
        # (prime.rs:35-47): defect 2^96, lift arity and opening t as below,
        # not direct, tau arity 8 + k, degree 2
        width = min(113, 126, 128 - k)
        self.pmin = 1 << (width - 1)
        self.pmax = min((1 << width) - 1, 1 + (2**128 - 1) // ((1 << k) + 1))
        tau_raw = (width - 1) - np.log2(8 + k)
        assert tau_raw >= 100 and (width - 1) - 1 >= 100, "grinding boundaries not restated"
        self.cfg = P.custom_johnson_config(sum(self.f), 1, 4)
        self.ood_bits = self.cfg["round0_grinding"]

    def security_hash(self) -> bytes:
        """hash_security_params (proof.rs:2000-2034)."""
        out = u64(len(PROFILE_NAME)) + PROFILE_NAME + (100).to_bytes(4, "little")
        out += S.u128b(self.pmin) + S.u128b(self.pmax) + b"\x00"
        out += (0).to_bytes(4, "little") * 3 + b"\x00"
        out += (0).to_bytes(4, "little") * 2  # forest and ring-switch grinding at lambda 100
        out += b"\x01" + int(self.ood_bits).to_bytes(4, "little")
        out += u64(100)
        return out

    def ligerito_hash(self) -> bytes:
        """hash_ligerito_config (proof.rs:2036-2067) over the prover config
        of the ladder (pcs/ligerito.rs:1441-1475)."""
        lv = self.cfg["levels"]
        out = u64(len(lv) - 1) + u64(lv[0]["msg_cols"]) + u64(self.cfg["initial_k"]) + u64(self.cfg["initial_k"])
        for vals in ([x["rate"] for x in lv], [x["msg_cols"] for x in lv[1:]], [x["k"] for x in lv[1:]],
                     [x["queries"] for x in lv], [x["grinding"] for x in lv], [x["fold_grinding"] for x in lv],
                     [x["ood"] for x in lv]):
            out += u64(len(vals)) + b"".join(u64(v) for v in vals)
        return out + b"\x01"

    def statement_binding(self, blocks, digest) -> bytes:
        """chain_statement_binding (chain.rs:1364-1383)."""
        out = STATEMENT_DOMAIN + b"".join(w.to_bytes(4, "little") for w in self.rel.initial_state)
        out += u64(len(blocks))
        out += b"".join(int(w).to_bytes(4, "little") for b in blocks for w in b)
        out += b"".join(int(w).to_bytes(4, "little") for w in digest)
        return O.blake3(out)

    def assignment_binding(self, root: bytes, stmt: bytes) -> bytes:
        """chain_assignment_binding (chain.rs:1385-1418)."""
        out = ASSIGNMENT_DOMAIN + root + self.rel.digest + self.map_digest + stmt
        for v in (sum(self.f), 1, self.cfg["initial_k"], self.n, self.k, self.h[0], self.h[1], 1,
                  self.f[0], self.f[1], 1):
            out += u64(v)
        out += b"\x00\x01"  # profile and hash codes of the commitment
        out += self.security_hash() + self.ligerito_hash()
        return O.blake3(out)


# ------------------------------------------------------------ witness
def instance_source_bits(block, hints) -> np.ndarray:
    """the committed cells of one instance: the block bits, then the hint
    bits (chain.rs:719-726)"""
    bits = [(int(block[i // 32]) >> (i % 32)) & 1 for i in range(512)]
    for v, w in hints:
        bits += [(v >> b) & 1 for b in range(w)]
    return np.array(bits, dtype=np.uint8)


def _pack_grid(cells: np.ndarray, t: int, s: int) -> np.ndarray:
    """flat cells (column-major: cell c 2^t + b) -> per-column u64 rows"""
    full = np.zeros(1 << (t + s), dtype=np.uint8)
    full[: len(cells)] = cells
    return np.packbits(full.reshape(-1, 64), axis=1, bitorder="little").view("<u8").reshape(-1)


class ChainWitness:
    """generate_sha256_chain_witnesses (chain.rs:635-702): the chain, every
    compression's hints, the packed source and the product tensor."""

    def __init__(self, batch: ChainBatch, blocks, flip_hint=None):
        """flip_hint: (instance, hint bit) to flip, a dishonest witness whose
        derived grid still follows the map but breaks a constraint"""
        assert len(blocks) == batch.n
        self.blocks = [[int(w) for w in b] for b in blocks]
        states = [list(batch.rel.initial_state)]
        src = [np.ones(1, dtype=np.uint8)]
        f_bars = []
        for b in self.blocks:
            hints, out = compression_hints(states[-1], b)
            states.append(out)
            cells = instance_source_bits(b, hints)
            if flip_hint is not None and flip_hint[0] == len(states) - 2:
                cells[512 + flip_hint[1]] ^= 1
            src.append(cells)
            st = [(states[-2][i // 32] >> (i % 32)) & 1 for i in range(256)]
            f_bars.append(np.concatenate([np.ones(1, np.uint8), cells[:512], np.array(st, np.uint8), cells[512:]]))
        self.states = states
        self.digest = states[-1]
        tf, sf = batch.f
        self.source_cells = np.concatenate(src)
        self.source_rows = _pack_grid(self.source_cells, tf, sf)
        # h_bar of every instance: M over its unchained f_bar (the lifts)
        m_rows = batch.rel.m_rows
        nnz = sum(len(r) for r in m_rows)
        ri = np.repeat(np.arange(len(m_rows)), [len(r) for r in m_rows])
        ci = np.fromiter((c for r in m_rows for c in r), dtype=np.int64, count=nnz)
        F = np.stack(f_bars, axis=1)  # F_BAR x N
        H = np.zeros((H_BAR, batch.n), dtype=np.uint8)
        np.bitwise_xor.at(H, ri, F[ci])
        self.h_bar = H
        # D[local, instance], local-major: cell = local N + instance
        D = np.zeros((LOCAL_STRIDE, batch.n), dtype=np.uint8)
        D[:H_BAR] = H
        out_bits = np.array([(self.digest[i // 32] >> (i % 32)) & 1 for i in range(256)], np.uint8)
        D[H_BAR:CHAIN_H_BAR, batch.n - 1] = out_bits
        th, sh = batch.h
        self.product_rows = _pack_grid(D.reshape(-1), th, sh)


# ------------------------------------------------------------ the PIOP
def pin_blocks(n: int):
    """the reference pin's blocks (tests/transcript_state_pins.rs:474-476)"""
    return [[((i * 0x9E3779B9) & MASK32) ^ j for j in range(16)] for i in range(n)]


def challenge_point(tr: P.Transcript, domain: bytes, n: int, q: int) -> list[int]:
    """challenge_point (linear.rs:306-317)"""
    S.spartan_message(tr, b"challenge-domain", domain)
    return [S.squeeze(tr, q) for _ in range(n)]


def collapse(rel: ChainRelation, weights, q: int) -> list[int]:
    """collapse_signed_columns (sumcheck/bridge/repeated.rs:4-22):
    beta[j] = sum_i weights[i] C[i][j] mod q"""
    return [sum(weights[r] * v for r, v in col) % q for col in rel.c_cols]


def public_words(blocks, digest, i: int) -> list[int]:
    """instance_words (chain.rs:221-228): the block, then the digest at the
    last instance and zeros elsewhere"""
    return list(blocks[i]) + (list(digest) if i + 1 == len(blocks) else [0] * 8)


def _outer(b: ChainBatch, root: bytes, blocks, digest, mp: int, round0):
    """prove_linear / verify_linear's outer transcript (linear.rs:398-558,
    586-705) up to the fork tag. `round0(tr)` runs Round 0's grinding,
    draw and value (prover: computes y; verifier: checks the proof's) and
    returns (point, y, nonce)."""
    n, k = b.n, b.k
    tr = P.Transcript()
    stmt = b.statement_binding(blocks, digest)
    binding = b.assignment_binding(root, stmt)
    # statement frames (chain.rs:962-990)
    frames = [(b"protocol", PROTOCOL_DOMAIN), (b"integer-relation", b.rel.digest),
              (b"chained-boolean-map", b.map_digest)]
    for tag, v in ((b"instance-vars", k), (b"instance-count", n), (b"assignment-row-vars", b.h[0]),
                   (b"assignment-column-vars", b.h[1]), (b"source-row-vars", b.f[0]),
                   (b"source-column-vars", b.f[1])):
        frames.append((tag, u64(v)))
    frames += [(b"public-sha256-chain", stmt), (b"assignment-oracle", binding)]
    for tag, payload in frames:
        S.spartan_message(tr, tag, payload)
    # the opener policy (configuration.rs:224-227), then Round 0 bound
    # before the prime (ood.rs:75-86, ligerito_flock.rs:1520-1549)
    tr.absorb_slice(S.POLICY_LABEL)
    tr.absorb_slice(P.ladder_digest(b.cfg))
    P.ood_round_header(tr, mp, b.ood_bits)
    opt, ood_y, r0_nonce = round0(tr)
    tr.absorb_slice(b"\x50" + g16(ood_y))
    # the runtime prime (prime.rs:233-262); the initial boundary is at
    # difficulty 0 and absorbs nothing
    S.spartan_message(tr, b"prime-domain", PRIME_DOMAIN)
    S.spartan_message(tr, b"log-instance-capacity", u64(k))
    S.spartan_message(tr, b"prime-min", S.u128b(b.pmin))
    S.spartan_message(tr, b"prime-max", S.u128b(b.pmax))
    q = P.sample_prime_context(tr, b.pmin, b.pmax, 128)
    S.spartan_message(tr, b"prime-q", S.u128b(q))
    S.spartan_message(tr, b"runtime-field-modulus", S.u128b(q))
    S.spartan_message(tr, b"projected-linear-relation", b.rel.digest)
    xi = challenge_point(tr, LOCAL_POINT_DOMAIN, LOCAL_VARS, q)
    eta = challenge_point(tr, INSTANCE_POINT_DOMAIN, k, q)
    alpha0 = challenge_point(tr, CONSTANT_ONE_DOMAIN, 1, q)[0]
    lam = challenge_point(tr, PUBLIC_IO_DOMAIN, PUBLIC_BITS + 1, q)
    alpha_pub = lam.pop()
    # the rank-one batching (chain.rs:1080-1161)
    coef = collapse(b.rel, S.eq_le(xi, q), q)
    coef[0] = (coef[0] + alpha0) % q
    for slot in range(PUBLIC_BITS):
        c = chain_public_h_column(slot)
        coef[c] = (coef[c] + alpha_pub * lam[slot]) % q
    u = S.eq_le(eta, q)[:n]
    mu = alpha0 * (sum(u) % q) % q
    for i in range(n):
        words = public_words(blocks, digest, i)
        val = 0
        for ws, word in enumerate(words):
            for bit in range(32):
                if (word >> bit) & 1:
                    val += lam[32 * ws + bit]
        mu = (mu + u[i] * alpha_pub % q * (val % q)) % q
    # chain_product_opening_claim (chain.rs:1201-1251)
    th, sh = b.h
    row_w = S.eq_le(eta[:th], q)
    high = S.eq_le(eta[th:], q)
    hi_n = len(high)
    col_w = [0] * (1 << sh)
    for c in range(min(1 << sh, CHAIN_H_BAR * hi_n)):
        col_w[c] = coef[c // hi_n] * high[c % hi_n] % q
    # the claim frame (linear.rs:324-368)
    hd = O.Blake3()
    hd.update(CLAIM_DOMAIN + binding)
    for pt in (xi, eta):
        hd.update(u64(len(pt)) + b"".join(S.u128b(x) for x in pt))
    hd.update(S.u128b(alpha0) + u64(len(lam)) + b"".join(S.u128b(x) for x in lam) + S.u128b(alpha_pub))
    hd.update(u64(len(row_w)) + b"".join(S.u128b(x) for x in row_w))
    hd.update(u64(len(col_w)) + b"".join(S.u128b(x) for x in col_w))
    hd.update(S.u128b(mu))
    claim_digest = hd.digest(32)
    S.spartan_message(tr, CLAIM_TAG, claim_digest)
    tag = tr.challenge(16) + tr.challenge(16)
    return tr, dict(q=q, xi=xi, eta=eta, alpha0=alpha0, lam=lam, alpha_pub=alpha_pub, coef=coef, mu=mu,
                    row_w=row_w, col_w=col_w, binding=binding, stmt=stmt, claim_digest=claim_digest, tag=tag,
                    ood_point=opt, ood_y=ood_y, r0_nonce=r0_nonce)


def transpose_weights(b: ChainBatch, u1, u2) -> np.ndarray:
    """VirtualStatement::transpose_query (src/bitz/virt.rs:153-199) over the
    chained map's column_rows (binary.rs:831-876): weight[src] = sum over
    the derived cells d of src of u1[d mod 2^t] u2[d div 2^t], as a Gf128
    array over the 2^m committed cells"""
    n, th = b.n, b.h[0]
    U1, U2 = P.gfa(u1), P.gfa(u2)
    out = np.zeros((1 << sum(b.f), 2), dtype=np.uint64)
    inst = np.arange(n, dtype=np.int64)
    for name, src_i, dst_i in (("local", inst, inst), ("prev", inst[:-1], inst[1:]),
                               ("first", inst[:1], inst[:1]), ("last", inst[-1:], inst[-1:])):
        mp = b.rel.maps[name]
        off = np.asarray(mp["offsets"], dtype=np.int64)
        fcol = np.repeat(np.arange(CHAIN_F_BAR, dtype=np.int64), np.diff(off))
        hrow = np.asarray(mp["indices"], dtype=np.int64)
        if len(hrow) == 0:
            continue
        src = np.where(fcol[:, None] == 0, 0, 1 + src_i[None, :] * CHAIN_F_INSTANCE + fcol[:, None] - 1)
        d = hrow[:, None] * n + dst_i[None, :]
        src, d = src.reshape(-1), d.reshape(-1)
        prod = O.mul_arrays(U1[d & ((1 << th) - 1)], U2[d >> th])
        np.bitwise_xor.at(out[:, 0], src, prod[:, 0])
        np.bitwise_xor.at(out[:, 1], src, prod[:, 1])
    return out


def virtual_frame(b: ChainBatch, q: int) -> bytes:
    """VirtualStatement::frame (virt.rs:132-139): the claim parameters on
    the derived grid, the committed shape, the map digest"""
    th, sh = b.h
    return u64(th) + u64(sh) + g16(q) + g16(P.generator()) + u64(b.f[0]) + u64(b.f[1]) + b.map_digest


def _pcs40(b: ChainBatch) -> bytes:
    return u64(sum(b.f)) + u64(b.cfg["levels"][0]["rate"]) + u64(b.cfg["initial_k"]) + u64(0) + u64(1)


class ChainProver:
    """prove_sha256_chain -> prove_linear (linear.rs:371-582) on the
    Blake3Transcript, the opening on this build's inner transcript."""

    def __init__(self, k: int, blocks, flip_hint=None):
        self.batch = ChainBatch(k)
        self.wit = ChainWitness(self.batch, blocks, flip_hint)
        tf, sf = self.batch.f
        self.pcs = P.OracleProver(self.wit.source_rows, tf, sf, 1, 4)
        assert self.pcs.cfg == self.batch.cfg

    def _round0(self, tr):
        bits = self.batch.ood_bits
        nonce = None
        if bits:
            seed = P.grinding_seed(tr, P.OOD_GRIND_DOMAIN, 0, bits)
            nonce = O.pow_find(seed[:16], seed[16:], bits)
            P.absorb_grinding_nonce(tr, nonce)
        opt = P.ood_point(P.gint(tr.challenge(16)), self.pcs.mp)
        return opt, O.mle_eval(self.pcs.packed, P.gfa(opt)), nonce

    def prove(self, with_opening: bool = True):
        """returns (outer transcript, dict of the PIOP values; "proof": the
        opening's bytes when with_opening)"""
        b, w = self.batch, self.wit
        tr, out = _outer(b, self.pcs.root, w.blocks, w.digest, self.pcs.mp, self._round0)
        if with_opening:
            out["proof"] = self.open(out)
        return tr, out

    def open(self, out) -> bytes:
        """prove_virtual_opening (bitz_opener.rs:1597-1624) ->
        BitZProver::prove_virtual (virt.rs:214-301)"""
        b = self.batch
        th, sh = b.h
        it = P.Inner(VIRTUAL_SESSION, out["tag"])
        it.public(VIRTUAL_LABEL)
        it.public(self.pcs.root)
        it.public(virtual_frame(b, out["q"]))
        it.public(_pcs40(b))
        it.public(P.claim_encoding(out["row_w"], out["col_w"], out["mu"]))
        folds = O.gf_to_ints(O.fold_columns(self.wit.product_rows, th, sh, P.gfa(out["row_w"])))
        u1, u2, target = self.pcs.reduce_claim(it, out["row_w"], folds, self.wit.product_rows, th, sh)
        wf = O.gf_to_ints(transpose_weights(b, u1, u2))
        return self.pcs.prove_lin(it, wf, [1], target, self.pcs.rows, self.pcs.m, 0, False, out["ood_point"],
                                  out["ood_y"], out["r0_nonce"])


def verify_chain(k: int, blocks, digest, root: bytes, proof: bytes) -> P.Transcript:
    """verify_sha256_chain (chain.rs:850-872) -> verify_linear: replays the
    outer transcript from the statement, the root and the proof's Round-0
    value, then checks the virtual opening; raises P.VerifyError, returns the
    outer transcript"""
    b = ChainBatch(k)
    if len(blocks) != b.n:
        raise P.VerifyError("block count")
    narg, hints, ood_y, r0_nonce = P.proof_tail_parse(proof)
    if ood_y is None:
        raise P.VerifyError("missing Round-0 claim")
    mp = sum(b.f) - P.LOG_PACKING

    def round0(tr):
        bits = b.ood_bits
        if bits:
            if r0_nonce is None:
                raise P.VerifyError("missing Round-0 nonce")
            seed = P.grinding_seed(tr, P.OOD_GRIND_DOMAIN, 0, bits)
            ok = O.pow_valid(seed[:16], seed[16:], r0_nonce, bits)
            P.absorb_grinding_nonce(tr, r0_nonce)
            if not ok:
                raise P.VerifyError("Round-0 proof of work")
        elif r0_nonce is not None:
            raise P.VerifyError("unexpected Round-0 nonce")
        return P.ood_point(P.gint(tr.challenge(16)), mp), ood_y, r0_nonce

    tr, out = _outer(b, root, [[int(x) for x in blk] for blk in blocks], [int(x) for x in digest], mp, round0)
    th, sh = b.h
    it = P.Inner(VIRTUAL_SESSION, out["tag"], narg)
    it.public(VIRTUAL_LABEL)
    it.public(root)
    it.public(virtual_frame(b, out["q"]))
    it.public(_pcs40(b))
    it.public(P.claim_encoding(out["row_w"], out["col_w"], out["mu"]))
    u1, u2, target = P.verify_reduce(it, th, sh, out["q"], out["row_w"], out["col_w"], out["mu"])
    wf = O.gf_to_ints(transpose_weights(b, u1, u2))
    P.verify_lin(it, root, sum(b.f), 0, b.cfg, wf, [1], target, False, out["ood_point"], ood_y, narg, hints)
    return tr
