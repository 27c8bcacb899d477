"""CPU restatement of the BitZ standalone opening: prover and verifier.

TEST INFRASTRUCTURE ONLY (see oracle.py). The prover restates the same
protocol the GPU library runs -- src/piop/spartan/protocol/bitz_opener.rs:672-747
(prove/verify_standalone) -> src/bitz/mod.rs:142-255 (BitZProver::prove,
BitZVerifier::verify) -> src/bitz/pcs.rs:311-507 (prove_lin, prove_mle,
verify_lin) -> vendor/flock-mod/crates/flock-core/src/pcs/ligerito.rs:3398-4420
(recursive prover / verifier) -- on the CPU oracle's primitives, so its proof
bytes can be compared with the GPU's. The verifier accepts or rejects a proof
from either.

Transcripts: the outer transcript is the reference's Blake3Transcript
(src/transcript/mod.rs:18-96). The inner BitZ transcript keeps the reference's
message order, labels and encodings on the same BLAKE3 framing, in place of
the spongefish 0.7.4 sponge (absent from the reference tree).
"""
from __future__ import annotations

import math
import struct

import numpy as np

import oracle as O

LOG_PACKING = 7
POW_PREFIX_PCS = b"bitz-pcs-pow-v1\0"


def u64(v: int) -> bytes:
    return int(v).to_bytes(8, "little")


def u32(v: int) -> bytes:
    return int(v).to_bytes(4, "little")


def g16(v: int) -> bytes:
    return int(v).to_bytes(16, "little")


def gint(b: bytes) -> int:
    return int.from_bytes(b, "little")


# --------------------------------------------------------------- transcripts
class Transcript:
    """Blake3Transcript: a BLAKE3 hash over the framed byte stream."""

    def __init__(self):
        self.buf = bytearray()
        self._h = O.Blake3()

    def absorb_inner(self, b: bytes) -> None:
        b = bytes(b)
        self.buf += b
        self._h.update(b)

    def absorb_slice(self, b: bytes) -> None:
        self.absorb_inner(b"\x06" + u64(len(b)) + bytes(b) + b"\x07")

    def challenge(self, n: int = 16) -> bytes:
        out = self._h.digest(n)
        self.absorb_inner(b"\x12" + out + b"\x34")
        return out

    def sampling_bytes(self, n: int) -> bytes:
        out = self._h.digest(n)
        self.absorb_inner(b"bitz/sampling-read/v1" + u64(n) + out)
        return out


class Inner:
    """The inner BitZ transcript: narg string (absorbed) + hint stream."""

    def __init__(self, session: bytes, instance: bytes, narg: bytes | None = None):
        self.t = Transcript()
        self.t.absorb_slice(b"bitz/v1")
        self.t.absorb_slice(session)
        self.t.absorb_slice(instance)
        self.narg = bytearray()
        self.hints = bytearray()
        self.read_narg = narg
        self.pos = 0

    def public(self, b: bytes) -> None:
        self.t.absorb_slice(b)

    def prover(self, b: bytes) -> None:
        self.t.absorb_slice(b)
        self.narg += b

    def read(self, n: int) -> bytes:
        """Verifier side: read and absorb the next n narg bytes."""
        if self.pos + n > len(self.read_narg):
            raise VerifyError("narg string exhausted")
        b = bytes(self.read_narg[self.pos:self.pos + n])
        self.pos += n
        self.t.absorb_slice(b)
        return b

    def gf(self) -> int:
        return gint(self.t.challenge(16))


class VerifyError(Exception):
    pass


# ------------------------------------------------------------ Ligerito ladder
# The embedded slim template every custom Johnson ladder starts from
# (vendor/flock-mod/crates/flock-core/configs/ligerito/m{22..35}_slim.toml:
# identical header strings, eta, query grinding and target at every m); the
# hash is forced to BLAKE3 (src/ligerito_flock/configuration.rs:146).
TEMPLATE = dict(analysis_version="johnson_ood_row_union_over_bchks25_thm_4_6", field="f128",
                hash="blake3", eta=0.02, grinding=16, target=100)


def _lig_log_a(r: int, eta: float, mc: int, il: int) -> float:
    """paper_johnson_log_a (pcs/ligerito.rs:700-740)."""
    rho = 2.0 ** -r
    sr = math.sqrt(rho)
    gamma = 1.0 - sr - eta
    m_param = float(max(math.ceil(sr / (2.0 * eta)), 3))
    half = m_param + 0.5
    num = 2.0 * half ** 5 + 3.0 * half * gamma * rho
    den = 3.0 * rho ** 1.5
    n = 2.0 ** (mc + r)
    return math.log2((num / den) * n + half / sr) + max(il - 1.0, 0.0)


def _lig_per_query(r: int, eta: float) -> float:
    """paper_per_query_bits (pcs/ligerito.rs:743-747)."""
    gamma = 1.0 - math.sqrt(2.0 ** -r) - eta
    return math.log2(1.0 / (1.0 - gamma))


def _lig_ood_bits(r: int, eta: float, mu: int, samples: int) -> float:
    """paper_ood_bits (pcs/ligerito.rs:820-853)."""
    log2_l = math.log2(1.0 / (2.0 * eta * math.sqrt(2.0 ** -r)))
    log2_mu = math.log2(float(mu))
    if samples == 0:
        return 128.0 - log2_l - log2_mu
    return float(samples) * (128.0 - log2_mu) - (2.0 * log2_l - 1.0)


def custom_johnson_config(m: int, r0: int, k0: int, target: int = 100) -> dict:
    """try_custom_johnson_config_bits (src/ligerito_flock.rs:535-643) on the
    slim template, with the level diagnostics the security config carries
    (expected_eps_pg/query/ood_bits: unrounded f64, as the Rust assigns them)
    and Round 0 (ood_round_bits / ood_round_params, :1384-1413)."""
    if not 20 <= m <= 35 or not 1 <= r0 <= 4:
        raise ValueError("custom Johnson ladders need m in [20, 35] and rate exponent in [1, 4]")
    log_n = m - LOG_PACKING
    if not 0 < k0 < log_n:
        raise ValueError("custom initial_k out of range")
    eta, grinding = TEMPLATE["eta"], TEMPLATE["grinding"]
    shapes = [(log_n - k0, k0, k0, r0)]
    n_run, rate = log_n - k0, r0
    while n_run > 5:
        kr = min(3, n_run)
        rate += 1
        shapes.append((n_run - kr, kr, kr, rate))
        n_run -= kr
    levels = []
    for i, (mc, il, kr, r) in enumerate(shapes):
        need = float(target - grinding)
        pq = _lig_per_query(r, eta)
        q = next((q for q in range(1, 10001) if float(q) * pq + 1e-3 >= need), None)
        if q is None:
            raise ValueError("Johnson query search did not converge")
        if q > 1 << (mc + r):
            raise ValueError("custom Johnson level is too thin for its query count")
        pg = 128.0 - _lig_log_a(r, eta, mc, il)
        qb = float(q) * pq
        ood = 0 if i == 0 else 1
        while True:
            ob = _lig_ood_bits(r, eta, mc + il, ood)
            if i == 0 or ob + 1e-3 >= float(target):
                break
            ood += 1
        levels.append(dict(rate=r, msg_cols=mc, lanes=il, k=kr, queries=q, grinding=grinding,
                           fold_grinding=int(max(math.ceil(float(target) - pg), 0.0)), ood=ood,
                           eps_pg=pg, eps_query=qb, eps_ood=ob))
    rho = 2.0 ** -r0
    lst = 1.0 / (2.0 * eta * math.sqrt(rho))
    pairs = max(lst * (lst - 1.0) / 2.0, 1.0)
    bits0 = 128.0 - math.log2(pairs) - math.log2(max(2.0 ** log_n - 1.0, 1.0))
    deficit = float(target) - bits0
    r0grind = 0 if deficit <= 0.0 else int(math.ceil(deficit))
    return dict(m=m, log_n=log_n, initial_k=k0, yr_log_n=n_run, target=target, levels=levels,
                round0_grinding=r0grind)


def _bc_str(s: str) -> bytes:
    return u64(len(s)) + s.encode()


def _bc_f64(x: float) -> bytes:
    return struct.pack("<d", x)


def security_config_bincode(cfg: dict) -> bytes:
    """bincode 1.3 (fixint, little endian) of LigeritoSecurityConfig
    (pcs/ligerito.rs:541-651): usize as u64, strings and vectors with a u64
    length, unit enum variants as u32 indices, Option as a 0/1 byte."""
    out = u64(cfg["m"]) + u64(cfg["log_n"]) + u64(cfg["initial_k"]) + u64(cfg["target"])
    out += _bc_str(TEMPLATE["analysis_version"]) + _bc_str(TEMPLATE["field"]) + _bc_str(TEMPLATE["hash"])
    out += u32(0)  # GrindingStep::PostCommitPreQueries
    out += u64(len(cfg["levels"]))
    for lv in cfg["levels"]:
        out += u64(lv["rate"]) + u64(lv["msg_cols"]) + u64(lv["lanes"]) + u64(lv["k"])
        out += u32(1)  # SoundnessRegime::JohnsonOod
        out += b"\x01" + _bc_f64(TEMPLATE["eta"])  # eta: Some
        out += b"\x00"  # proximity_loss: None
        out += u64(lv["queries"]) + u64(lv["grinding"]) + u64(lv["fold_grinding"]) + u64(lv["ood"])
        out += u64(cfg["target"])
        out += _bc_f64(lv["eps_pg"]) + _bc_f64(lv["eps_query"])
        out += b"\x01" + _bc_f64(lv["eps_ood"])
    out += u64(cfg["yr_log_n"])
    return out


def ladder_digest(cfg: dict) -> bytes:
    """ResolvedLigerito::digest: BLAKE3 of the security config's bincode
    (src/ligerito_flock/configuration.rs:157-159)."""
    return O.blake3(security_config_bincode(cfg))


# --------------------------------------------------------- statement frames
FRAME_DOMAIN = b"bitz/ligerito-flock/statement-frame/v1"
F_BYTES, F_U8, F_U64, F_U128, F_GF128 = 1, 2, 3, 4, 5


class StatementFrame:
    """StatementFrame (src/ligerito_flock.rs:1057-1160): two framed domain
    slices, then one 0x06 || tag || kind || count u64 || payload || 0x07 run
    per field, absorbed unframed."""

    def __init__(self, tr: Transcript, domain: bytes):
        self.tr = tr
        tr.absorb_slice(FRAME_DOMAIN)
        tr.absorb_slice(domain)

    def field(self, tag: int, kind: int, count: int, payload: bytes) -> None:
        self.tr.absorb_inner(b"\x06" + bytes([tag, kind]) + u64(count) + payload + b"\x07")

    def usize(self, tag: int, v: int) -> None:
        self.field(tag, F_U64, 1, u64(v))

    def usizes(self, tag: int, vs) -> None:
        self.field(tag, F_U64, len(vs), b"".join(u64(v) for v in vs))

    def byte(self, tag: int, v: int) -> None:
        self.field(tag, F_U8, 1, bytes([v]))


def bind_standalone_statement(tr: Transcript, root: bytes, t: int, s: int, cfg: dict) -> None:
    """bind_standalone_statement (bitz_opener.rs:589-609) ->
    absorb_standalone_mod_q_statement (src/ligerito_flock.rs:1245-1276): the
    commitment (root, m, rate, batch, profile Fast = 0, BLAKE3 = 1), the
    verifier configuration, the shape, the generator, the prime width and the
    Round-0 word, then the session label and the ladder digest."""
    lv = cfg["levels"]
    f = StatementFrame(tr, b"bitz/ligerito-flock/standalone-mod-q-statement/v1")
    f.field(0x01, F_BYTES, len(root), root)
    f.usize(0x02, cfg["m"])
    f.usize(0x03, lv[0]["rate"])
    f.usize(0x04, cfg["initial_k"])
    f.byte(0x05, 0)
    f.byte(0x06, 1)
    f.usizes(0x08, [x["rate"] for x in lv])
    f.usize(0x09, len(lv) - 1)
    f.usize(0x0a, lv[0]["msg_cols"])
    f.usize(0x0b, cfg["initial_k"])
    f.usize(0x0c, cfg["initial_k"])
    f.usizes(0x0d, [x["msg_cols"] for x in lv[1:]])
    f.usizes(0x0e, [x["k"] for x in lv[1:]])
    f.usizes(0x0f, [x["queries"] for x in lv])
    f.usizes(0x10, [x["grinding"] for x in lv])
    f.usizes(0x11, [x["fold_grinding"] for x in lv])
    f.usizes(0x12, [x["ood"] for x in lv])
    f.byte(0x13, 1)
    f.usize(0x20, t)
    f.usize(0x21, s)
    f.usize(0x22, 1)
    f.field(0x30, F_GF128, 1, g16(generator()))
    f.usize(0x31, standalone_q_bits(t))
    f.usize(0x32, 1 + cfg["round0_grinding"])
    tr.absorb_slice(b"bitz/bitz-opener/standalone/v1")
    tr.absorb_slice(ladder_digest(cfg))


def standalone_q_bits(t: int) -> int:
    """standalone_q_bits (src/ligerito_flock.rs:1286-1293)."""
    return min(113, 126 - t)


def absorb_standalone_claim(tr: Transcript, q: int, claimed: int) -> None:
    """absorb_standalone_mod_q_claim (src/ligerito_flock.rs:1278-1284)."""
    f = StatementFrame(tr, b"bitz/ligerito-flock/standalone-mod-q-claim/v1")
    f.field(0x30, F_U128, 2, g16(q) + g16(claimed))


# ------------------------------------------------------ Round-0 grinding
GRIND_DOMAIN = b"bitz/spartan/fiat-shamir-grinding/v1"
GRIND_NONCE_DOMAIN = b"bitz/spartan/fiat-shamir-grinding/nonce/v1"
OOD_GRIND_DOMAIN = b"bitz/core/ood-round-grinding/v1"


def grinding_seed(tr: Transcript, domain: bytes, index: int, bits: int) -> bytes:
    """derive_grinding_seed_in_domain (src/piop/spartan/grinding.rs:145-160):
    the frame, the domain, the round index u64 and the difficulty u32, then
    a 32-byte challenge."""
    if not 1 <= bits <= 256:
        raise ValueError("grinding difficulty out of range")
    tr.absorb_slice(GRIND_DOMAIN)
    tr.absorb_slice(domain)
    tr.absorb_slice(u64(index))
    tr.absorb_slice(u32(bits))
    return tr.challenge(32)


def absorb_grinding_nonce(tr: Transcript, nonce: int) -> None:
    """absorb_grinding_nonce (grinding.rs:272-275)."""
    tr.absorb_slice(GRIND_NONCE_DOMAIN)
    tr.absorb_slice(u64(nonce))


def ood_round_header(tr: Transcript, packed_vars: int, bits: int) -> None:
    """absorb_ood_round_header (src/ligerito_flock.rs:1488-1496)."""
    tr.absorb_slice(b"bitz/core/ood-round/v1")
    tr.absorb_slice(u64(packed_vars))
    tr.absorb_slice(u32(bits))


# --------------------------------------------------------- integers mod q
SMALL_PRIMES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53)


def prime_budgets(security: int, max_candidates: int) -> tuple[int, int]:
    """PrimeSearchPolicy::budgets (vendor/field/src/prime/sampling.rs:33-56):
    Miller-Rabin rounds and the rejection-draw budget."""
    if max_candidates == 0:
        raise ValueError("invalid prime search policy")
    sec = max(security + 16, 128)
    cand_bits = (max_candidates - 1).bit_length()
    rounds = (sec + cand_bits + 1) // 2
    draws = sec + cand_bits + rounds.bit_length()
    return rounds, draws


def sample_below(tr: Transcript, bound: int, max_draws: int) -> int:
    """sample_below (sampling.rs:110-138): masked rejection over whole
    little-endian u64 words, each read from the sampling stream."""
    if bound == 0:
        raise ValueError("empty range")
    bits = (bound - 1).bit_length()
    if bits == 0:
        return 0
    words = (bits + 63) // 64
    for _ in range(max_draws):
        cand = 0
        for w in range(words):
            cand |= gint(tr.sampling_bytes(8)) << (64 * w)
        cand &= (1 << bits) - 1
        if cand < bound:
            return cand
    raise RuntimeError("sampling exhausted")


def _strong_round(n: int, base: int, odd: int, twos: int) -> bool:
    x = pow(base, odd, n)
    if x == 1 or x == n - 1:
        return True
    for _ in range(1, twos):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def sample_prime_context(tr: Transcript, lo: int, hi: int, security: int = 128) -> int:
    """sample_prime_context (src/prime_sampling.rs:66-111) ->
    sample_prime_public (vendor/field/src/prime/sampling.rs:290-345): the
    interval and policy absorbed, then odd candidates 2(first + i) + 1 with
    i drawn below the odd count, trial division by the primes to 53, and
    `rounds` Miller-Rabin rounds with bases drawn from the stream."""
    first_odd = max(lo, 3) | 1
    if first_odd > hi:
        raise ValueError("no odd candidate")
    attempts = 1 if first_odd == hi - (1 - (hi & 1)) else 64 * hi.bit_length()
    rounds, draws = prime_budgets(security, attempts)
    tr.absorb_slice(b"bitz/shared-prime-sampling/v1")
    tr.absorb_slice(g16(lo))
    tr.absorb_slice(g16(hi))
    tr.absorb_slice(u32(security))
    tr.absorb_slice(u64(attempts))
    tr.absorb_slice(u64(draws))
    low = max(lo, 3)
    first = low >> 1
    last = hi >> 1 if hi & 1 else (hi >> 1) - 1
    count = last - first + 1
    for _ in range(attempts):
        cand = ((first + sample_below(tr, count, draws)) << 1) | 1
        small = None
        for p in SMALL_PRIMES:
            if cand == p:
                small = True
                break
            if cand % p == 0:
                small = False
                break
        if small is True:
            return cand
        if small is False:
            continue
        odd, twos = cand - 1, 0
        while odd % 2 == 0:
            odd //= 2
            twos += 1
        ok = True
        for _ in range(rounds):
            base = sample_below(tr, cand - 3, draws) + 2
            if not _strong_round(cand, base, odd, twos):
                ok = False
                break
        if ok:
            return cand
    raise RuntimeError("prime search exhausted")


def sample_prime(tr: Transcript, bits: int) -> int:
    """sample_prime_with_bits (src/prime_sampling.rs:125-137)."""
    if not 32 <= bits <= 120:
        raise ValueError("prime bit width must be in [32, 120]")
    top = 1 << (bits - 1)
    return sample_prime_context(tr, top, top | (top - 1), 128)


def sample_residue(tr: Transcript, q: int) -> int:
    """transcript_uniform_mod (src/prime_sampling.rs:11-33)."""
    lo = gint(tr.challenge(16))
    hi = gint(tr.challenge(16))
    return ((hi << 128) | lo) % q


def eq_table_mod_q(r: list[int], q: int) -> list[int]:
    """src/ligerito_flock.rs:1330-1347: each coordinate becomes the new lowest bit."""
    t = [1 % q]
    for c in r:
        nx = []
        for v in t:
            v1 = v * c % q
            nx += [(v - v1) % q, v1]
        t = nx
    return t


def smallest_generator() -> int:
    """src/pcs.rs:85-102."""
    factors = [3, 5, 17, 257, 641, 65537, 274177, 6700417, 67280421310721]
    order = (1 << 128) - 1
    n = 2
    while True:
        if all(O.power(n, order // p) != 1 for p in factors):
            return n
        n += 1


_GEN = None


def generator() -> int:
    global _GEN
    if _GEN is None:
        _GEN = smallest_generator()
    return _GEN


# ---------------------------------------------------------------- helpers
def gfa(xs) -> np.ndarray:
    return O.ints_to_gf(list(xs))


def eq_ints(r) -> list[int]:
    return O.gf_to_ints(O.eq_table(gfa(r))) if len(r) else [1]


def eq_factor(r: int, z: int) -> int:
    return O.mul(r, z) ^ O.mul(1 ^ r, 1 ^ z)


def rows_to_packed(rows: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1, 2)


def batch_claims(claims: list[int], bw: list[int]) -> int:
    return O.batch_claims(gfa(claims), gfa(bw))


def ceil_log2(n: int) -> int:
    k = 0
    while (1 << k) < n:
        k += 1
    return k


def sample_vec(it: Inner, n: int) -> list[int]:
    it.public(b"pcs/flock/sample-vector/v1")
    it.public(u64(n))
    return [it.gf() for _ in range(n)]


def sample_queries(it: Inner, block_len: int, count: int) -> list[int]:
    seen, out = set(), []
    while len(out) < count:
        q = (it.gf() & (2**64 - 1)) % block_len
        if q not in seen:
            seen.add(q)
            out.append(q)
    return sorted(out)


def proof_tail_parse(proof: bytes):
    pos = 0

    def take(n):
        nonlocal pos
        if pos + n > len(proof):
            raise VerifyError("truncated proof")
        b = proof[pos:pos + n]
        pos += n
        return b
    narg = take(gint(take(8)))
    hints = take(gint(take(8)))
    ood_y = None
    nonce = None
    if take(1) == b"\x01":
        ood_y = gint(take(16))
        if take(1) == b"\x01":
            nonce = gint(take(8))
    if pos != len(proof):
        raise VerifyError("trailing proof bytes")
    return narg, hints, ood_y, nonce


# ---------------------------------------------------------------- statement
def outer_statement(root: bytes, t: int, s: int, cfg: dict) -> Transcript:
    """The outer transcript up to Round 0: the bound statement, then Round
    0's header (prove_standalone, bitz_opener.rs:672-685)."""
    tr = Transcript()
    bind_standalone_statement(tr, root, t, s, cfg)
    ood_round_header(tr, t + s - LOG_PACKING, cfg["round0_grinding"])
    return tr


def ood_point(zeta: int, n: int) -> list[int]:
    pt, cur = [], zeta
    for _ in range(n):
        pt.append(cur)
        cur = O.mul(cur, cur)
    return pt


def instance(tr: Transcript, t: int, s: int):
    """sample_standalone_instance (src/ligerito_flock.rs:1305-1327)."""
    q = sample_prime(tr, standalone_q_bits(t))
    r1 = [sample_residue(tr, q) for _ in range(t)]
    r2 = [sample_residue(tr, q) for _ in range(s)]
    return q, eq_table_mod_q(r1, q), eq_table_mod_q(r2, q)


def claim_encoding(row_w, col_w, target) -> bytes:
    return (u64(len(row_w)) + b"".join(g16(w) for w in row_w) + u64(len(col_w)) +
            b"".join(g16(w) for w in col_w) + g16(target))


# =================================================================== prover
class OracleProver:
    """CPU prover of the standalone opening over rows (column-major u64
    words, 2^s columns x 2^(t-6) words)."""

    def __init__(self, rows: np.ndarray, t: int, s: int, log_inv_rate: int, initial_k: int):
        self.rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
        self.t, self.s = t, s
        self.m = t + s
        self.mp = self.m - LOG_PACKING
        self.r0, self.k0 = log_inv_rate, initial_k
        self.cfg = custom_johnson_config(self.m, log_inv_rate, initial_k)
        self.packed = rows_to_packed(self.rows)
        self.root, self.cw, self.tree = O.commit(self.packed, self.mp, initial_k, log_inv_rate)

    def _folds(self, row_w):
        return O.gf_to_ints(O.fold_columns(self.rows, self.t, self.s, gfa(row_w)))

    def _round0(self):
        tr = outer_statement(self.root, self.t, self.s, self.cfg)
        bits = self.cfg["round0_grinding"]
        nonce = None
        if bits:
            # grind_and_absorb (src/piop/spartan/grinding.rs:198-211), the
            # Round-0 boundary (src/ligerito_flock.rs:1523-1537)
            seed = grinding_seed(tr, OOD_GRIND_DOMAIN, 0, bits)
            nonce = O.pow_find(seed[:16], seed[16:], bits)
            absorb_grinding_nonce(tr, nonce)
        zeta = gint(tr.challenge(16))
        pt = ood_point(zeta, self.mp)
        y = O.mle_eval(self.packed, gfa(pt))
        tr.absorb_slice(b"\x50" + g16(y))  # absorb_ood_value (src/ligerito.rs:71-73)
        return tr, pt, y, nonce

    def evaluation(self) -> int:
        tr, _, _, _ = self._round0()
        q, row_w, col_w = instance(tr, self.t, self.s)
        folds = self._folds(row_w)
        return sum(c * (f % q) for c, f in zip(col_w, folds)) % q

    # ---- GKR (dense, gpgkr_prove) ----
    def _gkr(self, it: Inner, zeta: list[int], row_images: list[int], rows=None, t=None, s=None):
        """over the committed grid, or over (rows, t, s) for a virtual
        opening's derived grid"""
        if rows is None:
            rows, t, s = self.rows, self.t, self.s
        leaves = O.forest_leaves(rows, t, s, gfa(row_images))
        layers = [leaves]
        while layers[-1].shape[0] > (1 << s) * 2:
            layers.append(O.product_level(layers[-1]))
        internal = list(reversed(zeta))
        claim = 0
        for w in reversed(layers):
            L = len(internal)
            l, r = w[: 1 << L], w[1 << L:]  # read-only views: the folds write new arrays
            factor, nxt = 1, []
            ext = list(reversed(internal))
            for rnd, z in enumerate(internal):
                rem = L - 1 - rnd
                eq = O.eq_table(gfa(ext[:rem]))
                se, si = O.gkr_round_sums(l, r, eq, z == 0)
                it.prover(g16(O.mul(factor, se)) + g16(O.mul(factor, si)))
                rr = it.gf()
                nxt.append(rr)
                l, r = O.fold_halves(l, rr), O.fold_halves(r, rr)
                factor = O.mul(factor, eq_factor(rr, z))
            el, er = O.gf_to_ints(l)[0], O.gf_to_ints(r)[0]
            it.prover(g16(el) + g16(er))
            rc = it.gf()
            internal = [rc] + nxt
            claim = el ^ O.mul(rc, er ^ el)
        return list(reversed(internal)), claim

    def _sumcheck(self, it: Inner, a, w, target, point):
        while a.shape[0] > 1:
            a0, a1, a2 = O.round_polynomial(a, w)
            assert a1 ^ a2 == target, "inner sumcheck: claim does not match the committed bits"
            it.prover(g16(a0) + g16(a1) + g16(a2))
            r = it.gf()
            target = a0 ^ O.mul(r, a1 ^ O.mul(r, a2))
            point.append(r)
            a, w = O.fold_lsb(a, r), O.fold_lsb(w, r)
        return a, w, target

    def prove(self, claimed: int) -> bytes:
        t, s, mp = self.t, self.s, self.mp
        cfg = self.cfg
        tr, opt, ood_y, r0_nonce = self._round0()
        q, row_w, col_w = instance(tr, t, s)
        folds = self._folds(row_w)
        assert claimed < q
        absorb_standalone_claim(tr, q, claimed)
        tag = tr.challenge(16) + tr.challenge(16)
        it = Inner(b"bitz/bitz-opener/standalone/v1", tag)
        return self.open_claim(it, q, row_w, col_w, claimed, folds, opt, ood_y, r0_nonce)

    def open_claim(self, it: Inner, q: int, row_w, col_w, claimed: int, folds, opt, ood_y: int,
                   r0_nonce) -> bytes:
        """BitZProver::prove (src/bitz/pcs.rs) of the claim
        sum_{c,b} row_w[b] col_w[c] bit(c, b) = claimed (mod q) on the inner
        transcript `it`, with Round 0's point and value (opt, ood_y) already
        bound on the outer one; returns the proof bytes."""
        t, s = self.t, self.s
        it.public(self.root)
        it.public(u64(t) + u64(s) + g16(q) + g16(generator()))
        it.public(claim_encoding(row_w, col_w, claimed))
        u1, u2, target = self.reduce_claim(it, row_w, folds, self.rows, t, s)
        return self.prove_lin(it, u1, u2, target, self.rows, t, s, True, opt, ood_y, r0_nonce)

    def reduce_claim(self, it: Inner, row_w, folds, rows, t: int, s: int):
        """send_fold and the GKR reduction (src/bitz/fold.rs:237-263,
        reduce.rs:28-114) over a grid of bits (t, s): returns the factored
        inner-product claim (u1, u2, target)"""
        for f in folds:
            it.prover(g16(f))
        zeta = [it.gf() for _ in range(s)]
        row_images = O.gf_to_ints(O.pow_many(generator(), gfa(row_w)))
        point, leaf_claim = self._gkr(it, zeta, row_images, rows, t, s)
        alpha_c, alpha_b = point[:s], point[s:]
        eqb = eq_ints(alpha_b)
        u1 = [O.mul(a ^ 1, e) for a, e in zip(row_images, eqb)]
        u2 = eq_ints(alpha_c)
        return u1, u2, leaf_claim ^ 1

    def prove_lin(self, it: Inner, u1, u2, target: int, rows, t: int, s: int, bind: bool, opt, ood_y: int,
                  r0_nonce) -> bytes:
        """Pcs::prove_lin_rows (src/bitz/pcs.rs:337-377) of the inner-product
        claim sum u1[b] u2[c] bit(c, b) = target over the committed bits
        presented as the grid (rows, t, s): the statement (when `bind`), the
        inner sumcheck, the ring switch and Ligerito; returns the proof bytes"""
        pcs40 = u64(self.m) + u64(self.r0) + u64(self.k0) + u64(0) + u64(1)
        if bind:
            it.public(b"bitz/pcs/bit-inner-product/v2")
            it.public(self.root)
            it.public(pcs40)
            it.public(claim_encoding(u1, u2, target))
        it.public(b"bitz/pcs/inner-product-sumcheck/v1")
        comb = O.xi_combined_rows(rows, t, s, gfa(u2))
        sc_point = []
        _, wts, target = self._sumcheck(it, comb, gfa(u1), target, sc_point)
        row_scalar = O.gf_to_ints(wts)[0]
        folded = O.fold_rows_eq(rows, t, s, O.eq_table(gfa(sc_point)))
        cols = gfa([O.mul(row_scalar, c) for c in u2])
        fin, cw_, target = self._sumcheck(it, folded, cols, target, sc_point)
        evaluation = O.gf_to_ints(fin)[0]
        assert target == O.mul(evaluation, O.gf_to_ints(cw_)[0])
        it.prover(g16(evaluation))
        it.public(b"bitz/pcs/mle-opening/v1")
        it.public(self.root)
        it.public(pcs40)
        it.public(u64(len(sc_point)))
        for c in sc_point:
            it.public(g16(c))
        it.public(g16(evaluation))
        # ring switch
        prefix = eq_ints(sc_point[:LOG_PACKING])
        suffix = O.eq_table(gfa(sc_point[LOG_PACKING:]))
        claims = O.gf_to_ints(O.fold_1b_rows(self.packed, suffix))
        cc = 0
        for a, b in zip(prefix, claims):
            cc ^= O.mul(a, b)
        assert cc == evaluation
        it.public(b"bitz/pcs/mle-claims/v1")
        it.prover(b"".join(g16(c) for c in claims))
        it.public(b"bitz/pcs/ring-switch-challenges/v1")
        bw = eq_ints([it.gf() for _ in range(7)])
        packed_target = batch_claims(claims, bw)
        basis = O.fold_b128_elems(suffix, gfa(bw))
        it.public(b"bitz/bitz/round0-claim/v1")
        it.public(u64(len(opt)))
        for c in opt:
            it.public(g16(c))
        it.public(g16(ood_y))
        eta = it.gf()
        basis = O.lig_glue(basis, O.eq_table(gfa(opt)), eta)
        packed_target ^= O.mul(eta, ood_y)
        lig = self._ligerito(it, basis, packed_target)
        it.hints += u32(len(lig)) + lig
        out = u64(len(it.narg)) + bytes(it.narg) + u64(len(it.hints)) + bytes(it.hints)
        out += b"\x01" + g16(ood_y)
        out += (b"\x01" + u64(r0_nonce)) if r0_nonce is not None else b"\x00"
        return out

    # ---- Ligerito recursive prover ----
    def _ligerito(self, it: Inner, basis, target) -> bytes:
        cfg = self.cfg
        lv = cfg["levels"]
        first = [True]
        msgs, grind_nonces, fold_nonces, ood_values = [], [], [], []

        def observe(v):
            if first[0]:
                first[0] = False
                it.public(g16(v))
            else:
                it.prover(g16(v))

        def grind(bits):
            it.public(b"pcs/flock/pow/v1")
            it.public(u32(bits))
            seed = it.t.challenge(16)
            nonce = O.pow_find(POW_PREFIX_PCS, seed, bits)
            it.prover(u64(nonce))
            return nonce

        state = {"f": self.packed.copy(), "b": basis.copy()}

        def msg_obs(u0, u2):
            observe(u0)
            observe(u2)
            msgs.append((u0, u2))

        def fold_round(bits):
            if bits > 0:
                fold_nonces.append(grind(bits))
            r = it.gf()
            state["f"] = O.fold_lsb(state["f"], r)
            state["b"] = O.fold_lsb(state["b"], r)
            msg_obs(*O.lig_round_msg(state["f"], state["b"]))
            return r

        def commit_level(level):
            f = state["f"]
            n = f.shape[0].bit_length() - 1
            L = lv[level]
            return O.commit(f, n, L["lanes"], L["rate"])

        def ood_round(level, n_vars):
            for _ in range(lv[level]["ood"]):
                z = sample_vec(it, n_vars)
                eqz = O.eq_table(gfa(z))
                u0, u2 = O.lig_round_msg(state["f"], eqz)
                y = 0
                for a, c in zip(O.gf_to_ints(state["f"]), O.gf_to_ints(eqz)):
                    y ^= O.mul(a, c)
                observe(y)
                ood_values.append(y)
                msg_obs(u0, u2)
                beta = it.gf()
                state["b"] = O.lig_glue(state["b"], eqz, beta)

        def open_level(cw, tree, block_len, lanes, queries):
            cwr = cw.reshape(block_len, lanes, 2)
            rows = [O.gf_to_ints(cwr[q]) for q in queries]
            proof = O.merkle_multi_proof(tree, block_len, queries)
            return rows, proof

        def induce(n_log, rows, v_ch, queries, alpha):
            aw = eq_ints(alpha)[: len(queries)]
            eqv = eq_ints(v_ch)
            enforced = 0
            for row, a in zip(rows, aw):
                d = 0
                for x, e in zip(row, eqv):
                    d ^= O.mul(x, e)
                enforced ^= O.mul(d, a)
            b_new = O.induce_basis(n_log, queries, gfa(aw))
            msg_obs(*O.lig_round_msg(state["f"], b_new))
            beta = it.gf()
            state["b"] = O.lig_glue(state["b"], b_new, beta)

        it.public(b"flock-ligerito-basis-v0")
        observe(target)
        it.prover(u32(32) + self.root)
        msg_obs(*O.lig_round_msg(state["f"], state["b"]))
        r_lane = [fold_round(max(0, lv[0]["fold_grinding"] - j)) for j in range(cfg["initial_k"])]
        n1 = self.mp - cfg["initial_k"]
        root_prev, cw_prev, tree_prev = commit_level(1)
        it.prover(u32(32) + root_prev)
        roots = [root_prev]
        ood_round(1, n1)
        grind_nonces.append(grind(lv[0]["grinding"]))
        block0 = 1 << (lv[0]["msg_cols"] + lv[0]["rate"])
        q0 = sample_queries(it, block0, lv[0]["queries"])
        alpha0 = sample_vec(it, ceil_log2(lv[0]["queries"]))
        rows0, proof0 = open_level(self.cw, self.tree, block0, 1 << cfg["initial_k"], q0)
        induce(n1, rows0, r_lane, q0, alpha0)
        rec, prev = [], 1
        final = None
        yr = []
        r_steps = len(lv) - 1
        for i in range(r_steps):
            level_rs = [fold_round(max(0, lv[i + 1]["fold_grinding"] - j)) for j in range(lv[i + 1]["k"])]
            Lp = lv[prev]
            block_prev = 1 << (Lp["msg_cols"] + Lp["rate"])
            if i == r_steps - 1:
                yr = O.gf_to_ints(state["f"])
                for v in yr:
                    observe(v)
                grind_nonces.append(grind(lv[i + 1]["grinding"]))
                ql = sample_queries(it, block_prev, lv[i + 1]["queries"])
                final = open_level(cw_prev, tree_prev, block_prev, 1 << Lp["lanes"], ql)
                break
            n_next = state["f"].shape[0].bit_length() - 1
            root_n, cw_n, tree_n = commit_level(i + 2)
            it.prover(u32(32) + root_n)
            roots.append(root_n)
            ood_round(i + 2, n_next)
            grind_nonces.append(grind(lv[i + 1]["grinding"]))
            qi = sample_queries(it, block_prev, lv[i + 1]["queries"])
            alphai = sample_vec(it, ceil_log2(lv[i + 1]["queries"]))
            rows_i, proof_i = open_level(cw_prev, tree_prev, block_prev, 1 << Lp["lanes"], qi)
            induce(n_next, rows_i, level_rs, qi, alphai)
            rec.append((rows_i, proof_i))
            cw_prev, tree_prev, prev = cw_n, tree_n, i + 2
        return serialize_ligerito(self.root, (rows0, proof0), roots, rec, yr, final, msgs,
                                  grind_nonces, ood_values, fold_nonces)


def _ser_open(rows, proof) -> bytes:
    out = u64(len(rows))
    for row in rows:
        out += u64(len(row)) + b"".join(g16(x) for x in row)
    out += u64(len(proof)) + b"".join(bytes(p) for p in proof)
    return out


def serialize_ligerito(root, initial, roots, rec, yr, final, msgs, grind_nonces, ood_values,
                       fold_nonces) -> bytes:
    out = bytes(root) + _ser_open(*initial)
    out += u64(len(roots)) + b"".join(roots)
    out += u64(len(rec)) + b"".join(_ser_open(*r) for r in rec)
    out += u64(len(yr)) + b"".join(g16(v) for v in yr) + _ser_open(*final)
    out += u64(len(msgs)) + b"".join(g16(a) + g16(b) for a, b in msgs)
    out += u64(len(grind_nonces)) + b"".join(u64(n) for n in grind_nonces)
    out += u64(len(ood_values)) + b"".join(g16(v) for v in ood_values)
    out += u64(len(fold_nonces)) + b"".join(u64(n) for n in fold_nonces)
    return out


def with_extra_recursive_level(proof: bytes) -> bytes:
    """The proof with one more recursive root and opening appended to its
    Ligerito hint (copies of the last ones, counts and lengths edited to
    match): a verifier must reject it, every root and opening being bound
    to a level."""
    narg, hints, ood_y, nonce = proof_tail_parse(proof)
    p = parse_ligerito(hints[4:])
    lig = serialize_ligerito(p["initial_root"], p["initial"], p["roots"] + [p["roots"][-1]],
                             p["rec"] + [p["rec"][-1]], p["yr"], p["final"], p["msgs"], p["grind"],
                             p["ood"], p["fold_nonces"])
    hints = u32(len(lig)) + lig
    out = u64(len(narg)) + narg + u64(len(hints)) + hints + b"\x01" + g16(ood_y)
    out += (b"\x01" + u64(nonce)) if nonce is not None else b"\x00"
    return out


class _Reader:
    def __init__(self, b: bytes):
        self.b, self.pos = b, 0

    def take(self, n):
        if self.pos + n > len(self.b):
            raise VerifyError("truncated Ligerito proof")
        out = self.b[self.pos:self.pos + n]
        self.pos += n
        return out

    def u64(self):
        return gint(self.take(8))

    def gf(self):
        return gint(self.take(16))

    def open(self):
        rows = []
        for _ in range(self.u64()):
            rows.append([self.gf() for _ in range(self.u64())])
        proof = [self.take(32) for _ in range(self.u64())]
        return rows, proof


def parse_ligerito(b: bytes) -> dict:
    rd = _Reader(b)
    p = {"initial_root": rd.take(32), "initial": rd.open()}
    p["roots"] = [rd.take(32) for _ in range(rd.u64())]
    p["rec"] = [rd.open() for _ in range(rd.u64())]
    p["yr"] = [rd.gf() for _ in range(rd.u64())]
    p["final"] = rd.open()
    p["msgs"] = [(rd.gf(), rd.gf()) for _ in range(rd.u64())]
    p["grind"] = [rd.u64() for _ in range(rd.u64())]
    p["ood"] = [rd.gf() for _ in range(rd.u64())]
    p["fold_nonces"] = [rd.u64() for _ in range(rd.u64())]
    if rd.pos != len(b):
        raise VerifyError("trailing Ligerito proof bytes")
    return p


# ================================================================= verifier
def verify_standalone(root: bytes, t: int, s: int, log_inv_rate: int, initial_k: int,
                      claimed: int, proof: bytes) -> None:
    """Raises VerifyError unless the proof is accepted (BitZVerifier::verify
    and the succinct Ligerito verifier, with dense residual evaluation)."""
    m, mp = t + s, t + s - LOG_PACKING
    cfg = custom_johnson_config(m, log_inv_rate, initial_k)
    narg, hints, ood_y, r0_nonce = proof_tail_parse(proof)
    if ood_y is None:
        raise VerifyError("missing Round-0 claim")
    tr = outer_statement(root, t, s, cfg)
    bits = cfg["round0_grinding"]
    if bits:
        # verify_and_absorb (grinding.rs:219-236): the nonce is absorbed
        # before the verdict
        if r0_nonce is None:
            raise VerifyError("missing Round-0 nonce")
        seed = grinding_seed(tr, OOD_GRIND_DOMAIN, 0, bits)
        ok = O.pow_valid(seed[:16], seed[16:], r0_nonce, bits)
        absorb_grinding_nonce(tr, r0_nonce)
        if not ok:
            raise VerifyError("Round-0 proof of work")
    elif r0_nonce is not None:
        raise VerifyError("unexpected Round-0 nonce")
    zeta0 = gint(tr.challenge(16))
    opt = ood_point(zeta0, mp)
    tr.absorb_slice(b"\x50" + g16(ood_y))
    q, row_w, col_w = instance(tr, t, s)
    if claimed >= q:
        raise VerifyError("claim not reduced")
    absorb_standalone_claim(tr, q, claimed)
    tag = tr.challenge(16) + tr.challenge(16)
    it = Inner(b"bitz/bitz-opener/standalone/v1", tag, narg)
    verify_claim(it, root, t, s, cfg, q, row_w, col_w, claimed, opt, ood_y, narg, hints)


def verify_claim(it: Inner, root: bytes, t: int, s: int, cfg: dict, q: int, row_w, col_w, claimed: int, opt,
                 ood_y: int, narg: bytes, hints: bytes) -> None:
    """BitZVerifier::verify of the claim sum_{c,b} row_w[b] col_w[c]
    bit(c, b) = claimed (mod q) on the inner transcript `it` (reading the
    narg string), Round 0's point and value bound on the outer one."""
    it.public(root)
    it.public(u64(t) + u64(s) + g16(q) + g16(generator()))
    it.public(claim_encoding(row_w, col_w, claimed))
    u1, u2, target = verify_reduce(it, t, s, q, row_w, col_w, claimed)
    verify_lin(it, root, t, s, cfg, u1, u2, target, True, opt, ood_y, narg, hints)


def verify_reduce(it: Inner, t: int, s: int, q: int, row_w, col_w, claimed: int):
    """receive_fold and gkr_reduce_verify (fold.rs, gkr.rs:337-383,
    reduce.rs:28-58) over a grid (t, s): the factored claim (u1, u2, target)"""
    g = generator()
    folds = [gint(it.read(16)) for _ in range(1 << s)]
    bound = (1 << t) * (q - 1)
    if any(f > bound for f in folds):
        raise VerifyError("fold out of range")
    if sum(c * (f % q) for c, f in zip(col_w, folds)) % q != claimed:
        raise VerifyError("folds do not reconstruct the claim")
    zeta = [it.gf() for _ in range(s)]
    images = O.pow_many(g, gfa(folds))
    claim = O.mle_eval(images, gfa(zeta))
    # GKR (gpgkr_verify, gkr.rs:337-383)
    internal = list(reversed(zeta))
    for _ in range(t):
        prefix, nxt = 1, []
        for z in internal:
            b = it.read(32)
            se, si = gint(b[:16]), gint(b[16:])
            if z == 0:
                s0, s1 = claim, se
            else:
                s0, s1 = se, O.mul(claim ^ O.mul(1 ^ z, se), O.inv(z))
            rr = it.gf()
            nxt.append(rr)
            f = eq_factor(rr, z)
            claim = O.mul(f, s0 ^ O.mul(rr, (s1 ^ s0) ^ O.mul(rr ^ 1, si)))
            prefix = O.mul(prefix, f)
        b = it.read(32)
        el, er = gint(b[:16]), gint(b[16:])
        if O.mul(prefix, O.mul(el, er)) != claim:
            raise VerifyError("GKR layer check")
        rc = it.gf()
        internal = [rc] + nxt
        claim = el ^ O.mul(rc, er ^ el)
    point = list(reversed(internal))
    alpha_c, alpha_b = point[:s], point[s:]
    row_images = O.gf_to_ints(O.pow_many(g, gfa(row_w)))
    u1 = [O.mul(a ^ 1, e) for a, e in zip(row_images, eq_ints(alpha_b))]
    u2 = eq_ints(alpha_c)
    return u1, u2, claim ^ 1


def verify_lin(it: Inner, root: bytes, t: int, s: int, cfg: dict, u1, u2, target: int, bind: bool, opt,
               ood_y: int, narg: bytes, hints: bytes) -> None:
    """Pcs::verify_lin (src/bitz/pcs.rs:379-424) of sum u1[b] u2[c] bit(c, b)
    = target over the committed bits seen as the grid (t, s)"""
    m, mp = cfg["m"], cfg["m"] - LOG_PACKING
    pcs40 = u64(m) + u64(cfg["levels"][0]["rate"]) + u64(cfg["initial_k"]) + u64(0) + u64(1)
    if bind:
        it.public(b"bitz/pcs/bit-inner-product/v2")
        it.public(root)
        it.public(pcs40)
        it.public(claim_encoding(u1, u2, target))
    it.public(b"bitz/pcs/inner-product-sumcheck/v1")
    rows_w, cols_w = gfa(u1), gfa(u2)
    sc_point = []
    for _ in range(t + s):
        b = it.read(48)
        c0, c1, c2 = gint(b[:16]), gint(b[16:32]), gint(b[32:])
        if c1 ^ c2 != target:
            raise VerifyError("inner sumcheck round")
        r = it.gf()
        target = c0 ^ O.mul(r, c1 ^ O.mul(r, c2))
        sc_point.append(r)
        if rows_w.shape[0] > 1:
            rows_w = O.fold_lsb(rows_w, r)
        else:
            cols_w = O.fold_lsb(cols_w, r)
    evaluation = gint(it.read(16))
    if target != O.mul(evaluation, O.mul(O.gf_to_ints(rows_w)[0], O.gf_to_ints(cols_w)[0])):
        raise VerifyError("inner sumcheck final check")
    it.public(b"bitz/pcs/mle-opening/v1")
    it.public(root)
    it.public(pcs40)
    it.public(u64(len(sc_point)))
    for c in sc_point:
        it.public(g16(c))
    it.public(g16(evaluation))
    # ring switch (verify_mle)
    prefix = eq_ints(sc_point[:LOG_PACKING])
    it.public(b"bitz/pcs/mle-claims/v1")
    cb = it.read(128 * 16)
    claims = [gint(cb[16 * i:16 * i + 16]) for i in range(128)]
    cc = 0
    for a, b in zip(prefix, claims):
        cc ^= O.mul(a, b)
    if cc != evaluation:
        raise VerifyError("ring-switch claim check")
    it.public(b"bitz/pcs/ring-switch-challenges/v1")
    bw = eq_ints([it.gf() for _ in range(7)])
    packed_target = batch_claims(claims, bw)
    it.public(b"bitz/bitz/round0-claim/v1")
    it.public(u64(len(opt)))
    for c in opt:
        it.public(g16(c))
    it.public(g16(ood_y))
    eta = it.gf()
    packed_target ^= O.mul(eta, ood_y)
    # the dense initial basis b(x) = sum_i bit_i(suffix[x]) bw[i] + eta eq(opt)[x]
    suffix = O.eq_table(gfa(sc_point[LOG_PACKING:]))
    basis = O.lig_glue(O.fold_b128_elems(suffix, gfa(bw)), O.eq_table(gfa(opt)), eta)
    # the Ligerito proof from the hint stream
    if len(hints) < 4:
        raise VerifyError("missing Ligerito hint")
    ln = gint(hints[:4])
    if 4 + ln != len(hints):
        raise VerifyError("hint length")
    lig = parse_ligerito(hints[4:])
    _verify_ligerito(it, cfg, lig, root, mp, packed_target, basis)
    if it.pos != len(narg):
        raise VerifyError("trailing narg bytes")


def _verify_ligerito(it: Inner, cfg: dict, p: dict, root: bytes, log_n: int, target: int,
                     basis) -> None:
    lv = cfg["levels"]
    k0 = cfg["initial_k"]
    if p["initial_root"] != root:
        raise VerifyError("initial root")
    first = [True]
    msgs = iter(p["msgs"])
    fold_nonces = iter(p["fold_nonces"])
    grind_nonces = iter(p["grind"])
    ood_vals = iter(p["ood"])

    def observe(v):
        if first[0]:
            first[0] = False
            it.public(g16(v))
        else:
            if gint(it.read(16)) != v:
                raise VerifyError("observed value differs from the narg string")

    def next_msg():
        try:
            u0, u2 = next(msgs)
        except StopIteration:
            raise VerifyError("sumcheck transcript exhausted")
        observe(u0)
        observe(u2)
        return u0, u2

    def check_pow(it_nonces, bits):
        it.public(b"pcs/flock/pow/v1")
        it.public(u32(bits))
        seed = it.t.challenge(16)
        try:
            nonce = next(it_nonces)
        except StopIteration:
            raise VerifyError("missing nonce")
        if gint(it.read(8)) != nonce or not O.pow_valid(POW_PREFIX_PCS, seed, nonce, bits):
            raise VerifyError("proof of work")

    def quad(msg, t_r):
        u0, u2 = msg
        return (u0, t_r ^ u2, u2)

    def qeval(qd, r):
        c, b, a = qd
        return c ^ O.mul(r, b) ^ O.mul(O.mul(r, r), a)

    def qfold(p1, p2, alpha):
        return tuple(x ^ O.mul(alpha, y) for x, y in zip(p1, p2))

    def check_root_bytes(expected):
        b = it.read(36)
        if gint(b[:4]) != 32 or b[4:] != expected:
            raise VerifyError("root message")

    def verify_open(rows, proof, root_, block_len, lanes, queries):
        if len(rows) != len(queries) or any(len(r) != lanes for r in rows):
            raise VerifyError("opened rows shape")
        leaves = np.stack([np.frombuffer(O.leaf_cv(b"".join(g16(x) for x in row)), np.uint8)
                           for row in rows])
        pf = np.stack([np.frombuffer(x, np.uint8) for x in proof]) if proof else np.zeros((0, 32), np.uint8)
        if not O.merkle_verify_multi(root_, block_len, queries, leaves, pf):
            raise VerifyError("Merkle multi-proof")

    def enforced_sum(rows, v_ch, queries, alpha):
        aw = eq_ints(alpha)[: len(queries)]
        eqv = eq_ints(v_ch)
        tot = 0
        for row, a in zip(rows, aw):
            d = 0
            for x, e in zip(row, eqv):
                d ^= O.mul(x, e)
            tot ^= O.mul(d, a)
        return tot

    # components of the combined basis: (dense table, first challenge index, weight)
    comps = [(basis, 0, 1)]
    it.public(b"flock-ligerito-basis-v0")
    observe(target)
    check_root_bytes(root)
    t_r = target
    rq = quad(next_msg(), t_r)
    ris = []
    for j in range(k0):
        bits = max(0, lv[0]["fold_grinding"] - j)
        if bits > 0:
            check_pow(fold_nonces, bits)
        r = it.gf()
        ris.append(r)
        t_r = qeval(rq, r)
        rq = quad(next_msg(), t_r)
    if not p["roots"]:
        raise VerifyError("missing recursive roots")
    check_root_bytes(p["roots"][0])
    n_cur = log_n - k0

    def ood_mirror(level, n_vars):
        nonlocal rq, t_r
        for _ in range(lv[level]["ood"]):
            z = sample_vec(it, n_vars)
            try:
                y = next(ood_vals)
            except StopIteration:
                raise VerifyError("missing OOD value")
            observe(y)
            iq = quad(next_msg(), y)
            beta = it.gf()
            rq = qfold(rq, iq, beta)
            t_r ^= O.mul(beta, y)
            comps.append((O.eq_table(gfa(z)), len(ris), beta))

    ood_mirror(1, n_cur)
    check_pow(grind_nonces, lv[0]["grinding"])
    block0 = 1 << (lv[0]["msg_cols"] + lv[0]["rate"])
    q0 = sample_queries(it, block0, lv[0]["queries"])
    a0 = sample_vec(it, ceil_log2(lv[0]["queries"]))
    rows0, proof0 = p["initial"]
    verify_open(rows0, proof0, root, block0, 1 << k0, q0)
    es = enforced_sum(rows0, ris, q0, a0)
    iq = quad(next_msg(), es)
    beta = it.gf()
    rq = qfold(rq, iq, beta)
    t_r ^= O.mul(beta, es)
    comps.append((O.induce_basis(n_cur, q0, gfa(eq_ints(a0)[: len(q0)])), len(ris), beta))
    prev_root, prev = p["roots"][0], 1
    r_steps = len(lv) - 1
    rec = iter(p["rec"])
    next_root = 1
    for i in range(r_steps):
        level_rs = []
        for j in range(lv[i + 1]["k"]):
            bits = max(0, lv[i + 1]["fold_grinding"] - j)
            if bits > 0:
                check_pow(fold_nonces, bits)
            r = it.gf()
            ris.append(r)
            level_rs.append(r)
            t_r = qeval(rq, r)
            rq = quad(next_msg(), t_r)
        n_cur -= lv[i + 1]["k"]
        Lp = lv[prev]
        block_prev = 1 << (Lp["msg_cols"] + Lp["rate"])
        if i == r_steps - 1:
            if next(msgs, None) is not None or next(ood_vals, None) is not None \
                    or next(fold_nonces, None) is not None:
                raise VerifyError("unconsumed Ligerito transcript")
            yr = p["yr"]
            if len(yr) != 1 << n_cur:
                raise VerifyError("yr length")
            for v in yr:
                observe(v)
            check_pow(grind_nonces, lv[i + 1]["grinding"])
            ql = sample_queries(it, block_prev, lv[i + 1]["queries"])
            al = sample_vec(it, ceil_log2(lv[i + 1]["queries"]))
            rows_l, proof_l = p["final"]
            verify_open(rows_l, proof_l, prev_root, block_prev, 1 << Lp["lanes"], ql)
            esl = enforced_sum(rows_l, level_rs, ql, al)
            beta_l = it.gf()
            t_r ^= O.mul(beta_l, esl)
            comps.append((O.induce_basis(n_cur, ql, gfa(eq_ints(al)[: len(ql)])), len(ris), beta_l))
            # dense residual: fold each component by the challenges after it
            combined = [0] * len(yr)
            for tab, start, wgt in comps:
                v = tab
                for r in ris[start:]:
                    v = O.fold_lsb(v, r)
                vals = O.gf_to_ints(v)
                if len(vals) != len(yr):
                    raise VerifyError("residual length")
                for y in range(len(yr)):
                    combined[y] ^= O.mul(wgt, vals[y])
            inner = 0
            for a, c in zip(yr, combined):
                inner ^= O.mul(a, c)
            if inner != t_r:
                raise VerifyError("Ligerito final check")
            if next(grind_nonces, None) is not None:
                raise VerifyError("unconsumed grinding nonces")
            # every recursive root and opening is consumed (the flock
            # verifier rejects recursive_roots.len() != r, ligerito.rs:5298)
            if next_root != len(p["roots"]) or next(rec, None) is not None:
                raise VerifyError("unconsumed recursive roots or openings")
            return
        if next_root >= len(p["roots"]):
            raise VerifyError("missing recursive root")
        root_n = p["roots"][next_root]
        next_root += 1
        check_root_bytes(root_n)
        ood_mirror(i + 2, n_cur)
        check_pow(grind_nonces, lv[i + 1]["grinding"])
        qi = sample_queries(it, block_prev, lv[i + 1]["queries"])
        ai = sample_vec(it, ceil_log2(lv[i + 1]["queries"]))
        try:
            rows_i, proof_i = next(rec)
        except StopIteration:
            raise VerifyError("missing recursive opening")
        verify_open(rows_i, proof_i, prev_root, block_prev, 1 << Lp["lanes"], qi)
        es = enforced_sum(rows_i, level_rs, qi, ai)
        iq = quad(next_msg(), es)
        beta = it.gf()
        rq = qfold(rq, iq, beta)
        t_r ^= O.mul(beta, es)
        comps.append((O.induce_basis(n_cur, qi, gfa(eq_ints(ai)[: len(qi)])), len(ris), beta))
        prev_root, prev = root_n, i + 2
