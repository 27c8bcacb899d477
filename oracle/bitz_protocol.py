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

import numpy as np

import oracle as O

LOG_PACKING = 7
POW_PREFIX_PCS = b"bitz-pcs-pow-v1\0"
POW_PREFIX_NATIVE = b"bitz-native-pow1"


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

    def absorb_slice(self, b: bytes) -> None:
        self.buf += b"\x06" + u64(len(b)) + bytes(b) + b"\x07"

    def challenge(self, n: int = 16) -> bytes:
        out = O.blake3(bytes(self.buf), n)
        self.buf += b"\x12" + out + b"\x34"
        return out

    def sampling_bytes(self, n: int) -> bytes:
        out = O.blake3(bytes(self.buf), n)
        self.buf += b"bitz/sampling-read/v1" + u64(n) + out
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
def custom_johnson_config(m: int, r0: int, k0: int, target: int = 100) -> dict:
    """try_custom_johnson_config_bits (src/ligerito_flock.rs:535-665) with the
    Johnson formulas of pcs/ligerito.rs:700-930 (eta 0.02, grinding 16)."""
    if not 20 <= m <= 35 or not 1 <= r0 <= 4:
        raise ValueError("custom Johnson ladders need m in [20, 35] and rate exponent in [1, 4]")
    log_n = m - LOG_PACKING
    if not 0 < k0 < log_n:
        raise ValueError("custom initial_k out of range")
    eta, grinding = 0.02, 16
    shapes = [(log_n - k0, k0, k0, r0)]
    n_run, rate = log_n - k0, r0
    while n_run > 5:
        kr = min(3, n_run)
        rate += 1
        shapes.append((n_run - kr, kr, kr, rate))
        n_run -= kr

    def per_q(r):
        rho = 2.0 ** -r
        return math.log2(1.0 / (1.0 - (1.0 - math.sqrt(rho) - eta)))

    def log_a(r, mc, il):
        rho = 2.0 ** -r
        sr = math.sqrt(rho)
        gamma = 1.0 - sr - eta
        mp_ = max(math.ceil(sr / (2 * eta)), 3)
        half = mp_ + 0.5
        num = 2.0 * half ** 5 + 3.0 * half * gamma * rho
        den = 3.0 * rho ** 1.5
        n = 2.0 ** (mc + r)
        return math.log2((num / den) * n + half / sr) + max(0.0, il - 1.0)

    def ood_bits(r, mu, samples):
        rho = 2.0 ** -r
        l2 = math.log2(1.0 / (2 * eta * math.sqrt(rho)))
        if samples == 0:
            return 128.0 - l2 - math.log2(mu)
        return samples * (128.0 - math.log2(mu)) - (2 * l2 - 1)

    levels = []
    for i, (mc, il, kr, r) in enumerate(shapes):
        q = next(q for q in range(1, 10001) if q * per_q(r) + 1e-3 >= target - grinding)
        if q > 1 << (mc + r):
            raise ValueError("custom Johnson level is too thin for its query count")
        pg = 128.0 - log_a(r, mc, il)
        ood = 0 if i == 0 else 1
        if i > 0:
            while ood_bits(r, mc + il, ood) + 1e-3 < target:
                ood += 1
        levels.append(dict(rate=r, msg_cols=mc, lanes=il, k=kr, queries=q, grinding=grinding,
                           fold_grinding=max(0, math.ceil(target - pg)), ood=ood))
    rho = 2.0 ** -r0
    lst = 1.0 / (2 * eta * math.sqrt(rho))
    pairs = max(lst * (lst - 1) / 2, 1.0)
    bits0 = 128.0 - math.log2(pairs) - math.log2(max(2.0 ** log_n - 1, 1.0))
    r0grind = 0 if target - bits0 <= 0 else math.ceil(target - bits0)
    return dict(m=m, log_n=log_n, initial_k=k0, yr_log_n=n_run, target=target, levels=levels,
                round0_grinding=r0grind)


def config_frame(cfg: dict) -> bytes:
    words = [cfg["m"], cfg["initial_k"], cfg["yr_log_n"], cfg["target"]]
    for lv in cfg["levels"]:
        words += [lv["rate"], lv["msg_cols"], lv["lanes"], lv["k"], lv["queries"], lv["grinding"],
                  lv["fold_grinding"], lv["ood"]]
    return b"".join(u64(w) for w in words)


# --------------------------------------------------------- integers mod q
def miller_rabin(n: int) -> bool:
    bases = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79,
             83, 89]
    if n < 2:
        return False
    for p in bases:
        if n == p:
            return True
        if n % p == 0:
            return False
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in bases:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def sample_prime(tr: Transcript, bits: int) -> int:
    top = 1 << (bits - 1)
    tr.absorb_slice(b"bitz/shared-prime-sampling/v1")
    tr.absorb_slice(g16(top))
    tr.absorb_slice(g16(top | (top - 1)))
    for _ in range(64 * bits):
        cand = (gint(tr.sampling_bytes(16)) & (top - 1)) | top | 1
        if miller_rabin(cand):
            return cand
    raise RuntimeError("prime search exhausted")


def sample_residue(tr: Transcript, q: int) -> int:
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
    tr = Transcript()
    tr.absorb_slice(b"bitz/ligerito-flock/standalone-mod-q-statement/v1")
    tr.absorb_slice(root)
    tr.absorb_slice(config_frame(cfg))
    tr.absorb_slice(u64(t) + u64(s))
    tr.absorb_slice(g16(generator()))
    q_bits = min(113, 126 - t)
    tr.absorb_slice(u64(q_bits))
    tr.absorb_slice(u64(cfg["round0_grinding"] + 1))
    tr.absorb_slice(b"bitz/bitz-opener/standalone/v1")
    tr.absorb_slice(b"bitz/core/ood-round/v1")
    tr.absorb_slice(u64(t + s - LOG_PACKING))
    tr.absorb_slice(u32(cfg["round0_grinding"]))
    return tr


def ood_point(zeta: int, n: int) -> list[int]:
    pt, cur = [], zeta
    for _ in range(n):
        pt.append(cur)
        cur = O.mul(cur, cur)
    return pt


def instance(tr: Transcript, t: int, s: int):
    q = sample_prime(tr, min(113, 126 - t))
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
            seed = tr.challenge(16)
            nonce = O.pow_find(POW_PREFIX_NATIVE, seed, bits)
            tr.absorb_slice(u64(nonce))
        zeta = gint(tr.challenge(16))
        pt = ood_point(zeta, self.mp)
        y = O.mle_eval(self.packed, gfa(pt))
        tr.absorb_slice(g16(y))
        return tr, pt, y, nonce

    def evaluation(self) -> int:
        tr, _, _, _ = self._round0()
        q, row_w, col_w = instance(tr, self.t, self.s)
        folds = self._folds(row_w)
        return sum(c * (f % q) for c, f in zip(col_w, folds)) % q

    # ---- GKR (dense, gpgkr_prove) ----
    def _gkr(self, it: Inner, zeta: list[int], row_images: list[int]):
        t, s = self.t, self.s
        leaves = O.forest_leaves(self.rows, t, s, gfa(row_images))
        layers = [leaves]
        while layers[-1].shape[0] > (1 << s) * 2:
            layers.append(O.product_level(layers[-1]))
        internal = list(reversed(zeta))
        claim = 0
        for w in reversed(layers):
            L = len(internal)
            l, r = w[: 1 << L].copy(), w[1 << L:].copy()
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
        tr.absorb_slice(b"bitz/ligerito-flock/standalone-mod-q-claim/v1")
        tr.absorb_slice(g16(q) + g16(claimed))
        tag = tr.challenge(16) + tr.challenge(16)
        it = Inner(b"bitz/bitz-opener/standalone/v1", tag)
        g = generator()
        it.public(self.root)
        it.public(u64(t) + u64(s) + g16(q) + g16(g))
        it.public(claim_encoding(row_w, col_w, claimed))
        for f in folds:
            it.prover(g16(f))
        zeta = [it.gf() for _ in range(s)]
        row_images = O.gf_to_ints(O.pow_many(g, gfa(row_w)))
        point, leaf_claim = self._gkr(it, zeta, row_images)
        alpha_c, alpha_b = point[:s], point[s:]
        eqb = eq_ints(alpha_b)
        u1 = [O.mul(a ^ 1, e) for a, e in zip(row_images, eqb)]
        u2 = eq_ints(alpha_c)
        target = leaf_claim ^ 1
        pcs40 = u64(self.m) + u64(self.r0) + u64(self.k0) + u64(0) + u64(1)
        it.public(b"bitz/pcs/bit-inner-product/v2")
        it.public(self.root)
        it.public(pcs40)
        it.public(claim_encoding(u1, u2, target))
        it.public(b"bitz/pcs/inner-product-sumcheck/v1")
        comb = O.xi_combined_rows(self.rows, t, s, gfa(u2))
        sc_point = []
        _, wts, target = self._sumcheck(it, comb, gfa(u1), target, sc_point)
        row_scalar = O.gf_to_ints(wts)[0]
        folded = O.fold_rows_eq(self.rows, t, s, O.eq_table(gfa(sc_point)))
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
        seed = tr.challenge(16)
        if r0_nonce is None or not O.pow_valid(POW_PREFIX_NATIVE, seed, r0_nonce, bits):
            raise VerifyError("Round-0 proof of work")
        tr.absorb_slice(u64(r0_nonce))
    elif r0_nonce is not None:
        raise VerifyError("unexpected Round-0 nonce")
    zeta0 = gint(tr.challenge(16))
    opt = ood_point(zeta0, mp)
    tr.absorb_slice(g16(ood_y))
    q, row_w, col_w = instance(tr, t, s)
    if claimed >= q:
        raise VerifyError("claim not reduced")
    tr.absorb_slice(b"bitz/ligerito-flock/standalone-mod-q-claim/v1")
    tr.absorb_slice(g16(q) + g16(claimed))
    tag = tr.challenge(16) + tr.challenge(16)
    it = Inner(b"bitz/bitz-opener/standalone/v1", tag, narg)
    g = generator()
    it.public(root)
    it.public(u64(t) + u64(s) + g16(q) + g16(g))
    it.public(claim_encoding(row_w, col_w, claimed))
    # fold round (fold.rs receive_fold): range bound, reconstruction, zeta
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
    target = claim ^ 1
    pcs40 = u64(m) + u64(log_inv_rate) + u64(initial_k) + u64(0) + u64(1)
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
