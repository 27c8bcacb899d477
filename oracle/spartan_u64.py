"""CPU restatement of the u64 multiplication proof (x * y = lo + 2^64 hi,
MulLayout<u64>): statement binding, Round 0, the runtime prime, the Spartan
PIOP over F_q (outer cubic and inner quadratic sumchecks), bitification of
the terminal claim, and its BitZ opening.

TEST INFRASTRUCTURE ONLY (see oracle.py): the checker of the device prover
and the CPU baseline of the multiplication workload.

The outer (Blake3Transcript) stream follows the reference byte for byte:
  protocol::prove -> prove_with_opener -> bitz_opener::prove_direct
  (src/piop/spartan/protocol/mod.rs:1138-1170, bitz_opener.rs:439-494):
  bind_prover_statement (mod.rs:866-887): the assignment binding as a
    Spartan frame (u64_bitz.rs:135-171, binding.rs:185-208), the opener
    policy digest (configuration.rs:224-227), Round 0
    (ligerito_flock.rs:1520-1549);
  the session label and the ladder digest (bitz_opener.rs:461-462);
  prove_piop (mod.rs:1410-1506): the runtime prime (sample_mod_q,
    mod.rs:1852-1870), the Spartan PIOP (piop.rs:353-420): its statement
    (piop.rs:963-983, the matrix digest matrix.rs:1478-1513), tau, the
    outer sumcheck (sumcheck/outer/ordinary.rs, engine.rs, proof.rs:337-377),
    rho, the inner sumcheck (sumcheck/inner/mod.rs:80-150);
  the fork tag (bitz_opener.rs:377-384).
Every grinding boundary of the lambda = 100 profile is at difficulty 0 for
these shapes and absorbs nothing (mod.rs:1822-1832, grinding.rs:356-386).
The inner (opening) transcript is this build's Blake3-framed one, as in the
standalone opening.
"""
from __future__ import annotations

import math

import numpy as np

import bitz_protocol as P
import mul_witness as MW
import oracle as O
from bitz_protocol import g16, gint, u64

BINDING_DOMAIN = b"bitz/spartan-u64-bitz/assignment/v1-runtime"
BLOCK_ORDER = b"e0|x|y|zlo|zhi|zero|zero|zero"
STATEMENT_TAG = b"u64-mul-statement"
PRIME_DOMAIN = b"bitz/spartan-u64-mul/runtime-prime/v1"
BRIDGE_DOMAIN = b"bitz/spartan-u64-bitz/bitified-claim/v1"
FRAME_DOMAIN = b"bitz/spartan/transcript-frame/v2"
PIOP_DOMAIN = b"bitz/spartan/piop/v2"
ORACLE_DOMAIN = b"bitz/spartan/assignment-oracle/v1"
MATRIX_DOMAIN = b"bitz/spartan/constraint-matrices/v2"
POLICY_LABEL = b"bitz/ligerito-policy/early-ood/v1"
SESSION = b"bitz/bitz-opener/v1"
LIMB_BASE = 1 << 64
SLOTS = 256  # bit slots per multiplication: x, y, lo, hi, 64 each (u64_mul.rs:33-49)
LAMBDA = 100
LIGERITO_RATE, LIGERITO_K = 1, 4  # LigeritoSelection::JOHNSON (configuration.rs:24-27)


def u128b(v: int) -> bytes:
    return int(v).to_bytes(16, "little")


# ------------------------------------------------------------ transcript frames
def spartan_message(tr: P.Transcript, tag: bytes, payload: bytes) -> None:
    """absorb_spartan_message (src/piop/spartan/mod.rs:349-358)."""
    tr.absorb_slice(FRAME_DOMAIN + u64(len(tag)) + tag + u64(len(payload)) + payload)


def absorb_elements(tr: P.Transcript, vals) -> None:
    """absorb_field_elements (mod.rs:330-346): count, then each canonical
    16-byte encoding behind its length."""
    payload = u64(len(vals)) + b"".join(u64(16) + u128b(v) for v in vals)
    spartan_message(tr, b"field-elements", payload)


def squeeze(tr: P.Transcript, q: int) -> int:
    """squeeze_field (mod.rs:371-384): sample_below(q) over sampling reads
    (256 draws, vendor/field/src/prime/sampling.rs:373-376), then absorbed."""
    x = P.sample_below(tr, q, 256)
    absorb_elements(tr, [x])
    return x


# ------------------------------------------------------------ security profile
def log2_prime_count_lower_bound(b: int) -> float:
    """profile.rs:484-493."""
    ln2 = math.log(2.0)
    density = 1.0 / (b * ln2) - 0.63 / ((b - 1.0) * ln2)
    return b + math.log2(density)


def ceil_pos(v: float) -> int:
    return 0 if v <= 0.0 else int(math.ceil(v))


def security_params(row_vars: int, tau_arity: int, lam: int = LAMBDA) -> dict:
    """derive_params (profile.rs:506-646) for the single derived prime at the
    u64 relation's facts (u64_bitz.rs:56-66: defect < 2^130, lift arity and
    opening t = the row variables, a direct opening, degree 3)."""
    c_w = 126 - row_vars
    width = min(113, 126, 128 - row_vars, c_w)
    pmin = 1 << (width - 1)
    no_wrap_max = 1 + (2**128 - 1) // ((1 << row_vars) + 1)
    pmax = min((1 << width) - 1, no_wrap_max)
    divisors = 130 // (width - 1)
    projection_raw = log2_prime_count_lower_bound(width) - math.log2(divisors) if divisors else None
    tau_raw = (width - 1) - math.log2(max(tau_arity, 1))
    draw_raw = min(projection_raw, tau_raw) if projection_raw is not None else tau_raw
    initial = ceil_pos(lam - draw_raw)
    round_raw = (width - 1) - math.log2(3)
    piop = ceil_pos(lam - round_raw)
    forest = 0 if lam <= 126 else lam - 126
    return {"lambda": lam, "pmin": pmin, "pmax": pmax, "ligerito_target": lam, "initial": initial,
            "piop": piop, "terminal": piop, "forest": forest}


# ------------------------------------------------------------ statement
def assignment_binding(root: bytes, cfg: dict, sec: dict, lay: dict, mults: int) -> bytes:
    """bind_assignment (binding.rs:185-208) with the u64 relation's middle
    section (u64_bitz.rs:135-171)."""
    cap, g = lay["capacity"], lay["gate_vars"]
    h = O.Blake3()
    h.update(BINDING_DOMAIN + BLOCK_ORDER + root)
    # commitment params (binding.rs:114-121): m, rate, batch, Fast, BLAKE3
    h.update(u64(cfg["m"]) + u64(cfg["levels"][0]["rate"]) + u64(cfg["initial_k"]) + b"\x00\x01")
    h.update(u128b(sec["pmin"]) + u128b(sec["pmax"]))
    for v in (sec["lambda"], sec["ligerito_target"], sec["initial"], sec["piop"], sec["terminal"], sec["forest"]):
        h.update(u64(v))
    h.update(u128b(LIMB_BASE))
    for v in (mults, cap, 5 * cap, 8 * cap, g, 5, 8, 64, 0, 64, 128, 192, 256, lay["row_vars"],
              lay["col_vars"], 1):
        h.update(u64(v))
    return h.digest(32)


def matrix_digest(q: int, mults: int, cap: int) -> bytes:
    """constraint_matrix_digest (matrix.rs:1478-1513) of the u64 matrices
    (u64_mul.rs:155-176, selector_matrix mul.rs:476-503): rows = the
    multiplications, columns = 5 * capacity, per CSR row its length, then
    (column, coefficient encoding) pairs."""
    one, base = u64(16) + u128b(1), u64(16) + u128b(LIMB_BASE)
    h = O.Blake3()
    h.update(MATRIX_DOMAIN + u64(16) + u128b(q) + u64(mults) + u64(5 * cap))
    for label, blocks in ((b"A", [(1, one)]), (b"B", [(2, one)]), (b"C", [(3, one), (4, base)])):
        h.update(label)
        parts = []
        for i in range(mults):
            parts.append(u64(len(blocks)))
            for blk, enc in blocks:
                parts.append(u64(blk * cap + i) + enc)
        h.update(b"".join(parts))
    return h.digest(32)


# ------------------------------------------------------------ F_q helpers
def eq_le(point, q: int) -> list[int]:
    """eq_le_table_fq_fast_with (bitify.rs:234-266): coordinate i is bit i."""
    t = [1 % q]
    for c in point:
        one = [v * c % q for v in t]
        t = [(v - o) % q for v, o in zip(t, one)] + one
    return t


def cubic_from_evals(e, q: int):
    """coefficients of the cubic through (0, e0) .. (3, e3)."""
    inv2, inv6 = pow(2, -1, q), pow(6, -1, q)
    e0, e1, e2, e3 = e
    c3 = (e3 - 3 * e2 + 3 * e1 - e0) * inv6 % q
    c2 = (-e3 + 4 * e2 - 5 * e1 + 2 * e0) * inv2 % q
    c1 = (2 * e3 - 9 * e2 + 18 * e1 - 11 * e0) * inv6 % q
    return [e0 % q, c1, c2, c3]


def horner(c, x: int, q: int) -> int:
    acc = 0
    for v in reversed(c):
        acc = (acc * x + v) % q
    return acc


def fold(v, r: int, q: int):
    return [(v[2 * i] + r * (v[2 * i + 1] - v[2 * i])) % q for i in range(len(v) // 2)]


def outer_sumcheck(tr: P.Transcript, q: int, tau, A, B, C, strict: bool = True):
    """Spartan's outer sumcheck of sum_x eq(tau, x)(A(x)B(x) - C(x)) = 0, the
    lowest variable first: per round the cubic's four coefficients absorbed,
    a challenge squeezed; then [A, B, C](r_x) absorbed (outer/engine.rs,
    ordinary.rs:150-214, proof.rs:337-377)."""
    eq = eq_le(tau, q)
    claim, point, polys = 0, [], []
    for _ in range(len(tau)):
        e = [0, 0, 0, 0]
        for i in range(0, len(A), 2):
            a0, a1, b0, b1, c0, c1, w0, w1 = A[i], A[i + 1], B[i], B[i + 1], C[i], C[i + 1], eq[i], eq[i + 1]
            da, db, dc, dw = a1 - a0, b1 - b0, c1 - c0, w1 - w0
            for k in range(4):
                e[k] += (w0 + k * dw) * ((a0 + k * da) * (b0 + k * db) - (c0 + k * dc))
        c = cubic_from_evals([x % q for x in e], q)
        assert not strict or (c[0] + sum(c)) % q == claim, "outer sumcheck: round sum"
        absorb_elements(tr, c)
        r = squeeze(tr, q)
        claim = horner(c, r, q)
        point.append(r)
        polys.append(c)
        A, B, C, eq = fold(A, r, q), fold(B, r, q), fold(C, r, q), fold(eq, r, q)
    assert not strict or claim == eq[0] * (A[0] * B[0] - C[0]) % q, "outer sumcheck: terminal"
    absorb_elements(tr, [A[0], B[0], C[0]])
    return point, polys, (A[0], B[0], C[0])


def inner_sumcheck(tr: P.Transcript, q: int, claim: int, W, Z):
    """Spartan's inner sumcheck of sum_y W(y) Z(y), the lowest variable
    first: per round [c0, c1, c2] absorbed (c1 from the claim), a challenge
    squeezed (sumcheck/inner/mod.rs:80-150)."""
    point, polys = [], []
    while len(W) > 1:
        c0 = c2 = 0
        for i in range(0, len(W), 2):
            w0, z0 = W[i], Z[i]
            c0 += w0 * z0
            c2 += (W[i + 1] - w0) * (Z[i + 1] - z0)
        c0, c2 = c0 % q, c2 % q
        c1 = (claim - 2 * c0 - c2) % q
        absorb_elements(tr, [c0, c1, c2])
        r = squeeze(tr, q)
        claim = (c0 + r * (c1 + r * c2)) % q
        point.append(r)
        polys.append([c0, c1, c2])
        W, Z = fold(W, r, q), fold(Z, r, q)
    assert claim == W[0] * Z[0] % q, "inner sumcheck: terminal"
    return point, polys, W[0], Z[0], claim


# ------------------------------------------------------------ bitification
def bitify(q: int, point, scale: int, value: int, lay: dict):
    """bitify (bitify.rs:119-219) with scale on the rows, then the dense row
    weights (:268-318) and column weights (:338-352): row (slot << h) | gh,
    column = the low gate bits (mul.rs:140-150)."""
    g, s, t = lay["gate_vars"], lay["col_vars"], lay["row_vars"]
    gate, sel = point[:g], point[g:]

    def block_factor(code):
        acc = 1
        for bit, c in enumerate(sel):
            acc = acc * (c if (code >> bit) & 1 else (1 - c)) % q
        return acc

    const = block_factor(0)
    for c in gate:
        const = const * (1 - c) % q
    claimed = (value - scale * const) % q
    factors = [block_factor(code) for code in (1, 2, 3, 4)]
    if all(f == 0 for f in factors):
        raise ValueError("constant-only claim (a zero-probability point) is not handled")
    if scale in (0, 1):
        rows_f, col_scale = factors, scale
    else:
        rows_f, col_scale = [scale * f % q for f in factors], 1
    h = g - s
    eq_hi = eq_le(gate[s:], q)
    words = [0] * SLOTS
    for b, f in enumerate(rows_f):
        sc = f
        for k in range(64):
            words[64 * b + k] = sc
            sc = sc * 2 % q
    row_w = [0] * (1 << t)
    for slot in range(SLOTS):
        ws = words[slot]
        if ws:
            base = slot << h
            for gh, e in enumerate(eq_hi):
                row_w[base + gh] = ws * e % q
    col_w = [col_scale * e % q for e in eq_le(gate[:s], q)] if col_scale else [0] * (1 << s)
    return {"gate_point": gate, "rows": rows_f, "col_scale": col_scale, "claimed": claimed,
            "row_w": row_w, "col_w": col_w}


def bridge_digest(binding: bytes, q: int, mdigest: bytes, lay: dict, mults: int, point, scale: int, value: int,
                  bit: dict) -> bytes:
    """bridge_digest (bitify.rs:365-410) with the u64 relation's constants
    (u64_bitz.rs:173-202)."""
    cap, g, s, t = lay["capacity"], lay["gate_vars"], lay["col_vars"], lay["row_vars"]
    h = O.Blake3()
    h.update(BRIDGE_DOMAIN + binding + u64(16) + u128b(q) + mdigest + u128b(q))
    h.update(u128b(LIMB_BASE))
    for v in (mults, cap, 5 * cap, 8 * cap, g, 5, 8, t, s, 1, 0, 64, 128, 192, 64, 256):
        h.update(u64(v))
    h.update(bytes([1, 0, 1, 2, 3, 4, 5, 6, 7]))
    h.update(u64(len(point)) + b"".join(u128b(c) for c in point))
    h.update(u128b(scale) + u128b(value))
    gp = bit["gate_point"]
    h.update(u64(s) + b"".join(u128b(c) for c in gp[:s]))
    h.update(u64(g - s) + b"".join(u128b(c) for c in gp[s:]))
    h.update(b"\x00" + b"".join(u128b(f) for f in bit["rows"]))
    h.update(u128b(bit["col_scale"]) + u128b(bit["claimed"]))
    return h.digest(32)


# ------------------------------------------------------------ prover
def pin_inputs(n: int):
    """the reference's transcript-pin witness (tests/transcript_state_pins.rs:
    269-289): x_i = i * 0x9e3779b97f4a7c15 | 1, y_i = i * 0xc2b2ae3d27d4eb4f | 1
    (wrapping)."""
    m = 2**64 - 1
    xs = [((i * 0x9E3779B97F4A7C15) & m) | 1 for i in range(n)]
    ys = [((i * 0xC2B2AE3D27D4EB4F) & m) | 1 for i in range(n)]
    return xs, ys


class SpartanU64Prover:
    """The u64 multiplication relation over 2^k multiplications (x, y),
    products computed here, committed as the 256-slot bit grid."""

    def __init__(self, xs, ys, split_shift: int = 0, products=None):
        """products: (lo, hi) limbs to commit instead of the true ones (a
        dishonest witness, for rejection tests)"""
        self.n = len(xs)
        self.xs = [int(v) for v in xs]
        self.ys = [int(v) for v in ys]
        self.lo = [(a * b) & (2**64 - 1) for a, b in zip(self.xs, self.ys)]
        self.hi = [(a * b) >> 64 for a, b in zip(self.xs, self.ys)]
        self.honest = products is None
        if products is not None:
            self.lo, self.hi = [int(v) for v in products[0]], [int(v) for v in products[1]]
        self.lay = MW.layout(64, self.n, split_shift)
        t, s = self.lay["row_vars"], self.lay["col_vars"]
        rows = MW.bit_rows(self.lay, self.xs, self.ys, self.lo, self.hi)
        self.pcs = P.OracleProver(rows, t, s, LIGERITO_RATE, LIGERITO_K)
        self.cfg = self.pcs.cfg
        self.tau_arity = max((self.n - 1).bit_length(), 0) if self.n > 1 else 0
        self.sec = security_params(t, max(self.tau_arity, 1))

    def prove(self, with_opening: bool = True):
        """returns (outer transcript, proof dict); the opening's bytes when
        with_opening"""
        lay, cfg, sec = self.lay, self.cfg, self.sec
        assert sec["initial"] == sec["piop"] == sec["terminal"] == 0, "grinding boundaries not restated"
        cap, g = lay["capacity"], lay["gate_vars"]
        mults = self.n
        tr = P.Transcript()
        binding = assignment_binding(self.pcs.root, cfg, sec, lay, mults)
        spartan_message(tr, STATEMENT_TAG, binding)
        tr.absorb_slice(POLICY_LABEL)
        tr.absorb_slice(P.ladder_digest(cfg))
        # Round 0 (prove_ood_round_packed, ligerito_flock.rs:1520-1549)
        mp = self.pcs.mp
        bits = cfg["round0_grinding"]
        P.ood_round_header(tr, mp, bits)
        r0_nonce = None
        if bits:
            seed = P.grinding_seed(tr, P.OOD_GRIND_DOMAIN, 0, bits)
            r0_nonce = O.pow_find(seed[:16], seed[16:], bits)
            P.absorb_grinding_nonce(tr, r0_nonce)
        zeta = gint(tr.challenge(16))
        opt = P.ood_point(zeta, mp)
        ood_y = O.mle_eval(self.pcs.packed, P.gfa(opt))
        tr.absorb_slice(b"\x50" + g16(ood_y))
        tr.absorb_slice(SESSION)
        tr.absorb_slice(P.ladder_digest(cfg))
        # runtime prime (sample_mod_q, mod.rs:1852-1870)
        spartan_message(tr, b"prime-domain", PRIME_DOMAIN)
        spartan_message(tr, b"prime-min", u128b(sec["pmin"]))
        spartan_message(tr, b"prime-max", u128b(sec["pmax"]))
        q = P.sample_prime_context(tr, sec["pmin"], sec["pmax"], 128)
        spartan_message(tr, b"prime-q", u128b(q))
        # the Spartan statement (piop.rs:963-983)
        mdigest = matrix_digest(q, mults, cap)
        spartan_message(tr, b"protocol", PIOP_DOMAIN)
        spartan_message(tr, b"field-modulus", u128b(q))
        spartan_message(tr, b"matrix-statement", mdigest)
        spartan_message(tr, ORACLE_DOMAIN, binding)
        rows_pow2 = 1 << self.tau_arity
        tau = [squeeze(tr, q) for _ in range(self.tau_arity)]
        A = [v % q for v in self.xs] + [0] * (rows_pow2 - mults)
        B = [v % q for v in self.ys] + [0] * (rows_pow2 - mults)
        C = [(lo | (hi << 64)) % q for lo, hi in zip(self.lo, self.hi)] + [0] * (rows_pow2 - mults)
        rx, outer_polys, (va, vb, vc) = outer_sumcheck(tr, q, tau, A, B, C, strict=self.honest)
        rho = squeeze(tr, q)
        claim = (va + rho * vb + rho * rho % q * vc) % q
        # inner: W(b, gate) = coef_b eq(r_x, gate) over live rows; Z the assignment
        eqx = eq_le(rx, q)
        coef = [0, 1, rho, rho * rho % q, rho * rho % q * LIMB_BASE % q, 0, 0, 0]
        W = [0] * (8 * cap)
        Z = [0] * (8 * cap)
        Z[0] = 1
        for b in range(1, 5):
            cb = coef[b]
            vals = (self.xs, self.ys, self.lo, self.hi)[b - 1]
            for i in range(mults):
                W[b * cap + i] = cb * eqx[i] % q
                Z[b * cap + i] = vals[i] % q
        ry, inner_polys, w_ry, z_ry, final = inner_sumcheck(tr, q, claim, W, Z)
        bit = bitify(q, ry, w_ry, final, lay)
        bridge = bridge_digest(binding, q, mdigest, lay, mults, ry, w_ry, final, bit)
        tag = tr.challenge(16) + tr.challenge(16)
        out = {"q": q, "tau": tau, "rx": rx, "outer": outer_polys, "abc": (va, vb, vc), "rho": rho, "ry": ry,
               "inner": inner_polys, "scale": w_ry, "value": final, "z_ry": z_ry, "bitified": bit,
               "bridge": bridge, "binding": binding, "matrix_digest": mdigest, "tag": tag, "ood_y": ood_y,
               "r0_nonce": r0_nonce}
        if with_opening:
            it = P.Inner(SESSION, tag)
            it.public(bridge)
            folds = O.gf_to_ints(O.fold_columns(self.pcs.rows, lay["row_vars"], lay["col_vars"],
                                                 P.gfa(bit["row_w"])))
            out["opening"] = self.pcs.open_claim(it, q, bit["row_w"], bit["col_w"], bit["claimed"], folds, opt,
                                                 ood_y, r0_nonce)
        return tr, out


# ------------------------------------------------------------ proof bytes
def encode_proof(out: dict) -> bytes:
    """the Spartan payload as 16-byte field elements (the reference's size
    accounting, mod.rs:1071-1077: four per outer round, Az Bz Cz, three per
    inner round; no nonces at difficulty 0), then the BitZ opening bytes"""
    b = b"".join(u128b(c) for p in out["outer"] for c in p)
    b += b"".join(u128b(v) for v in out["abc"])
    b += b"".join(u128b(c) for p in out["inner"] for c in p)
    return b + out["opening"]


def eq_prefix_sum(a, b, live: int, q: int) -> int:
    """sum_{g < live} eq(a, g) eq(b, g) over n-bit g (coordinate i = bit i),
    by the bits of live from the top"""
    n = len(a)
    h0 = [(1 - x) * (1 - y) % q for x, y in zip(a, b)]
    h1 = [x * y % q for x, y in zip(a, b)]
    full = [1]
    for i in range(n):
        full.append(full[-1] * (h0[i] + h1[i]) % q)
    if live >= 1 << n:
        return full[n]
    total, prefix = 0, 1
    for i in reversed(range(n)):
        if (live >> i) & 1:
            total = (total + prefix * h0[i] % q * full[i]) % q
            prefix = prefix * h1[i] % q
        else:
            prefix = prefix * h0[i] % q
    return total


def verify_u64(root: bytes, n: int, split_shift: int, proof: bytes) -> P.Transcript:
    """Replays the prover's outer transcript from the statement (root, the
    multiplication count, the split) and the proof; raises P.VerifyError
    unless every sumcheck round, the outer terminal check, the inner
    terminal against the public weights, and the BitZ opening of the
    bitified claim hold. Returns the outer transcript."""
    lay = MW.layout(64, n, split_shift)
    t, s, g, cap = lay["row_vars"], lay["col_vars"], lay["gate_vars"], lay["capacity"]
    cfg = P.custom_johnson_config(t + s, LIGERITO_RATE, LIGERITO_K)
    tau_arity = (n - 1).bit_length() if n > 1 else 0
    sec = security_params(t, max(tau_arity, 1))
    pos = 0

    def take(k):
        nonlocal pos
        if pos + k > len(proof):
            raise P.VerifyError("truncated proof")
        v = proof[pos:pos + k]
        pos += k
        return v

    outer_raw = take(16 * (4 * tau_arity + 3))
    inner_raw = take(16 * 3 * (g + 3))
    narg, hints, ood_y, r0_nonce = P.proof_tail_parse(proof[pos:])
    if ood_y is None:
        raise P.VerifyError("missing Round-0 claim")
    tr = P.Transcript()
    binding = assignment_binding(root, cfg, sec, lay, n)
    spartan_message(tr, STATEMENT_TAG, binding)
    tr.absorb_slice(POLICY_LABEL)
    tr.absorb_slice(P.ladder_digest(cfg))
    mp = t + s - P.LOG_PACKING
    bits = cfg["round0_grinding"]
    P.ood_round_header(tr, mp, bits)
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
    opt = P.ood_point(gint(tr.challenge(16)), mp)
    tr.absorb_slice(b"\x50" + g16(ood_y))
    tr.absorb_slice(SESSION)
    tr.absorb_slice(P.ladder_digest(cfg))
    spartan_message(tr, b"prime-domain", PRIME_DOMAIN)
    spartan_message(tr, b"prime-min", u128b(sec["pmin"]))
    spartan_message(tr, b"prime-max", u128b(sec["pmax"]))
    q = P.sample_prime_context(tr, sec["pmin"], sec["pmax"], 128)
    spartan_message(tr, b"prime-q", u128b(q))
    mdigest = matrix_digest(q, n, cap)
    spartan_message(tr, b"protocol", PIOP_DOMAIN)
    spartan_message(tr, b"field-modulus", u128b(q))
    spartan_message(tr, b"matrix-statement", mdigest)
    spartan_message(tr, ORACLE_DOMAIN, binding)
    tau = [squeeze(tr, q) for _ in range(tau_arity)]
    pos = 0
    proof_outer = outer_raw

    def rd(buf, k):
        nonlocal pos
        vals = []
        for _ in range(k):
            v = gint(buf[pos:pos + 16])
            pos += 16
            if v >= q:
                raise P.VerifyError("non-canonical field element")
            vals.append(v)
        return vals

    claim, rx = 0, []
    for _ in range(tau_arity):
        c = rd(proof_outer, 4)
        if (c[0] + sum(c)) % q != claim:
            raise P.VerifyError("outer sumcheck round")
        absorb_elements(tr, c)
        r = squeeze(tr, q)
        claim = horner(c, r, q)
        rx.append(r)
    va, vb, vc = rd(proof_outer, 3)
    eq_tr = 1
    for a, b in zip(tau, rx):
        eq_tr = eq_tr * ((1 - a) * (1 - b) + a * b) % q
    if claim != eq_tr * (va * vb - vc) % q:
        raise P.VerifyError("outer sumcheck terminal")
    absorb_elements(tr, [va, vb, vc])
    rho = squeeze(tr, q)
    claim = (va + rho * vb + rho * rho % q * vc) % q
    pos = 0
    ry = []
    for _ in range(g + 3):
        c0, c1, c2 = rd(inner_raw, 3)
        if (2 * c0 + c1 + c2) % q != claim:
            raise P.VerifyError("inner sumcheck round")
        absorb_elements(tr, [c0, c1, c2])
        r = squeeze(tr, q)
        claim = (c0 + r * (c1 + r * c2)) % q
        ry.append(r)
    # the public weight at r_y: coef_b eq(sel, b) over the blocks, times
    # sum_{gate < n} eq(r_x, gate) eq(r_gate, gate)
    gate, sel = ry[:g], ry[g:]
    coef = [0, 1, rho, rho * rho % q, rho * rho % q * LIMB_BASE % q, 0, 0, 0]
    wsel = 0
    for code in range(8):
        f = coef[code]
        for bit, c in enumerate(sel):
            f = f * (c if (code >> bit) & 1 else (1 - c)) % q
        wsel = (wsel + f) % q
    rx_full = rx + [0] * (g - len(rx))
    scale = wsel * eq_prefix_sum(rx_full, gate, n, q) % q
    bit = bitify(q, ry, scale, claim, lay)
    bridge = bridge_digest(binding, q, mdigest, lay, n, ry, scale, claim, bit)
    tag = tr.challenge(16) + tr.challenge(16)
    it = P.Inner(SESSION, tag, narg)
    it.public(bridge)
    P.verify_claim(it, root, t, s, cfg, q, bit["row_w"], bit["col_w"], bit["claimed"], opt, ood_y, narg, hints)
    return tr
