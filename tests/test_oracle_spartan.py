"""The u64 multiplication proof's outer transcript against the reference's
own golden pins.

tests/transcript_state_pins.rs:163-172 of the reference records the BLAKE3
state of the prover's outer Blake3Transcript after `protocol::prove` of
2^15 u64 multiplications (x_i = i * 0x9e3779b97f4a7c15 | 1,
y_i = i * 0xc2b2ae3d27d4eb4f | 1, wrapping; :269-289), at the default split
and at split shift +1. That state covers every byte the outer transcript
absorbs: the assignment binding, the ladder policy digest, Round 0, the
session, the runtime prime search, the Spartan statement (matrix digest),
tau, every outer and inner sumcheck message and challenge, and the fork
tag. The oracle's restatement must reproduce both digests exactly.
"""
import random

import pytest

import bitz_protocol as P

import oracle as O
import spartan_u64 as S

PINS = {
    0: "430c3b2411088ae3818c681c01f2b22a99041899a30f4547e142d5b230748603",
    1: "26f5e94960932bb7251be7f833231c38dfeb4242b4f1d7c7b09df1028fbbfecc",
}


@pytest.mark.parametrize("shift", [0, 1])
def test_outer_transcript_equals_reference_pin(shift):
    xs, ys = S.pin_inputs(1 << 15)
    pr = S.SpartanU64Prover(xs, ys, split_shift=shift)
    tr, out = pr.prove(with_opening=False)
    assert out["q"].bit_length() == min(113, 126 - pr.lay["row_vars"])
    assert O.blake3(bytes(tr.buf), 32).hex() == PINS[shift]


def test_security_params_at_lambda_100():
    """derive_params' closed forms: all grinding boundaries off at these
    shapes, the prime interval [2^(w-1), 2^w - 1] with w = min(113, 126 - t)"""
    for t, tau in ((14, 15), (16, 20), (13, 15)):
        sec = S.security_params(t, tau)
        w = min(113, 126 - t)
        assert (sec["pmin"], sec["pmax"]) == (1 << (w - 1), (1 << w) - 1)
        assert sec["initial"] == sec["piop"] == sec["terminal"] == sec["forest"] == 0


def test_sumchecks_and_bitification_on_a_ragged_batch():
    """a batch that does not fill its capacity (live rows < 2^g): the
    prover's own checks (round sums, terminal claims) hold, and the bitified
    row/column weights reproduce the adjusted claim over the committed bits"""
    rng = random.Random(9)
    n = 3000
    xs = [rng.getrandbits(64) for _ in range(n)]
    ys = [rng.getrandbits(64) for _ in range(n)]
    xs[0], ys[0] = 2**64 - 1, 2**64 - 1
    pr = S.SpartanU64Prover(xs, ys)
    _, out = pr.prove(with_opening=False)
    q, bit = out["q"], out["bitified"]
    lay = pr.lay
    t, s = lay["row_vars"], lay["col_vars"]
    words = 1 << (t - 6)
    rows = pr.pcs.rows
    acc = 0
    for c in range(1 << s):
        cw = bit["col_w"][c]
        if cw == 0:
            continue
        col = 0
        for b in range(1 << t):
            if (int(rows[c * words + b // 64]) >> (b % 64)) & 1:
                col += bit["row_w"][b]
        acc = (acc + cw * (col % q)) % q
    assert acc == bit["claimed"]
    # and the claim is the MLE of the assignment at r_y minus the constant cell
    z = out["z_ry"]
    assert out["value"] == out["scale"] * z % q


def _batch(n, seed):
    rng = random.Random(seed)
    xs = [rng.getrandbits(64) for _ in range(n)]
    ys = [rng.getrandbits(64) for _ in range(n)]
    xs[1], ys[1] = 2**64 - 1, 2**64 - 1
    xs[2] = 0
    return xs, ys


@pytest.fixture(scope="module")
def honest_proof():
    xs, ys = _batch(3000, 21)
    pr = S.SpartanU64Prover(xs, ys)
    tr, out = pr.prove()
    return pr, tr, S.encode_proof(out)


def test_full_proof_verifies(honest_proof):
    """the CPU verifier replays the same outer transcript and accepts"""
    pr, tr, proof = honest_proof
    vt = S.verify_u64(pr.pcs.root, pr.n, 0, proof)
    assert bytes(vt.buf) == bytes(tr.buf)


@pytest.mark.parametrize("where", ["outer", "abc", "inner", "narg", "tail"])
def test_tampered_proof_rejected(honest_proof, where):
    pr, _, proof = honest_proof
    g = pr.lay["gate_vars"]
    outer_len = 16 * 4 * pr.tau_arity
    pos = {"outer": 40, "abc": outer_len + 20, "inner": outer_len + 48 + 16 * 3 * (g // 2) + 3,
           "narg": outer_len + 48 + 16 * 3 * (g + 3) + 8 + 100, "tail": len(proof) - 3}[where]
    bad = bytearray(proof)
    bad[pos] ^= 1
    with pytest.raises(P.VerifyError):
        S.verify_u64(pr.pcs.root, pr.n, 0, bytes(bad))


def test_wrong_statement_rejected(honest_proof):
    pr, _, proof = honest_proof
    with pytest.raises(P.VerifyError):
        S.verify_u64(bytes(32), pr.n, 0, proof)
    with pytest.raises(P.VerifyError):
        S.verify_u64(pr.pcs.root, pr.n - 1, 0, proof)


def test_wrong_product_rejected():
    """a committed z_hi off by one in one gate (the reference's own
    rejection test, u64_bitz.rs:272-310): the rows no longer satisfy
    x y = lo + 2^64 hi and the proof is rejected"""
    xs, ys = _batch(3000, 22)
    lo = [(a * b) & (2**64 - 1) for a, b in zip(xs, ys)]
    hi = [(a * b) >> 64 for a, b in zip(xs, ys)]
    hi[3] ^= 1
    pr = S.SpartanU64Prover(xs, ys, products=(lo, hi))
    _, out = pr.prove()
    with pytest.raises(P.VerifyError):
        S.verify_u64(pr.pcs.root, pr.n, 0, S.encode_proof(out))


def test_eq_prefix_sum_brute_force():
    rng = random.Random(4)
    q = (1 << 61) - 1
    a = [rng.randrange(q) for _ in range(6)]
    b = [rng.randrange(q) for _ in range(6)]
    ea, eb = S.eq_le(a, q), S.eq_le(b, q)
    for live in (0, 1, 5, 17, 32, 63, 64):
        assert S.eq_prefix_sum(a, b, live, q) == sum(x * y for x, y in zip(ea[:live], eb[:live])) % q
