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
