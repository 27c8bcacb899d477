"""The u64 multiplication proof on the device (bitz_prove_mul_u64 /
bitz_verify_mul_u64): Spartan's outer and inner sumchecks over the runtime
prime field, bitification and the BitZ opening of the bitified claim.

* At the reference's pinned instance (2^15 multiplications,
  tests/transcript_state_pins.rs:163-172, split shifts 0 and +1) the device
  prover's outer transcript state equals the reference's pinned digest, and
  its root and proof bytes equal the CPU restatement's (oracle/spartan_u64.py,
  itself pinned to the same digests).
* Ragged batches (n not a power of two) give the oracle's bytes too.
* The multiplication workload of 2^20 u64 pairs (t = 16, s = 12): the
  library's verifier and the CPU verifier accept the device proof.
* A wrong committed product and tampered bytes are rejected by both verifiers.
"""
import random

import numpy as np
import pytest

import bitz_protocol as P
import spartan_u64 as S
from gpu_util import ctx, log

pytestmark = pytest.mark.gpu

PINS = {
    0: "430c3b2411088ae3818c681c01f2b22a99041899a30f4547e142d5b230748603",
    1: "26f5e94960932bb7251be7f833231c38dfeb4242b4f1d7c7b09df1028fbbfecc",
}


def _dev(vals):
    import torch
    return torch.from_numpy(np.array(vals, dtype=np.uint64).view(np.int64)).to("cuda")


def _prove(xs, ys, shift=0, lo=None, hi=None):
    from bitz_amd.mul import prove_u64
    return prove_u64(ctx(), _dev(xs), _dev(ys), shift, None if lo is None else _dev(lo),
                     None if hi is None else _dev(hi))


@pytest.mark.parametrize("shift", [0, 1])
def test_reference_pin_and_oracle_bytes(shift):
    from bitz_amd.mul import verify_u64
    xs, ys = S.pin_inputs(1 << 15)
    log(f"u64 multiplication proof, 2^15 pin instance, shift {shift}")
    root, state, proof = _prove(xs, ys, shift)
    assert state.hex() == PINS[shift]
    pr = S.SpartanU64Prover(xs, ys, split_shift=shift)
    _, out = pr.prove()
    assert root == pr.pcs.root
    assert proof == S.encode_proof(out)
    assert verify_u64(ctx(), root, 1 << 15, proof, shift) == state


@pytest.mark.parametrize("n", [3000, 5000, 40000])
def test_ragged_batches_equal_oracle(n):
    rng = random.Random(n)
    xs = [rng.getrandbits(64) for _ in range(n)]
    ys = [rng.getrandbits(64) for _ in range(n)]
    xs[0] = ys[0] = 2**64 - 1
    log(f"u64 multiplication proof, n={n}")
    root, state, proof = _prove(xs, ys)
    pr = S.SpartanU64Prover(xs, ys)
    tr, out = pr.prove()
    assert root == pr.pcs.root
    assert state == P.O.blake3(bytes(tr.buf), 32)
    assert proof == S.encode_proof(out)


def test_config2_workload_verifies():
    """2^20 uniform u64 pairs from a seed (the multiplication workload):
    both verifiers accept the device proof"""
    from bitz_amd.mul import verify_u64
    n = 1 << 20
    rng = np.random.default_rng(20)
    xs = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    ys = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    log("u64 multiplication proof, 2^20 pairs")
    import torch
    x = torch.from_numpy(xs.view(np.int64)).to("cuda")
    y = torch.from_numpy(ys.view(np.int64)).to("cuda")
    from bitz_amd.mul import prove_u64
    root, state, proof = prove_u64(ctx(), x, y)
    assert verify_u64(ctx(), root, n, proof) == state
    log(f"CPU verifier on the 2^20 proof ({len(proof)} bytes)")
    vt = S.verify_u64(root, n, 0, proof)
    assert P.O.blake3(bytes(vt.buf), 32) == state


def test_wrong_product_and_tampering_rejected():
    from bitz_amd import _lib
    from bitz_amd.mul import verify_u64
    n = 5000
    rng = random.Random(77)
    xs = [rng.getrandbits(64) for _ in range(n)]
    ys = [rng.getrandbits(64) for _ in range(n)]
    lo = [(a * b) & (2**64 - 1) for a, b in zip(xs, ys)]
    hi = [(a * b) >> 64 for a, b in zip(xs, ys)]
    hi[3] ^= 1
    log("u64 multiplication proof of a wrong product")
    root, _, proof = _prove(xs, ys, 0, lo, hi)
    with pytest.raises(_lib.BitzError):
        verify_u64(ctx(), root, n, proof)
    with pytest.raises(P.VerifyError):
        S.verify_u64(root, n, 0, proof)
    log("u64 multiplication proof, tampered bytes")
    root, _, proof = _prove(xs, ys)
    for pos in (5, 16 * 4 * 13 + 7, 16 * (4 * 13 + 3) + 30, len(proof) // 2, len(proof) - 2):
        bad = bytearray(proof)
        bad[pos] ^= 0x10
        with pytest.raises(_lib.BitzError):
            verify_u64(ctx(), root, n, bytes(bad))
    with pytest.raises(_lib.BitzError):
        verify_u64(ctx(), root, n - 1, proof)
