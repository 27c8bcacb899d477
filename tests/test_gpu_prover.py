"""End-to-end parity of the GPU standalone opening: the commitment root, the
honest claim and the proof bytes equal the CPU restatement's at m = 20, and
the CPU verifier accepts the GPU proof at m = 20 and 24 and at the
2^20-coefficient workload (n = 26: t = 15, s = 11, rate 1/4, initial_k 4), rejecting tampered
bytes."""
import numpy as np
import pytest

import bitz_protocol as P
from gpu_util import ctx, log

pytestmark = pytest.mark.gpu


def _rows(t, s, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)


@pytest.mark.parametrize("device_fs", [1, 0])
@pytest.mark.parametrize("t,s,rate,k0,seed",
                         [(11, 9, 2, 4, 5), (12, 8, 1, 4, 6), (10, 10, 3, 3, 7), (14, 6, 2, 4, 8),
                          (14, 10, 3, 4, 9)])
def test_gpu_proof_bytes_equal_oracle(t, s, rate, k0, seed, device_fs):
    """with the GKR transcript on the device (default) and on the host; m = 24
    at rate 1/8 runs Round 0 with one bit of grinding (the others none); the
    column counts cover s > 8 (the right half carries the weights of column
    bits 8..s-1, the round sums those of bits 0..7), s = 8 and s < 8 (the
    round sums carry every column weight)"""
    from bitz_amd import _lib
    from bitz_amd.device import Commitment
    rows = _rows(t, s, seed)
    log(f"prover parity t={t} s={s} rate=1/{1 << rate} k0={k0} device_fs={device_fs}")
    c = ctx()
    _lib.call("bitz_ctx_set_device_transcript", c.ptr, device_fs)
    try:
        com = Commitment(c, rows, t, s, k0, rate)
        pr = P.OracleProver(rows, t, s, rate, k0)
        assert com.root == pr.root
        claimed = com.evaluation()
        assert claimed == pr.evaluation()
        proof = com.prove(claimed)
        want = pr.prove(claimed)
        assert proof == want
        P.verify_standalone(com.root, t, s, rate, k0, claimed, proof)
    finally:
        _lib.call("bitz_ctx_set_device_transcript", c.ptr, 0)


def test_gpu_proof_verifies_at_2p20_coefficients():
    from bitz_amd.device import Commitment
    t, s = 15, 11
    coeffs = _rows(t, s, 2024)  # 2^20 u64 coefficients = the committed bit rows
    log("prover n=26 (2^20 u64 coefficients)")
    com = Commitment(ctx(), coeffs, t, s, 4, 2)
    claimed = com.evaluation()
    proof = com.prove(claimed)
    P.verify_standalone(com.root, t, s, 2, 4, claimed, proof)
    bad = bytearray(proof)
    bad[len(proof) // 3] ^= 4
    with pytest.raises(P.VerifyError):
        P.verify_standalone(com.root, t, s, 2, 4, claimed, bytes(bad))
    with pytest.raises(P.VerifyError):
        P.verify_standalone(com.root, t, s, 2, 4, claimed ^ 1, proof)
