"""The CPU restatement of the standalone opening proves and its verifier
accepts, at the smallest ladder shape the reference accepts (m = 20: t = 11,
s = 9, rate 1/4, initial_k 4); tampering anywhere in the proof, a wrong
claim, a wrong root or a wrong shape is rejected (the reference's own
rejection tests, tests/completeness_random.rs:67-92)."""
import numpy as np
import pytest

import bitz_protocol as P


@pytest.fixture(scope="module")
def proved():
    t, s = 11, 9
    rng = np.random.default_rng(5)
    rows = rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)
    pr = P.OracleProver(rows, t, s, 2, 4)
    claimed = pr.evaluation()
    return pr, claimed, pr.prove(claimed)


def test_honest_proof_verifies(proved):
    pr, claimed, proof = proved
    P.verify_standalone(pr.root, pr.t, pr.s, 2, 4, claimed, proof)


def test_wrong_claim_root_or_shape_rejected(proved):
    pr, claimed, proof = proved
    with pytest.raises(P.VerifyError):
        P.verify_standalone(pr.root, pr.t, pr.s, 2, 4, claimed ^ 1, proof)
    bad_root = bytes([pr.root[0] ^ 1]) + pr.root[1:]
    with pytest.raises(P.VerifyError):
        P.verify_standalone(bad_root, pr.t, pr.s, 2, 4, claimed, proof)
    with pytest.raises((P.VerifyError, ValueError)):
        P.verify_standalone(pr.root, pr.t + 1, pr.s - 1, 2, 4, claimed, proof)


def test_tampering_rejected(proved):
    pr, claimed, proof = proved
    narg_len = int.from_bytes(proof[:8], "little")
    n = len(proof)
    # folds, GKR messages, sumcheck, ring-switch claims (narg) and the Ligerito
    # hint (rows, Merkle paths, yr, messages, nonces) and the Round-0 tail
    positions = [9, 9 + 16 * 512 + 5, 9 + narg_len // 2, 9 + narg_len - 3, 8 + narg_len + 8 + 40,
                 8 + narg_len + 8 + 600, n - 300, n - 40, n - 20, n - 2]
    for pos in positions:
        bad = bytearray(proof)
        bad[pos] ^= 0x10
        with pytest.raises(P.VerifyError):
            P.verify_standalone(pr.root, pr.t, pr.s, 2, 4, claimed, bytes(bad))
    with pytest.raises(P.VerifyError):
        P.verify_standalone(pr.root, pr.t, pr.s, 2, 4, claimed, proof + b"\0")
    with pytest.raises(P.VerifyError):
        P.verify_standalone(pr.root, pr.t, pr.s, 2, 4, claimed, proof[:-1])


def test_dishonest_claim_is_not_accepted(proved):
    pr, claimed, _ = proved
    wrong = claimed - 1 if claimed else 1
    bad_proof = pr.prove(wrong)  # the folds still reconstruct the honest value
    with pytest.raises(P.VerifyError):
        P.verify_standalone(pr.root, pr.t, pr.s, 2, 4, wrong, bad_proof)


def test_extra_recursive_level_rejected(proved):
    """one more recursive root and opening than the ladder's levels
    (the flock verifier's recursive_roots.len() != r check)"""
    pr, claimed, proof = proved
    bad = P.with_extra_recursive_level(proof)
    assert bad != proof
    with pytest.raises(P.VerifyError, match="unconsumed recursive"):
        P.verify_standalone(pr.root, pr.t, pr.s, 2, 4, claimed, bad)
