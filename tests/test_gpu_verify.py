"""The library's verifier (bitz_verify_standalone: transcript checks on the
host, dense tables on the device) against the CPU restatement's verifier:
both accept the honest proofs of the parity shapes (m = 20) and of the
2^20-coefficient workload (n = 26), and both reject the same tampered proofs
(the positions of tests/test_oracle_protocol.py), a wrong claim, root or
shape, trailing and missing bytes, and a dishonest claim's proof."""
import numpy as np
import pytest

import bitz_protocol as P
from gpu_util import ctx, log

pytestmark = pytest.mark.gpu


def _rows(t, s, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)


def _gpu_accepts(root, t, s, rate, k0, claimed, proof) -> bool:
    from bitz_amd._lib import BitzError
    from bitz_amd.device import verify_standalone
    try:
        verify_standalone(ctx(), root, t, s, rate, k0, claimed, proof)
        return True
    except BitzError as e:
        assert e.code in (-5, -1), str(e)  # rejected (VERIFY) or a shape the ladder refuses
        return False


def _cpu_accepts(root, t, s, rate, k0, claimed, proof) -> bool:
    try:
        P.verify_standalone(root, t, s, rate, k0, claimed, proof)
        return True
    except (P.VerifyError, ValueError):
        return False


@pytest.fixture(scope="module", params=[(11, 9, 2, 4, 5), (12, 8, 1, 4, 6), (10, 10, 3, 3, 7)],
                ids=["t11s9r2", "t12s8r1", "t10s10r3"])
def proved(request):
    from bitz_amd.device import Commitment
    t, s, rate, k0, seed = request.param
    com = Commitment(ctx(), _rows(t, s, seed), t, s, k0, rate)
    claimed = com.evaluation()
    proof = com.prove(claimed)
    yield (t, s, rate, k0), com, claimed, proof
    com.close()


def test_honest_proof_accepted(proved):
    (t, s, rate, k0), com, claimed, proof = proved
    log(f"verify honest t={t} s={s} rate={rate} k0={k0}")
    assert _gpu_accepts(com.root, t, s, rate, k0, claimed, proof)


def test_rejections_agree_with_cpu_verifier(proved):
    (t, s, rate, k0), com, claimed, proof = proved
    log(f"verify tampered t={t} s={s}")
    narg_len = int.from_bytes(proof[:8], "little")
    n = len(proof)
    positions = [9, 9 + 16 * (1 << s) // 2 + 5, 9 + narg_len // 2, 9 + narg_len - 3, 8 + narg_len + 8 + 40,
                 8 + narg_len + 8 + 600, n - 300, n - 40, n - 20, n - 2]
    cases = []
    for pos in positions:
        bad = bytearray(proof)
        bad[pos] ^= 0x10
        cases.append((com.root, t, s, claimed, bytes(bad)))
    cases += [(com.root, t, s, claimed ^ 1, proof), (bytes([com.root[0] ^ 1]) + com.root[1:], t, s, claimed, proof),
              (com.root, t, s, claimed, proof + b"\0"), (com.root, t, s, claimed, proof[:-1]),
              (com.root, t, s, claimed, b""), (com.root, t + 1, s - 1, claimed, proof)]
    for root, tt, ss, c, pf in cases:
        assert not _gpu_accepts(root, tt, ss, rate, k0, c, pf)
        assert not _cpu_accepts(root, tt, ss, rate, k0, c, pf)


def test_dishonest_claim_rejected(proved):
    (t, s, rate, k0), com, claimed, _ = proved
    wrong = claimed - 1 if claimed else 1
    bad_proof = com.prove(wrong)
    assert not _gpu_accepts(com.root, t, s, rate, k0, wrong, bad_proof)


def test_verifies_2p20_coefficient_proof():
    from bitz_amd.device import Commitment
    t, s = 15, 11
    log("verify n=26")
    com = Commitment(ctx(), _rows(t, s, 2024), t, s, 4, 2)
    claimed = com.evaluation()
    proof = com.prove(claimed)
    assert _gpu_accepts(com.root, t, s, 2, 4, claimed, proof)
    bad = bytearray(proof)
    bad[len(proof) // 3] ^= 4
    assert not _gpu_accepts(com.root, t, s, 2, 4, claimed, bytes(bad))
    assert not _gpu_accepts(com.root, t, s, 2, 4, claimed ^ 1, proof)
    com.close()


def test_random_byte_flips_rejected(proved):
    """40 seeded single-bit flips over the whole proof: every one is rejected
    by the library's verifier (and by the CPU verifier for a sample)"""
    (t, s, rate, k0), com, claimed, proof = proved
    rng = np.random.default_rng(7 + t)
    positions = rng.choice(len(proof), size=40, replace=False)
    for i, pos in enumerate(positions):
        bad = bytearray(proof)
        bad[int(pos)] ^= 1 << int(rng.integers(0, 8))
        assert not _gpu_accepts(com.root, t, s, rate, k0, claimed, bytes(bad)), f"flip at {pos} accepted"
        if i < 4:
            assert not _cpu_accepts(com.root, t, s, rate, k0, claimed, bytes(bad))


def test_extra_recursive_level_rejected(proved):
    """a proof with one more recursive root and opening appended to its
    Ligerito hint is rejected by both verifiers"""
    (t, s, rate, k0), com, claimed, proof = proved
    log(f"verify extra recursive level t={t} s={s}")
    bad = P.with_extra_recursive_level(proof)
    assert not _cpu_accepts(com.root, t, s, rate, k0, claimed, bad)
    assert not _gpu_accepts(com.root, t, s, rate, k0, claimed, bad)
