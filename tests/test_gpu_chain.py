"""The SHA-256 chain proof on the device (bitz_prove_sha256_chain) against
the CPU restatement (oracle/sha256_chain.py), which reproduces the
reference's pinned outer transcript digest for 2^7 chained compressions
(tests/transcript_state_pins.rs:471-497).

* At the pinned instance the device's outer transcript state equals the
  reference pin, and its root, digest and proof bytes equal the oracle's.
* Random blocks at 2^8 compressions: the same equalities, and the CPU
  verifier accepts the device proof and rejects it for another digest.
"""
import numpy as np
import pytest

import oracle as O
import sha256_chain as C
from gpu_util import ctx, log

pytestmark = pytest.mark.gpu

PIN_LAMBDA100 = "56d2484895a33caa0e6741b60c22198eba0cd8de1b84cb8f9abe47603de3cc32"


def _prove(blocks):
    from bitz_amd.chain import prove_chain
    return prove_chain(ctx(), np.array(blocks, dtype=np.uint32))


def test_reference_pin_and_oracle_bytes():
    blocks = C.pin_blocks(128)
    log("sha256 chain, 2^7 pin instance")
    root, digest, state, proof = _prove(blocks)
    assert state.hex() == PIN_LAMBDA100
    pr = C.ChainProver(7, blocks)
    tr, out = pr.prove()
    assert digest == pr.wit.digest
    assert root == pr.pcs.root
    assert O.blake3(bytes(tr.buf), 32) == state
    assert proof == out["proof"]


def test_random_chain_equals_oracle_and_verifies():
    rng = np.random.default_rng(88)
    blocks = rng.integers(0, 2**32, size=(256, 16), dtype=np.uint64).astype(np.uint32)
    log("sha256 chain, 2^8 random blocks")
    root, digest, state, proof = _prove(blocks)
    pr = C.ChainProver(8, blocks.tolist())
    tr, out = pr.prove()
    assert root == pr.pcs.root and digest == pr.wit.digest
    assert proof == out["proof"]
    vt = C.verify_chain(8, blocks.tolist(), digest, root, proof)
    assert O.blake3(bytes(vt.buf), 32) == state
    bad = list(digest)
    bad[0] ^= 1
    with pytest.raises(C.P.VerifyError):
        C.verify_chain(8, blocks.tolist(), bad, root, proof)
