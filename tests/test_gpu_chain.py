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


def test_library_verifier_accepts_and_rejects():
    """bitz_verify_sha256_chain replays the pinned outer transcript and checks
    the virtual opening; a wrong digest, a wrong block and tampered bytes
    are rejected"""
    from bitz_amd import _lib
    from bitz_amd.chain import verify_chain
    blocks = np.array(C.pin_blocks(128), dtype=np.uint32)
    log("sha256 chain, library verifier at the pin instance")
    root, digest, state, proof = _prove(blocks)
    assert verify_chain(ctx(), blocks, digest, root, proof) == state
    bad = list(digest)
    bad[5] ^= 1 << 30
    with pytest.raises(_lib.BitzError):
        verify_chain(ctx(), blocks, bad, root, proof)
    b2 = blocks.copy()
    b2[100, 7] ^= 4
    with pytest.raises(_lib.BitzError):
        verify_chain(ctx(), b2, digest, root, proof)
    for pos in (12, 16 * 300 + 5, len(proof) // 3, len(proof) - 30, len(proof) - 1):
        t = bytearray(proof)
        t[pos] ^= 0x20
        with pytest.raises(_lib.BitzError):
            verify_chain(ctx(), blocks, digest, root, bytes(t))
    with pytest.raises(_lib.BitzError):
        verify_chain(ctx(), blocks, digest, root, proof[:-1])


def _be_words(b: bytes):
    return [int.from_bytes(b[4 * i: 4 * i + 4], "big") for i in range(len(b) // 4)]


def test_config4_chain_of_2p12_compressions():
    """2^12 chained compressions over a seeded random 256 KiB message: the
    chain's digest is SHA-256's state before the padding block (hashlib
    pins it through the final compression), and the library verifier
    accepts the proof"""
    import hashlib
    from bitz_amd.chain import verify_chain
    rng = np.random.default_rng(12)
    msg = rng.integers(0, 256, size=64 << 12, dtype=np.uint8).tobytes()
    blocks = np.array([_be_words(msg[64 * i: 64 * i + 64]) for i in range(1 << 12)], dtype=np.uint32)
    log("sha256 chain, 2^12 compressions")
    root, digest, state, proof = _prove(blocks)
    pad = b"\x80" + bytes(55) + (8 * len(msg)).to_bytes(8, "big")
    assert C.compress(digest, _be_words(pad)) == _be_words(hashlib.sha256(msg).digest())
    log(f"library verifier on the 2^12 proof ({len(proof)} bytes)")
    assert verify_chain(ctx(), blocks, digest, root, proof) == state
