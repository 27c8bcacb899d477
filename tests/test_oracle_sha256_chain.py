"""The SHA-256 chain proof's CPU restatement (oracle/sha256_chain.py).

The reference pins the outer transcript of `prove_sha256_chain` over 2^7
chained compressions (blocks[i][j] = (i * 0x9e3779b9) ^ j from the standard
initial state, tests/transcript_state_pins.rs:471-497) to the BLAKE3 state
digest 56d24848... That digest covers every byte the prover absorbs before
the opening fork: the compiled compression relation (the constraint matrix
and the four chained maps through their canonical digests), the chained map
digest, the commitment root of the packed source rows, the security and
ladder bindings, Round 0, the runtime prime, every challenge, the collapsed
local rows, the batched public statement and the rank-one claim's weights
and value. The restatement must reproduce it exactly.

The other tests pin what the digest leaves implicit: the native compression
against hashlib, the relation's shape and its satisfaction by an honest
witness (and failure for a wrong hint bit), the claim as a brute-force sum
over the product tensor, and the public slots' placement.
"""
import hashlib
import random

import numpy as np
import pytest

import oracle as O
import sha256_chain as C

PIN_LAMBDA100 = "56d2484895a33caa0e6741b60c22198eba0cd8de1b84cb8f9abe47603de3cc32"


def _words(b: bytes):
    return [int.from_bytes(b[4 * i: 4 * i + 4], "big") for i in range(len(b) // 4)]


def test_native_compression_against_hashlib():
    """SHA-256 of an N-block message is the compression of the padding
    block after the chain over the message blocks"""
    rng = random.Random(5)
    for nblocks in (1, 3, 8):
        msg = bytes(rng.getrandbits(8) for _ in range(64 * nblocks))
        state = list(C.IV)
        for i in range(nblocks):
            state = C.compress(state, _words(msg[64 * i: 64 * i + 64]))
        pad = b"\x80" + bytes(55) + (8 * len(msg)).to_bytes(8, "big")
        assert C.compress(state, _words(pad)) == _words(hashlib.sha256(msg).digest())


def _f_bar(state, block):
    hints, out = C.compression_hints(state, block)
    f = [1] + [(block[i // 32] >> (i % 32)) & 1 for i in range(512)]
    f += [(state[i // 32] >> (i % 32)) & 1 for i in range(256)]
    for v, w in hints:
        f += [(v >> b) & 1 for b in range(w)]
    return f, out


def _residuals(f):
    m_rows, c_rows, _ = C.relation()
    h = [sum(f[c] for c in row) & 1 for row in m_rows]
    return [sum(v * h[c] for c, v in r) for r in c_rows]


def test_relation_shape_and_satisfaction():
    m_rows, c_rows, out_bits = C.relation()
    assert len(m_rows) == C.H_BAR == 20457 and len(c_rows) == C.CONSTRAINTS == 184
    assert m_rows[0] == [0] and C.HINT_BITS == 6376
    # the public cells sit where sha256_public_h_column puts them
    # (sha256/constraints.rs:62-71): state lifts 1.., block lifts 257..,
    # the output sums' lifts 20193 + 33 w
    assert m_rows[1] == [1 + 512] and m_rows[257] == [1]
    assert m_rows[20193] == [C.OUTPUT_F_BASE]
    rng = random.Random(11)
    state = [rng.getrandbits(32) for _ in range(8)]
    block = [rng.getrandbits(32) for _ in range(16)]
    f, out = _f_bar(state, block)
    assert all(r == 0 for r in _residuals(f))
    # the circuit's output bits are the final sums' low hint bits
    vals = [f[next(iter(b.s)) + 1] for b in out_bits]
    assert [sum(vals[32 * w + b] << b for b in range(32)) for w in range(8)] == out
    # one wrong hint bit breaks a constraint
    for pos in (1 + 768 + 5, 1 + 768 + 48 * 34 + 17, len(f) - 3):
        g = list(f)
        g[pos] ^= 1
        assert any(r != 0 for r in _residuals(g))


def test_chain_maps_fix_the_public_statement():
    """chain_map_fixes_public_statement (chain.rs:1306-1345): every public
    derived cell is read from exactly its committed cell, block slots
    through `local`, terminal slots through `last`"""
    rel = C.ChainRelation()
    for slot in range(C.PUBLIC_BITS):
        hc = C.chain_public_h_column(slot)
        src = 1 + slot if slot < 512 else C.chain_output_f_column((slot - 512) // 32, slot % 32)
        owner = "local" if slot < 512 else "last"
        for name in ("local", "prev", "first", "last"):
            rows = rel.maps[name]["rows"][hc]
            assert rows == ([src] if name == owner else [])


@pytest.fixture(scope="module")
def pinned():
    pr = C.ChainProver(7, C.pin_blocks(128))
    tr, out = pr.prove(with_opening=False)
    return pr, tr, out


def test_outer_transcript_equals_reference_pin(pinned):
    _, tr, _ = pinned
    assert O.blake3(bytes(tr.buf), 32).hex() == PIN_LAMBDA100


def test_product_claim_by_brute_force(pinned):
    """the rank-one claim sum_{b,c} row_w[b] col_w[c] D(c, b) = mu (mod q)
    over the packed product tensor: the relation's collapsed rows vanish on
    the honest witness and the constant and public terms give mu"""
    pr, _, out = pinned
    q = out["q"]
    th, sh = pr.batch.h
    bits = np.unpackbits(pr.wit.product_rows.view(np.uint8), bitorder="little").reshape(1 << sh, 1 << th)
    acc = 0
    rw = out["row_w"]
    for c in np.nonzero(bits.any(axis=1))[0]:
        cw = out["col_w"][c]
        if cw:
            acc += cw * (sum(rw[b] for b in np.nonzero(bits[c])[0]) % q)
    assert acc % q == out["mu"]
    # the runtime prime lies in [2^112, 2^113) (derive_params at 2^7 instances)
    assert out["q"].bit_length() == 113


def test_chain_states_and_digest(pinned):
    pr, _, _ = pinned
    w = pr.wit
    assert len(w.states) == 129 and w.states[0] == C.IV
    for i in (0, 64, 127):
        assert w.states[i + 1] == C.compress(w.states[i], w.blocks[i])
    # source cells: the constant, then block and hint bits per instance
    assert w.source_cells[0] == 1 and len(w.source_cells) == 1 + 128 * C.CHAIN_F_INSTANCE


@pytest.fixture(scope="module")
def full_proof():
    pr = C.ChainProver(7, C.pin_blocks(128))
    tr, out = pr.prove(with_opening=True)
    return pr, tr, out


def test_full_chain_proof_verifies(full_proof):
    """the CPU verifier replays the pinned outer transcript and accepts the
    virtual opening"""
    pr, tr, out = full_proof
    assert O.blake3(bytes(tr.buf), 32).hex() == PIN_LAMBDA100
    vt = C.verify_chain(7, pr.wit.blocks, pr.wit.digest, pr.pcs.root, out["proof"])
    assert bytes(vt.buf) == bytes(tr.buf)


def test_wrong_statement_and_tampering_rejected(full_proof):
    pr, _, out = full_proof
    proof = out["proof"]
    bad_digest = list(pr.wit.digest)
    bad_digest[3] ^= 1 << 7
    with pytest.raises(C.P.VerifyError):
        C.verify_chain(7, pr.wit.blocks, bad_digest, pr.pcs.root, proof)
    blocks = [list(b) for b in pr.wit.blocks]
    blocks[40][2] ^= 1
    with pytest.raises(C.P.VerifyError):
        C.verify_chain(7, blocks, pr.wit.digest, pr.pcs.root, proof)
    for pos in (20, 16 * 200 + 3, len(proof) // 2, len(proof) - 40):
        bad = bytearray(proof)
        bad[pos] ^= 0x04
        with pytest.raises(C.P.VerifyError):
            C.verify_chain(7, pr.wit.blocks, pr.wit.digest, pr.pcs.root, bytes(bad))


def test_unsatisfying_witness_rejected():
    """one flipped hint bit: the derived grid still follows the map, the
    relation's collapsed rows no longer vanish, and the folds miss the claim"""
    pr = C.ChainProver(7, C.pin_blocks(128), flip_hint=(77, 300))
    _, out = pr.prove(with_opening=True)
    with pytest.raises(C.P.VerifyError):
        C.verify_chain(7, pr.wit.blocks, pr.wit.digest, pr.pcs.root, out["proof"])


def test_derived_grid_inconsistent_with_commitment_cannot_be_proved():
    """virt.rs:605-620: a product tensor that is not M f fails the opening's
    inner-product check in the prover"""
    pr = C.ChainProver(7, C.pin_blocks(128))
    _, out = pr.prove(with_opening=False)
    rows = pr.wit.product_rows.copy()
    rows[12345] ^= np.uint64(1 << 9)
    pr.wit.product_rows = rows
    with pytest.raises(AssertionError):
        pr.open(out)
