"""Pins the oracle's Merkle tree / multi-proof (restating
vendor/flock-mod/crates/flock-core/src/merkle.rs) and PCS commit (restating
pcs/commit.rs:180-315) from the pinned BLAKE3 leaf/parent primitives and the
pinned NTT: the tree is rebuilt node by node, multi-proofs verify for random
position sets and fail when tampered, and the commit's codeword is the full
NTT of the zero-padded message with the tree over its 16*2^k-byte leaves."""
import random

import numpy as np
import pytest

import oracle as O


def test_tree_matches_hand_built():
    rng = random.Random(1)
    for n_leaves, leaf in [(1, 64), (2, 32), (8, 256), (16, 100), (64, 128)]:
        data = bytes(rng.getrandbits(8) for _ in range(n_leaves * leaf))
        tree = O.merkle_tree(data, n_leaves)
        level = [O.leaf_cv(data[i * leaf:(i + 1) * leaf]) for i in range(n_leaves)]
        flat = list(level)
        while len(level) > 1:
            level = [O.parent_cv(level[2 * i], level[2 * i + 1]) for i in range(len(level) // 2)]
            flat += level
        assert [bytes(h) for h in tree] == flat


@pytest.mark.parametrize("n_leaves", [2, 16, 1024])
def test_multi_proof_roundtrip_and_tamper(n_leaves):
    rng = random.Random(n_leaves)
    data = bytes(rng.getrandbits(8) for _ in range(n_leaves * 64))
    tree = O.merkle_tree(data, n_leaves)
    root = bytes(tree[-1])
    for trial in range(10):
        k = rng.randint(1, min(40, n_leaves))
        pos = [rng.randrange(n_leaves) for _ in range(k)]
        proof = O.merkle_multi_proof(tree, n_leaves, pos)
        sp = sorted(set(pos))
        leaves = np.stack([tree[p] for p in sp])
        assert O.merkle_verify_multi(root, n_leaves, sp, leaves, proof)
        if len(proof):
            bad = proof.copy()
            bad[rng.randrange(len(bad)), rng.randrange(32)] ^= 1
            assert not O.merkle_verify_multi(root, n_leaves, sp, leaves, bad)
            assert not O.merkle_verify_multi(root, n_leaves, sp, leaves, proof[:-1])
        bad_leaves = leaves.copy()
        bad_leaves[0, 0] ^= 0x80
        assert not O.merkle_verify_multi(root, n_leaves, sp, bad_leaves, proof)
    # all leaves opened: empty proof
    allp = list(range(n_leaves))
    assert len(O.merkle_multi_proof(tree, n_leaves, allp)) == 0


@pytest.mark.parametrize("log_msg,log_batch,r", [(6, 2, 1), (8, 4, 2), (9, 3, 3), (10, 4, 2)])
def test_commit_is_ntt_then_merkle(log_msg, log_batch, r):
    rng = np.random.default_rng(log_msg * 7 + r)
    msg = rng.integers(0, 2**64, size=(1 << log_msg, 2), dtype=np.uint64)
    root, cw, tree = O.commit(msg, log_msg, log_batch, r)
    padded = np.zeros_like(cw)
    padded[: 1 << log_msg] = msg
    k_code = log_msg - log_batch + r
    assert np.array_equal(cw, O.ntt_forward(padded, k_code, log_batch))
    assert np.array_equal(tree, O.merkle_tree(cw.view(np.uint8).reshape(-1), 1 << k_code))
    assert root == bytes(tree[-1])
    # determinism and sensitivity (commit.rs:484-551)
    assert O.commit(msg, log_msg, log_batch, r)[0] == root
    msg2 = msg.copy()
    msg2[3, 1] ^= 1 << 40
    assert O.commit(msg2, log_msg, log_batch, r)[0] != root


def test_commit_rejects_bad_geometry():
    msg = np.zeros((16, 2), np.uint64)
    with pytest.raises(ValueError):
        O.commit(msg, 4, 4, 1)  # log_batch must leave a position variable
    with pytest.raises(ValueError):
        O.commit(msg, 4, 2, 0)  # rate must be < 1
