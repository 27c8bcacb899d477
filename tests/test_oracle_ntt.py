"""Pins the oracle's additive NTT (restating
vendor/flock-mod/crates/flock-core/src/ntt/additive_ntt_f128.rs) without the
reference: (1) brute force -- the forward transform evaluates the novel-basis
polynomial sum_j c_j X_j at the domain point whose integer encoding is the
output index, with X_j = prod_{i in bits(j)} What_i and What_i the normalized
subspace polynomials built from their recursive definition (:20-31); (2) the
inverse round trip and linearity (:1099-1160); (3) interleaved == per-lane;
(4) replicate-from-layer-r == full NTT of the zero-padded message
(pcs/commit.rs:443-482)."""
import random

import numpy as np
import pytest

import oracle as O


def _rand(rng, n):
    return np.array([[rng.getrandbits(64), rng.getrandbits(64)] for _ in range(n)], np.uint64)


def _to_int(a):
    return [int(x[0]) | (int(x[1]) << 64) for x in a]


def _what(i, x):
    """Normalized subspace polynomial What_i(x) over basis beta_k = X^k."""
    def w(k, z):
        if k == 0:
            return z
        prev = w(k - 1, z)
        return O.mul(prev, prev ^ w(k - 1, 1 << (k - 1)))
    return O.mul(w(i, x), O.inv(w(i, 1 << i)))


@pytest.mark.parametrize("log_d", [1, 2, 3, 4, 5])
def test_forward_is_novel_basis_evaluation(log_d):
    rng = random.Random(100 + log_d)
    n = 1 << log_d
    coeffs = _rand(rng, n)
    out = _to_int(O.ntt_forward(coeffs, log_d, 0))
    c = _to_int(coeffs)
    what = [[_what(i, p) for i in range(log_d)] for p in range(n)]
    for p in range(n):
        acc = 0
        for j in range(n):
            basis = 1
            for i in range(log_d):
                if (j >> i) & 1:
                    basis = O.mul(basis, what[p][i])
            acc ^= O.mul(c[j], basis)
        assert out[p] == acc, (log_d, p)


def test_twiddle_rows_are_normalized():
    for ell in (1, 4, 9):
        ev = O.ntt_evals(ell)
        assert all(row[0] == 1 for row in ev)
        # evals[i][k] = What_i(beta_{i+k})
        for i in range(min(ell, 4)):
            for k in range(ell - i):
                assert ev[i][k] == _what(i, 1 << (i + k))


@pytest.mark.parametrize("log_d,log_lanes", [(1, 0), (4, 0), (6, 2), (8, 4), (10, 3)])
def test_roundtrip_and_linearity(log_d, log_lanes):
    rng = random.Random(log_d * 31 + log_lanes)
    n = 1 << (log_d + log_lanes)
    a, b = _rand(rng, n), _rand(rng, n)
    fa = O.ntt_forward(a, log_d, log_lanes)
    assert np.array_equal(O.ntt_inverse(fa, log_d, log_lanes), a)
    assert np.array_equal(O.ntt_forward(O.ntt_inverse(a, log_d, log_lanes), log_d, log_lanes), a)
    fb = O.ntt_forward(b, log_d, log_lanes)
    assert np.array_equal(O.ntt_forward(a ^ b, log_d, log_lanes), fa ^ fb)
    assert not O.ntt_forward(np.zeros_like(a), log_d, log_lanes).any()


def test_interleaved_equals_per_lane():
    rng = random.Random(5)
    log_d, log_lanes = 7, 3
    lanes = 1 << log_lanes
    data = _rand(rng, (1 << log_d) * lanes)
    inter = O.ntt_forward(data, log_d, log_lanes)
    for lane in range(lanes):
        col = np.ascontiguousarray(data[lane::lanes])
        assert np.array_equal(O.ntt_forward(col, log_d, 0), inter[lane::lanes])


@pytest.mark.parametrize("r", [1, 2, 3])
def test_replicate_equals_zero_padded_full_ntt(r):
    rng = random.Random(r)
    log_msg, log_lanes = 6, 2
    msg = _rand(rng, 1 << log_msg)
    log_d = log_msg - log_lanes + r
    padded = np.zeros(((1 << (log_msg + r)), 2), np.uint64)
    padded[: 1 << log_msg] = msg
    full = O.ntt_forward(padded, log_d, log_lanes)
    repl = np.concatenate([msg] * (1 << r))
    part = O.ntt_forward(repl, log_d, log_lanes, start_layer=r)
    assert np.array_equal(full, part)
