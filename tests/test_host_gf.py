"""The library's host GF(2^128) loops (bitz-poc_amd/csrc/host_gf.cpp: the
last rounds of each sumcheck run on the host) against the oracle, for every
multiply path this CPU has: the integer carry-less product (0), PCLMULQDQ
(1), and the dispatched entry points (2: four elements per 512-bit register
with VPCLMULQDQ where present, at sizes that are multiples of four). The
oracle's field is itself pinned by the NIST GCM vectors
(test_oracle_gf128.py)."""
import ctypes as C
import os
import random
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def hg(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("hostgf") / "libhostgf.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so,
                    os.path.join(HERE, "host_gf_harness.cpp")], check=True)
    lib = C.CDLL(so)
    P = C.c_void_p
    lib.hg_has_clmul.restype = C.c_int
    lib.hg_has_vpclmul.restype = C.c_int
    lib.hg_mul.argtypes = [C.c_int, P, P, P, C.c_uint64]
    lib.hg_eq_table.argtypes = [C.c_int, P, C.c_int, P]
    lib.hg_round_sums.argtypes = [C.c_int, P, P, C.c_uint64, P, C.c_int, P]
    lib.hg_fold.argtypes = [C.c_int, P, C.c_uint64, P]
    lib.hg_inner_round.argtypes = [C.c_int, P, P, C.c_uint64, P]
    lib.hg_fold_lsb.argtypes = [C.c_int, P, C.c_uint64, P]
    lib.hg_drop_top.argtypes = [P, C.c_uint64]
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _paths(hg):
    return [0, 1, 2] if hg.hg_has_clmul() else [0, 2]


def _rand(rng, n):
    return [rng.getrandbits(128) for _ in range(n)]


def test_mul_matches_oracle(hg):
    rng = random.Random(11)
    edge = [0, 1, 2, (1 << 128) - 1, 1 << 127, (1 << 64) - 1, 1 << 64, 0xF << 60, (0xF << 60) << 64]
    a = edge + _rand(rng, 300) + edge[::-1]
    b = edge[::-1] + _rand(rng, 300) + edge
    A, B = O.ints_to_gf(a), O.ints_to_gf(b)
    want = [O.mul(x, y) for x, y in zip(a, b)]
    for path in _paths(hg):
        out = np.zeros_like(A)
        hg.hg_mul(path, _p(A), _p(B), _p(out), len(a))
        assert O.gf_to_ints(out) == want, f"path {path}"


@pytest.mark.parametrize("n", [0, 1, 3, 8])
def test_eq_table_matches_oracle(hg, n):
    rng = random.Random(n)
    r = O.ints_to_gf(_rand(rng, n)) if n else np.zeros((1, 2), np.uint64)
    want = O.gf_to_ints(O.eq_table(r[:n])) if n else [1]
    for path in _paths(hg):
        out = np.zeros((1 << n, 2), np.uint64)
        hg.hg_eq_table(path, _p(r), n, _p(out))
        assert O.gf_to_ints(out) == want


@pytest.mark.parametrize("h,send_one", [(1, 0), (1, 1), (16, 0), (1024, 1), (1024, 0)])
def test_round_sums_match_oracle(hg, h, send_one):
    rng = random.Random(h * 2 + send_one)
    l = O.ints_to_gf(_rand(rng, 2 * h))
    r = O.ints_to_gf(_rand(rng, 2 * h))
    w = O.ints_to_gf(_rand(rng, h))
    want = list(O.gkr_round_sums(l, r, w, bool(send_one)))
    for path in _paths(hg):
        out = np.zeros((2, 2), np.uint64)
        hg.hg_round_sums(path, _p(l), _p(r), h, _p(w), send_one, _p(out))
        assert O.gf_to_ints(out) == want


@pytest.mark.parametrize("h", [1, 7, 512])
def test_fold_matches_oracle(hg, h):
    rng = random.Random(100 + h)
    v = O.ints_to_gf(_rand(rng, 2 * h))
    rho = rng.getrandbits(128)
    want = O.gf_to_ints(O.fold_halves(v, rho))
    for path in _paths(hg):
        x = v.copy()
        hg.hg_fold(path, _p(x), h, _p(O.ints_to_gf([rho])))
        assert O.gf_to_ints(x[:h]) == want


@pytest.mark.parametrize("n", [2, 8, 2048])
def test_inner_round_matches_oracle(hg, n):
    rng = random.Random(200 + n)
    m = O.ints_to_gf(_rand(rng, n))
    w = O.ints_to_gf(_rand(rng, n))
    c0, c1, c2 = O.round_polynomial(m, w)
    for path in _paths(hg):
        out = np.zeros((3, 2), np.uint64)
        hg.hg_inner_round(path, _p(m), _p(w), n // 2, _p(out))
        v0, v1, v2 = O.gf_to_ints(out)
        assert (v0, v0 ^ v1 ^ v2, v2) == (c0, c1, c2)


@pytest.mark.parametrize("n", [2, 6, 1024])
def test_fold_lsb_matches_oracle(hg, n):
    rng = random.Random(300 + n)
    v = O.ints_to_gf(_rand(rng, n))
    r = rng.getrandbits(128)
    want = O.gf_to_ints(O.fold_lsb(v, r))
    for path in _paths(hg):
        x = v.copy()
        hg.hg_fold_lsb(path, _p(x), n // 2, _p(O.ints_to_gf([r])))
        assert O.gf_to_ints(x[: n // 2]) == want


def test_drop_top_gives_the_smaller_eq_table(hg):
    # eq over r[0..n) with its halves added is eq over r[0..n-1): the two
    # weights of the last coordinate, r and 1 + r, sum to one
    rng = random.Random(400)
    r = _rand(rng, 7)
    t = O.eq_table(O.ints_to_gf(r)).copy()
    hg.hg_drop_top(_p(t), 64)
    assert O.gf_to_ints(t[:64]) == O.gf_to_ints(O.eq_table(O.ints_to_gf(r[:6])))
