"""The library's succinct evaluation of the Ligerito verifier's basis at the
residual (bitz-poc_amd/csrc/residual.cpp), against the oracle's dense tables
folded least significant coordinate first: the ring-switch basis
sum_i bw_i bit_i(eq(z, x)) (fold_b128_elems of the eq table), the induced
basis of a query set (induce_basis, pinned through the NTT) and a scaled eq
table, each at several splits between folded and residual coordinates. The
dense and succinct forms agreeing at every binary residual point is the
identity the reference's own tests check
(pcs/ligerito.rs:7366-7420, pcs/ring_switch.rs:2797-2835)."""
import ctypes as C
import os
import random
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def rh(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("residual") / "libresidual.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so,
                    os.path.join(HERE, "residual_harness.cpp")], check=True)
    lib = C.CDLL(so)
    P = C.c_void_p
    lib.rh_transpose128.argtypes = [P, P]
    lib.rh_ring_switch.argtypes = [P, C.c_int, P, C.c_int, P, P]
    lib.rh_induced.argtypes = [C.c_int, P, C.c_int, P, P, P, P, C.c_int, P]
    lib.rh_eq_point.argtypes = [P, C.c_int, P, C.c_int, P, P]
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _g(xs):
    return np.ascontiguousarray(O.ints_to_gf(xs))


def _fold_all(table, rs):
    v = table
    for r in rs:
        v = O.fold_lsb(v, r)
    return O.gf_to_ints(v)


def test_transpose128(rh):
    rng = random.Random(5)
    xs = [rng.getrandbits(128) for _ in range(128)]
    out = np.zeros((128, 2), np.uint64)
    rh.rh_transpose128(_p(_g(xs)), _p(out))
    got = O.gf_to_ints(out)
    for b in range(128):
        want = sum(((xs[i] >> b) & 1) << i for i in range(128))
        assert got[b] == want


@pytest.mark.parametrize("nz,nr", [(1, 0), (1, 1), (6, 3), (9, 0), (9, 5), (9, 9), (11, 7)])
def test_ring_switch_residual_equals_dense_fold(rh, nz, nr):
    rng = random.Random(100 * nz + nr)
    z = [rng.getrandbits(128) for _ in range(nz)]
    r = [rng.getrandbits(128) for _ in range(nr)]
    bw = [rng.getrandbits(128) for _ in range(128)]
    dense = O.fold_b128_elems(O.eq_table(_g(z)), _g(bw))
    want = _fold_all(dense, r)
    out = np.zeros((1 << (nz - nr), 2), np.uint64)
    rh.rh_ring_switch(_p(_g(z)), nz, _p(_g(r if r else [0])), nr, _p(_g(bw)), _p(out))
    assert O.gf_to_ints(out) == want


def test_ring_switch_sparse_weights(rh):
    """bw = e_j picks bit j of eq(z, x): each of a few bit positions alone"""
    rng = random.Random(9)
    nz, nr = 7, 4
    z = [rng.getrandbits(128) for _ in range(nz)]
    r = [rng.getrandbits(128) for _ in range(nr)]
    eqz = O.gf_to_ints(O.eq_table(_g(z)))
    for j in (0, 1, 63, 64, 100, 127):
        bw = [0] * 128
        bw[j] = 1
        dense = _g([(e >> j) & 1 for e in eqz])
        want = _fold_all(dense, r)
        out = np.zeros((1 << (nz - nr), 2), np.uint64)
        rh.rh_ring_switch(_p(_g(z)), nz, _p(_g(r)), nr, _p(_g(bw)), _p(out))
        assert O.gf_to_ints(out) == want


@pytest.mark.parametrize("log_n,yr_log,nq", [(1, 0, 3), (5, 2, 7), (8, 0, 20), (8, 3, 33), (8, 8, 5), (10, 4, 64)])
def test_induced_residual_equals_dense_fold(rh, log_n, yr_log, nq):
    rng = random.Random(log_n * 31 + yr_log * 7 + nq)
    queries = np.array(sorted(rng.sample(range(1 << (log_n + 2)), nq)), np.uint64)
    aw = [rng.getrandbits(128) for _ in range(nq)]
    r = [rng.getrandbits(128) for _ in range(log_n - yr_log)]
    want = _fold_all(O.induce_basis(log_n, queries, _g(aw)), r)
    sks = O.gf_to_ints(O.eval_sk_at_vks(log_n))
    inv = [O.inv(x) for x in sks]
    out = np.zeros((1 << yr_log, 2), np.uint64)
    rh.rh_induced(log_n, _p(queries), nq, _p(_g(aw)), _p(_g(sks)), _p(_g(inv)), _p(_g(r if r else [0])),
                  yr_log, _p(out))
    assert O.gf_to_ints(out) == want


@pytest.mark.parametrize("nz,yr_log", [(1, 1), (4, 0), (7, 3), (9, 9)])
def test_eq_point_residual_equals_dense_fold(rh, nz, yr_log):
    rng = random.Random(nz * 13 + yr_log)
    z = [rng.getrandbits(128) for _ in range(nz)]
    r = [rng.getrandbits(128) for _ in range(nz - yr_log)]
    scale = rng.getrandbits(128)
    dense = O.gf_to_ints(O.eq_table(_g(z)))
    want = _fold_all(_g([O.mul(scale, e) for e in dense]), r)
    out = np.zeros((1 << yr_log, 2), np.uint64)
    rh.rh_eq_point(_p(_g(z)), nz, _p(_g(r if r else [0])), yr_log, _p(_g([scale])), _p(out))
    assert O.gf_to_ints(out) == want
