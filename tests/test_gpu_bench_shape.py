"""The opening's bit-pattern kernels at the benchmarked shape (2^24 u64
coefficients: t = 17, s = 13, n = 30), against the dense kernels they
replace, which test_gpu_open.py pins against the oracle. The dense leaves
(1 + (A[y] - 1) bit(c, y), index (y << s) | c) are built on the device with
torch from the same bits; the leaf layer's rounds 0 and 1 and its double
fold, level 1's round 0 and fold (with and without carried column weights)
and levels 1..3 are compared bit for bit on the device, and so are the
table rounds of levels 0-3 and level 4 from the row words. Then GPU proofs
at n = 28 and n = 30 are checked by the CPU verifier."""
import time

import numpy as np
import pytest

import bitz_protocol as P
from gpu_util import ctx, log, rand_gf, to_dev

pytestmark = pytest.mark.gpu

T, S = 17, 13


def _eq_parts(point, s_c):
    """tensor weights over the point: eq_c over the low s_c coordinates,
    eq_y over the rest (the layout of the GKR dense rounds)"""
    import oracle as O
    pt = O.ints_to_gf(point)
    return O.eq_table(pt[:s_c]), O.eq_table(pt[s_c:])


def _rand_ints(k, seed):
    rng = np.random.default_rng(seed)
    return [int.from_bytes(rng.bytes(16), "little") for _ in range(k)]


@pytest.fixture(scope="module")
def bench_shape():
    import torch
    from bitz_amd.device import OpenOps
    ops = OpenOps(ctx())
    rng = np.random.default_rng(2401)
    rows = torch.from_numpy(rng.integers(0, 2**64, size=(1 << S) << (T - 6), dtype=np.uint64).view(np.int64))
    rows = rows.to("cuda")
    A = rand_gf(1 << T, 2402)
    A[::9] = [1, 0]
    dA = to_dev(A)
    packed = ops.pack_columns(rows, T, S)
    # dense leaves on the device: bit (c, y) = bit y % 64 of word y / 64 of column c
    shifts = torch.arange(64, device="cuda", dtype=torch.int64)
    bits = ((rows.view(1 << S, 1 << (T - 6), 1) >> shifts) & 1).to(torch.bool).view(1 << S, 1 << T)
    bits = bits.t().contiguous()  # [y, c]
    one = torch.tensor([1, 0], dtype=torch.int64, device="cuda")
    leaves = torch.where(bits.unsqueeze(-1), dA.view(1 << T, 1, 2), one.view(1, 1, 2)).reshape(-1, 2).contiguous()
    del bits
    yield ops, rows, packed, dA, leaves
    del leaves
    torch.cuda.empty_cache()


def test_leaf_layer_rounds_0_and_1(bench_shape):
    import torch
    ops, rows, packed, dA, leaves = bench_shape
    log(f"bench shape t={T} s={S}: leaf rounds 0-1 against the dense kernels")
    half = 1 << (T + S - 1)
    h0, h1 = half >> 1, half >> 2
    l, r = leaves[:half].clone(), leaves[half:].clone()
    point = _rand_ints(T + S - 2, 1)
    eq_c, eq_y = _eq_parts(point, S)
    for so in (False, True):
        want = ops.gkr_round_sums(l, r, h0, to_dev(eq_c), S, to_dev(eq_y), so)
        got = ops.gkr_leaf_round_sums(packed, dA, T, S, to_dev(eq_c), S, to_dev(eq_y), so)
        assert tuple(got) == tuple(want)
    rho0, rho1 = _rand_ints(2, 2)
    gl, gr = ops.gkr_leaf_fold(packed, dA, T, S, rho0)
    ops.fold_halves_(l, h0, rho0)
    ops.fold_halves_(r, h0, rho0)
    assert torch.equal(gl[:h0], l[:h0]) and torch.equal(gr[:h0], r[:h0])
    del gl, gr
    point1 = _rand_ints(T + S - 3, 3)
    eq_c1, eq_y1 = _eq_parts(point1, S)
    for so in (False, True):
        want = ops.gkr_round_sums(l, r, h1, to_dev(eq_c1), S, to_dev(eq_y1), so)
        got = ops.gkr_leaf_round1_sums(packed, dA, T, S, rho0, to_dev(eq_c1), to_dev(eq_y1), so)
        assert tuple(got) == tuple(want)
    fl, fr = ops.gkr_leaf_fold2(packed, dA, T, S, rho0, rho1)
    ops.fold_halves_(l, h1, rho1)
    ops.fold_halves_(r, h1, rho1)
    assert torch.equal(fl, l[:h1]) and torch.equal(fr, r[:h1])


def test_levels_and_level1_round0(bench_shape):
    import torch
    ops, rows, packed, dA, leaves = bench_shape
    log(f"bench shape t={T} s={S}: levels 1-3 and level-1 round 0 against the dense kernels")
    lvl1 = ops.product_level(leaves)
    assert torch.equal(ops.gkr_level1(packed, dA, T, S), lvl1)
    lvl2 = ops.product_level(lvl1)
    assert torch.equal(ops.gkr_level2(packed, dA, T, S), lvl2)
    lvl3 = ops.product_level(lvl2)
    del lvl2
    assert torch.equal(ops.gkr_level3(packed, dA, T, S), lvl3)
    del lvl3
    h = 1 << (T + S - 3)
    l, r = lvl1[:2 * h].clone(), lvl1[2 * h:].clone()
    del lvl1
    point = _rand_ints(T + S - 3, 4)
    eq_c, eq_y = _eq_parts(point, S)
    for so in (False, True):
        want = ops.gkr_round_sums(l, r, h, to_dev(eq_c), S, to_dev(eq_y), so)
        got = ops.gkr_level1_round0_sums(packed, dA, T, S, to_dev(eq_c), to_dev(eq_y), so)
        assert tuple(got) == tuple(want)
    rho = _rand_ints(1, 5)[0]
    ops.fold_halves_(l, h, rho)
    ops.fold_halves_(r, h, rho)
    gl, gr = ops.gkr_level1_fold(packed, dA, T, S, rho)
    assert torch.equal(gl, l[:h]) and torch.equal(gr, r[:h])
    del gl, gr
    # with the column weights carried by the right half
    wl, wr = ops.gkr_level1_fold(packed, dA, T, S, rho, to_dev(eq_c))
    assert torch.equal(wl, l[:h])
    want_r = ctx().gf_mul(r[:h].contiguous(), to_dev(np.tile(eq_c, (h >> S, 1))))
    assert torch.equal(wr, want_r)


@pytest.mark.parametrize("level", [3, 2, 1, 0])
def test_table_round_at_bench_shape(bench_shape, level):
    """the table round of the level-l layer (sums, plain and eq_hi-weighted
    folds) and level 4 from the row words, against the dense kernels on the
    dense level folded by the same challenges"""
    import torch
    from test_gpu_nibble import _jit_fold, _jit_round_sums, _nibble_rows
    ops, rows, packed, dA, leaves = bench_shape
    log(f"bench shape t={T} s={S}: table round of level {level} against the dense kernels")
    nr = _nibble_rows(packed, T, S)
    lvl = leaves
    for _ in range(level):
        lvl = ops.product_level(lvl)
    half = lvl.shape[0] // 2
    l, r = lvl[:half].clone(), lvl[half:].clone()
    del lvl
    j = 3 - level
    chal = _rand_ints(j + 1, 40 + level)
    h = half
    for i in range(j):
        h //= 2
        ops.fold_halves_(l, h, chal[i])
        ops.fold_halves_(r, h, chal[i])
    assert h == 1 << (T + S - 4)
    point = _rand_ints(T + S - 5, 50 + level)
    eq_c, eq_y = _eq_parts(point, S)
    hh = h // 2
    for so in (False, True):
        want = ops.gkr_round_sums(l[:h].contiguous(), r[:h].contiguous(), hh, to_dev(eq_c), S, to_dev(eq_y), so)
        got = _jit_round_sums(level, nr, dA, T, S, chal[:j], to_dev(eq_c), to_dev(eq_y), so)
        assert got == tuple(want)
    import oracle as O
    eq_hi = O.eq_table(O.ints_to_gf(_rand_ints(S - 8, 60 + level)))
    wl, wr = _jit_fold(level, nr, dA, T, S, chal, to_dev(eq_hi))
    ops.fold_halves_(l, hh, chal[j])
    ops.fold_halves_(r, hh, chal[j])
    assert torch.equal(wl, l[:hh])
    del wl
    cols_hi = torch.from_numpy(np.repeat(eq_hi, 256, axis=0).view(np.int64)).to("cuda")
    assert torch.equal(wr, ctx().gf_mul(r[:hh].contiguous(), cols_hi.repeat(hh >> S, 1).contiguous()))
    del wr, l, r
    if level == 3:
        from bitz_amd import _lib
        from test_gpu_nibble import _vp
        lv = leaves
        for _ in range(4):
            lv = ops.product_level(lv)
        out = ctx().empty_gf(1 << (T + S - 4))
        _lib.call("bitz_gkr_level4_dev", ctx().ptr, _vp(nr), _vp(dA), T, S, _vp(out))
        assert torch.equal(out, lv)
    torch.cuda.empty_cache()


@pytest.mark.parametrize("t,s", [(16, 12), (17, 13)], ids=["n28", "n30"])
def test_cpu_verifier_accepts_gpu_proof(t, s):
    """the CPU restatement's verifier, independent of the library's kernels,
    accepts the GPU proof at n = 28 and at the benchmarked n = 30, and
    rejects a flipped byte"""
    from bitz_amd.device import Commitment
    rng = np.random.default_rng(t * 100 + s)
    rows = rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)
    log(f"prover n={t + s} t={t} s={s}, CPU verifier")
    com = Commitment(ctx(), rows, t, s, 4, 2)
    try:
        claimed = com.evaluation()
        proof = com.prove(claimed)
        t0 = time.perf_counter()
        P.verify_standalone(com.root, t, s, 2, 4, claimed, proof)
        log(f"CPU verifier accepted the n={t + s} proof ({len(proof)} bytes) in {time.perf_counter() - t0:.1f} s")
        bad = bytearray(proof)
        bad[len(proof) // 2] ^= 1
        with pytest.raises(P.VerifyError):
            P.verify_standalone(com.root, t, s, 2, 4, claimed, bytes(bad))
    finally:
        com.close()
