"""The low GKR levels off 16-bit row words (bitz_gkr_nibble_*): for the
leaf, level-1 and level-2 layers, after their first 2 - l rounds folded by
random challenges, round 2 - l's sums and the fold into halves of 2^(t+s-4)
entries, bit-exact against the oracle's dense GKR (forest_leaves,
product_level, fold_halves, gkr_round_sums) at small shapes, and against the
device's dense kernels (pinned against the oracle in test_gpu_open.py) at the
2^20-coefficient shape (t = 15, s = 11)."""
import ctypes as C
import random

import numpy as np
import pytest

import oracle as O
from gpu_util import ctx, log, rand_gf, to_dev, to_host

pytestmark = pytest.mark.gpu


def _rand_ints(k, seed):
    rng = random.Random(seed)
    return [rng.getrandbits(128) for _ in range(k)]


def _b16(vals):
    return (C.c_uint8 * max(16 * len(vals), 16)).from_buffer_copy(
        b"".join(int(v).to_bytes(16, "little") for v in vals) or bytes(16))


def _vp(x):
    return C.c_void_p(x.data_ptr()) if hasattr(x, "data_ptr") else C.cast(x, C.c_void_p)


def _nibble_rows(packed, t, s):
    import torch
    from bitz_amd import _lib
    nr = torch.empty((1 << (t - 4)) << s, dtype=torch.int16, device="cuda")
    _lib.call("bitz_gkr_nibble_rows_dev", ctx().ptr, _vp(packed), t, s, _vp(nr))
    return nr


def _round_sums(level, nr, dA, t, s, chal, eq_c, eq_y, send_one):
    from bitz_amd import _lib
    out = (C.c_uint8 * 32)()
    _lib.call("bitz_gkr_nibble_round_sums_dev", ctx().ptr, level, _vp(nr), _vp(dA), t, s, _b16(chal), _vp(eq_c),
              _vp(eq_y), 1 if send_one else 0, C.cast(out, C.c_void_p))
    b = bytes(out)
    return int.from_bytes(b[:16], "little"), int.from_bytes(b[16:], "little")


def _fold(level, nr, dA, t, s, chal, eq_hi=None):
    from bitz_amd import _lib
    n = 1 << (t + s - 4)
    l, r = ctx().empty_gf(n), ctx().empty_gf(n)
    _lib.call("bitz_gkr_nibble_fold_dev", ctx().ptr, level, _vp(nr), _vp(dA), t, s, _b16(chal),
              _vp(eq_hi) if eq_hi is not None else None, _vp(l), _vp(r))
    return l, r


def _rows(t, s, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)


def _setup(t, s, seed):
    import torch
    from bitz_amd.device import OpenOps
    rows = _rows(t, s, seed)
    A = rand_gf(1 << t, seed + 1)
    A[::5] = [1, 0]
    ops = OpenOps(ctx())
    packed = ops.pack_columns(torch.from_numpy(rows.view(np.int64)).to("cuda"), t, s)
    dA = to_dev(A)
    return rows, A, ops, packed, dA, _nibble_rows(packed, t, s)


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("t,s", [(7, 1), (8, 3), (9, 6), (10, 8), (11, 9), (12, 5)])
def test_nibble_layers_against_oracle(level, t, s):
    rows, A, ops, packed, dA, nr = _setup(t, s, 100 * t + s)
    log(f"nibble layer level={level} t={t} s={s}")
    lvl = O.forest_leaves(rows, t, s, A)
    for _ in range(level):
        lvl = O.product_level(lvl)
    half = lvl.shape[0] // 2
    l, r = lvl[:half].copy(), lvl[half:].copy()
    k = 2 - level
    chal = _rand_ints(k + 1, 7 * t + level)
    h = half
    for i in range(k):
        h //= 2
        l, r = O.fold_halves(l, chal[i]), O.fold_halves(r, chal[i])
    point = _rand_ints(t + s - 4, 11 * t + s + level)
    pt = O.ints_to_gf(point)
    eq_c, eq_y = O.eq_table(pt[:s]), O.eq_table(pt[s:])
    full_eq = O.eq_table(pt)
    for so in (False, True):
        got = _round_sums(level, nr, dA, t, s, chal[:k], to_dev(eq_c), to_dev(eq_y), so)
        assert got == O.gkr_round_sums(l, r, full_eq, so)
    gl, gr = _fold(level, nr, dA, t, s, chal)
    fl, fr = O.fold_halves(l, chal[k]), O.fold_halves(r, chal[k])
    assert np.array_equal(to_host(gl), fl)
    assert np.array_equal(to_host(gr), fr)
    if s > 8:
        eq_hi = O.eq_table(O.ints_to_gf(_rand_ints(s - 8, 3 * t)))
        wl, wr = _fold(level, nr, dA, t, s, chal, to_dev(eq_hi))
        assert np.array_equal(to_host(wl), fl)
        cols_hi = np.repeat(eq_hi, 256, axis=0)
        assert np.array_equal(to_host(wr), O.mul_arrays(fr, np.tile(cols_hi, (fr.shape[0] >> s, 1))))


@pytest.mark.parametrize("level", [0, 1, 2])
def test_nibble_layers_at_2p20_coefficients(level):
    """t = 15, s = 11: against the device's dense kernels over the dense
    halves (level 1 and 2 from gkr_level1 / product_level, leaves from the
    level-1 kernel's inputs rebuilt with torch)"""
    import torch
    t, s = 15, 11
    rows, A, ops, packed, dA, nr = _setup(t, s, 4242)
    log(f"nibble layer level={level} at t={t} s={s}")
    rows_d = torch.from_numpy(rows.view(np.int64)).to("cuda")
    shifts = torch.arange(64, device="cuda", dtype=torch.int64)
    bits = ((rows_d.view(1 << s, 1 << (t - 6), 1) >> shifts) & 1).to(torch.bool).view(1 << s, 1 << t).t()
    one = torch.tensor([1, 0], dtype=torch.int64, device="cuda")
    lvl = torch.where(bits.unsqueeze(-1), dA.view(1 << t, 1, 2), one.view(1, 1, 2)).reshape(-1, 2).contiguous()
    for _ in range(level):
        lvl = ops.product_level(lvl)
    half = lvl.shape[0] // 2
    l, r = lvl[:half].clone(), lvl[half:].clone()
    k = 2 - level
    chal = _rand_ints(k + 1, 99 + level)
    h = half
    for i in range(k):
        h //= 2
        ops.fold_halves_(l, h, chal[i])
        ops.fold_halves_(r, h, chal[i])
    point = _rand_ints(t + s - 4, 5 + level)
    pt = O.ints_to_gf(point)
    eq_c, eq_y = O.eq_table(pt[:s]), O.eq_table(pt[s:])
    hh = h // 2
    for so in (False, True):
        want = ops.gkr_round_sums(l[:h].contiguous(), r[:h].contiguous(), hh, to_dev(eq_c), s, to_dev(eq_y), so)
        got = _round_sums(level, nr, dA, t, s, chal[:k], to_dev(eq_c), to_dev(eq_y), so)
        assert got == tuple(want)
    gl, gr = _fold(level, nr, dA, t, s, chal)
    ops.fold_halves_(l, hh, chal[k])
    ops.fold_halves_(r, hh, chal[k])
    assert torch.equal(gl, l[:hh]) and torch.equal(gr, r[:hh])


def _jit_round_sums(level, nr, dA, t, s, chal, eq_c, eq_y, send_one):
    from bitz_amd import _lib
    out = (C.c_uint8 * 32)()
    _lib.call("bitz_gkr_jit_round_sums_dev", ctx().ptr, level, _vp(nr), _vp(dA), t, s, _b16(chal), _vp(eq_c),
              _vp(eq_y), 1 if send_one else 0, C.cast(out, C.c_void_p))
    b = bytes(out)
    return int.from_bytes(b[:16], "little"), int.from_bytes(b[16:], "little")


def _jit_fold(level, nr, dA, t, s, chal, eq_hi=None):
    from bitz_amd import _lib
    n = 1 << (t + s - 5)
    l, r = ctx().empty_gf(n), ctx().empty_gf(n)
    _lib.call("bitz_gkr_jit_fold_dev", ctx().ptr, level, _vp(nr), _vp(dA), t, s, _b16(chal),
              _vp(eq_hi) if eq_hi is not None else None, _vp(l), _vp(r))
    return l, r


@pytest.mark.parametrize("level", [0, 1, 2, 3])
@pytest.mark.parametrize("t,s", [(7, 1), (8, 3), (9, 6), (10, 8), (11, 9), (12, 5)])
def test_table_round_against_oracle(level, t, s):
    """Round 3 - l of the level-l layer from its per-row tables (the sums and
    the fold into halves of 2^(t+s-5) entries) against the oracle's dense
    layer folded by the same challenges"""
    rows, A, ops, packed, dA, nr = _setup(t, s, 100 * t + s + 7)
    log(f"table round level={level} t={t} s={s}")
    lvl = O.forest_leaves(rows, t, s, A)
    for _ in range(level):
        lvl = O.product_level(lvl)
    half = lvl.shape[0] // 2
    l, r = lvl[:half].copy(), lvl[half:].copy()
    j = 3 - level
    chal = _rand_ints(j + 1, 13 * t + level)
    for i in range(j):
        l, r = O.fold_halves(l, chal[i]), O.fold_halves(r, chal[i])
    assert l.shape[0] == 1 << (t + s - 4)
    point = _rand_ints(t + s - 5, 17 * t + s + level)
    pt = O.ints_to_gf(point)
    eq_c, eq_y = O.eq_table(pt[:s]), O.eq_table(pt[s:])
    full_eq = O.eq_table(pt)
    for so in (False, True):
        got = _jit_round_sums(level, nr, dA, t, s, chal[:j], to_dev(eq_c), to_dev(eq_y), so)
        assert got == O.gkr_round_sums(l, r, full_eq, so)
    gl, gr = _jit_fold(level, nr, dA, t, s, chal)
    fl, fr = O.fold_halves(l, chal[j]), O.fold_halves(r, chal[j])
    assert np.array_equal(to_host(gl), fl)
    assert np.array_equal(to_host(gr), fr)
    if s > 8:
        eq_hi = O.eq_table(O.ints_to_gf(_rand_ints(s - 8, 5 * t + level)))
        wl, wr = _jit_fold(level, nr, dA, t, s, chal, to_dev(eq_hi))
        assert np.array_equal(to_host(wl), fl)
        cols_hi = np.repeat(eq_hi, 256, axis=0)
        assert np.array_equal(to_host(wr), O.mul_arrays(fr, np.tile(cols_hi, (fr.shape[0] >> s, 1))))


@pytest.mark.parametrize("t,s", [(7, 1), (9, 6), (11, 9)])
def test_level4_from_row_words_against_oracle(t, s):
    from bitz_amd import _lib
    rows, A, ops, packed, dA, nr = _setup(t, s, 31 * t + s)
    log(f"level 4 from row words t={t} s={s}")
    lvl = O.forest_leaves(rows, t, s, A)
    for _ in range(4):
        lvl = O.product_level(lvl)
    out = ctx().empty_gf(1 << (t + s - 4))
    _lib.call("bitz_gkr_level4_dev", ctx().ptr, _vp(nr), _vp(dA), t, s, _vp(out))
    assert np.array_equal(to_host(out), lvl)


def _dense_level(ops, rows, dA, t, s, level):
    import torch
    rows_d = torch.from_numpy(rows.view(np.int64)).to("cuda")
    shifts = torch.arange(64, device="cuda", dtype=torch.int64)
    bits = ((rows_d.view(1 << s, 1 << (t - 6), 1) >> shifts) & 1).to(torch.bool).view(1 << s, 1 << t).t()
    one = torch.tensor([1, 0], dtype=torch.int64, device="cuda")
    lvl = torch.where(bits.unsqueeze(-1), dA.view(1 << t, 1, 2), one.view(1, 1, 2)).reshape(-1, 2).contiguous()
    del bits
    for _ in range(level):
        lvl = ops.product_level(lvl)
    return lvl


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_table_round_at_2p20_coefficients(level):
    """t = 15, s = 11: the table round's sums and folds (plain and with the
    eq_hi-weighted right half) against the device's dense kernels, and level
    4 from the row words against three product levels over level 1"""
    import torch
    t, s = 15, 11
    rows, A, ops, packed, dA, nr = _setup(t, s, 5151 + level)
    log(f"table round level={level} at t={t} s={s}")
    lvl = _dense_level(ops, rows, dA, t, s, level)
    half = lvl.shape[0] // 2
    l, r = lvl[:half].clone(), lvl[half:].clone()
    del lvl
    j = 3 - level
    chal = _rand_ints(j + 1, 77 + level)
    h = half
    for i in range(j):
        h //= 2
        ops.fold_halves_(l, h, chal[i])
        ops.fold_halves_(r, h, chal[i])
    assert h == 1 << (t + s - 4)
    point = _rand_ints(t + s - 5, 9 + level)
    pt = O.ints_to_gf(point)
    eq_c, eq_y = O.eq_table(pt[:s]), O.eq_table(pt[s:])
    hh = h // 2
    for so in (False, True):
        want = ops.gkr_round_sums(l[:h].contiguous(), r[:h].contiguous(), hh, to_dev(eq_c), s, to_dev(eq_y), so)
        got = _jit_round_sums(level, nr, dA, t, s, chal[:j], to_dev(eq_c), to_dev(eq_y), so)
        assert got == tuple(want)
    gl, gr = _jit_fold(level, nr, dA, t, s, chal)
    eq_hi = O.eq_table(O.ints_to_gf(_rand_ints(s - 8, 3 + level)))
    wl, wr = _jit_fold(level, nr, dA, t, s, chal, to_dev(eq_hi))
    ops.fold_halves_(l, hh, chal[j])
    ops.fold_halves_(r, hh, chal[j])
    assert torch.equal(gl, l[:hh]) and torch.equal(gr, r[:hh])
    assert torch.equal(wl, l[:hh])
    cols_hi = torch.from_numpy(np.repeat(eq_hi, 256, axis=0).view(np.int64)).to("cuda")
    want_r = ctx().gf_mul(r[:hh].contiguous(), cols_hi.repeat(hh >> s, 1).contiguous())
    assert torch.equal(wr, want_r)
    if level == 0:
        from bitz_amd import _lib
        lv4 = _dense_level(ops, rows, dA, t, s, 4)
        out = ctx().empty_gf(1 << (t + s - 4))
        _lib.call("bitz_gkr_level4_dev", ctx().ptr, _vp(nr), _vp(dA), t, s, _vp(out))
        assert torch.equal(out, lv4)
