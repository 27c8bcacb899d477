"""GPU parity of the opening-side kernels against the CPU oracle, bit-exact,
through the C ABI: eq tables, integer column folds and images, the GKR round
sums / folds (materialised and leaf layer), inner-sumcheck tables and rounds,
ring-switch folds, Ligerito sumcheck messages/folds, the induced basis and the
proof-of-work scan. Sizes include the 2^20-coefficient workload's shapes
(t=15, s=11, m_p=19)."""
import functools
import random

import numpy as np
import pytest

import oracle as O
from gpu_util import ctx, log, rand_gf, to_dev, to_host

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=1)
def ops():
    from bitz_amd.device import OpenOps
    return OpenOps(ctx())


def rand_ints(k, seed):
    rng = random.Random(seed)
    return [rng.getrandbits(128) for _ in range(k)]


def rand_rows(t, s, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)


def dev_u64(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to("cuda")


@pytest.mark.parametrize("n", [0, 1, 5, 8, 9, 12, 13, 14, 15, 19, 22])
def test_eq_table(n):
    r = rand_ints(n, n)
    r[:1] = [0] if n > 3 else r[:1]
    log(f"eq_table n={n}")
    got = to_host(ops().eq_table(r))
    assert np.array_equal(got, O.eq_table(O.ints_to_gf(r)))


@pytest.mark.parametrize("t,s", [(7, 0), (9, 3), (12, 8), (15, 11)])
def test_fold_columns_and_images(t, s):
    rows = rand_rows(t, s, t + s)
    w = rand_gf(1 << t, 5)
    w[::7] = [2**64 - 1, 2**64 - 1]  # force carries and wrap-around
    log(f"fold_columns t={t} s={s}")
    got = to_host(ops().fold_columns(dev_u64(rows), t, s, to_dev(w)))
    want = O.fold_columns(rows, t, s, w)
    assert np.array_equal(got, want)
    g = 0x1F2E3D4C5B6A7988_0123456789ABCDEF
    k = min(1 << s, 512)
    imgs = to_host(ops().pow_images(g, to_dev(want[:k])))
    assert np.array_equal(imgs, O.pow_many(g, want[:k]))


@pytest.mark.parametrize("t,s", [(7, 0), (8, 6), (9, 7), (12, 9), (13, 5), (17, 3), (10, 11), (8, 13)])
def test_pack_and_sumcheck_tables(t, s):
    rows = rand_rows(t, s, 3)
    packed = ops().pack_columns(dev_u64(rows), t, s)
    # pack_columns_from_rows (src/ligerito.rs:655-690): word g * 2^t + b holds
    # bit b of columns 64g..64g+63, column 64g+j at bit j
    cols, groups = 1 << s, ((1 << s) + 63) // 64
    bits = np.unpackbits(rows.reshape(cols, -1).view(np.uint8), axis=1, bitorder="little")
    bits = np.concatenate([bits, np.zeros((groups * 64 - cols, 1 << t), np.uint8)])
    want = np.packbits(bits.reshape(groups, 64, 1 << t).transpose(0, 2, 1), axis=2, bitorder="little")
    want = np.ascontiguousarray(want).view(np.uint64).reshape(-1)
    assert np.array_equal(packed.cpu().numpy().view(np.uint64).reshape(-1), want)
    w2 = rand_gf(1 << s, 1)
    eq = rand_gf(1 << t, 2)
    log(f"xi/fold_rows t={t} s={s}")
    assert np.array_equal(to_host(ops().xi_combined(packed, t, s, to_dev(w2))),
                          O.xi_combined_rows(rows, t, s, w2))
    assert np.array_equal(to_host(ops().fold_rows_eq(dev_u64(rows), t, s, to_dev(eq))),
                          O.fold_rows_eq(rows, t, s, eq))


@pytest.mark.parametrize("n", [2, 64, 1 << 15, 1 << 19])
def test_round_poly_and_fold(n):
    m, w = rand_gf(n, 1), rand_gf(n, 2)
    log(f"round_poly n={n}")
    a0, a1, q = ops().round_poly(to_dev(m), to_dev(w))
    want = O.round_polynomial(m, w)  # [a0, a0+a1+q, q]
    assert [a0, a0 ^ a1 ^ q, q] == want
    r = rand_ints(1, n)[0]
    assert np.array_equal(to_host(ops().fold_lsb(to_dev(m), r)), O.fold_lsb(m, r))


def _eq_parts(point, s_c):
    """tensor weights over `len(point)` coordinates: eq_c over the low s_c,
    eq_y over the rest (the layout of the GKR dense rounds)."""
    pt = O.ints_to_gf(point)
    return O.eq_table(pt[:s_c]), O.eq_table(pt[s_c:])


@pytest.mark.parametrize("log_h,s_c", [(0, 0), (3, 3), (5, 2), (10, 4), (18, 11)])
@pytest.mark.parametrize("send_one", [False, True])
def test_gkr_round_sums(log_h, s_c, send_one):
    h = 1 << log_h
    l, r = rand_gf(2 * h, log_h), rand_gf(2 * h, log_h + 100)
    point = rand_ints(log_h, 7)
    eq_c, eq_y = _eq_parts(point, s_c)
    log(f"gkr_round_sums h=2^{log_h} s_c={s_c} send_one={send_one}")
    got = ops().gkr_round_sums(to_dev(l), to_dev(r), h, to_dev(eq_c), s_c, to_dev(eq_y), send_one)
    full_eq = O.eq_table(O.ints_to_gf(point))
    assert tuple(got) == O.gkr_round_sums(l, r, full_eq, send_one)
    rho = rand_ints(1, 9)[0]
    v = to_dev(l)
    ops().fold_halves_(v, h, rho)
    assert np.array_equal(to_host(v)[:h], O.fold_halves(l, rho))


@pytest.mark.parametrize("log_h,s_c", [(0, 0), (3, 3), (5, 2), (10, 4), (17, 11), (20, 13)])
@pytest.mark.parametrize("send_one", [False, True])
def test_gkr_fold_round_sums(log_h, s_c, send_one):
    h = 1 << log_h
    l, r = rand_gf(4 * h, log_h + 3), rand_gf(4 * h, log_h + 200)
    point = rand_ints(log_h, 17)
    eq_c, eq_y = _eq_parts(point, s_c)
    rho = rand_ints(1, 23)[0]
    log(f"gkr_fold_round_sums h=2^{log_h} s_c={s_c} send_one={send_one}")
    dl, dr = to_dev(l), to_dev(r)
    got = ops().gkr_fold_round_sums_(dl, dr, h, rho, to_dev(eq_c), s_c, to_dev(eq_y), send_one)
    fl, fr = O.fold_halves(l, rho), O.fold_halves(r, rho)
    assert np.array_equal(to_host(dl)[:2 * h], fl)
    assert np.array_equal(to_host(dr)[:2 * h], fr)
    full_eq = O.eq_table(O.ints_to_gf(point))
    assert tuple(got) == O.gkr_round_sums(fl, fr, full_eq, send_one)


@pytest.mark.parametrize("t,s", [(7, 0), (7, 3), (9, 6), (10, 8), (15, 11), (13, 14)])
def test_gkr_leaf_layer(t, s):
    rows = rand_rows(t, s, 11)
    A = rand_gf(1 << t, 12)
    A[::7] = [1, 0]
    packed = ops().pack_columns(dev_u64(rows), t, s)
    dA = to_dev(A)
    log(f"gkr leaf layer t={t} s={s}")
    if t + s <= 20:
        leaves = O.forest_leaves(rows, t, s, A)
        half = 1 << (t + s - 1)
        l, r = leaves[:half], leaves[half:]
        # level 1 = products of the two halves
        assert np.array_equal(to_host(ops().gkr_level1(packed, dA, t, s)), O.product_level(leaves))
    point = rand_ints(t + s - 2, 3)
    s_c = min(s, t + s - 2)
    eq_c, eq_y = _eq_parts(point, s_c)
    for send_one in (False, True):
        got = ops().gkr_leaf_round_sums(packed, dA, t, s, to_dev(eq_c), s_c, to_dev(eq_y), send_one)
        if t + s <= 20:
            full_eq = O.eq_table(O.ints_to_gf(point))
            assert tuple(got) == O.gkr_round_sums(l, r, full_eq, send_one)
    rho = rand_ints(1, 4)[0]
    gl, gr = ops().gkr_leaf_fold(packed, dA, t, s, rho)
    if t + s <= 20:
        assert np.array_equal(to_host(gl), O.fold_halves(l, rho))
        assert np.array_equal(to_host(gr), O.fold_halves(r, rho))
    else:
        # full-size: spot-check rows of the fold against leaves rebuilt on the host
        q = 1 << (t - 2)
        glh, grh = to_host(gl), to_host(gr)
        for y in (0, 1, q - 1):
            for c in (0, 5, (1 << s) - 1):
                def leaf(row):
                    bit = (int(rows[c * (1 << (t - 6)) + row // 64]) >> (row % 64)) & 1
                    return O.gf_to_ints(A[row:row + 1])[0] if bit else 1
                e = leaf(y) ^ O.mul(rho, leaf(y) ^ leaf(y + q))
                o = leaf(y + 2 * q) ^ O.mul(rho, leaf(y + 2 * q) ^ leaf(y + 3 * q))
                x = (y << s) | c
                assert O.gf_to_ints(glh[x:x + 1])[0] == e
                assert O.gf_to_ints(grh[x:x + 1])[0] == o


@pytest.mark.parametrize("t,s", [(7, 0), (7, 3), (9, 6), (10, 8), (15, 11), (13, 14)])
def test_gkr_leaf_round1(t, s):
    """round 1 of the leaf layer and the double fold, from bit patterns: against
    the oracle's dense fold + round sums on small shapes, and against the
    device's pinned dense kernels (round-0 fold, fold_halves, round sums) at
    the full-size shape"""
    rows = rand_rows(t, s, 31)
    A = rand_gf(1 << t, 32)
    A[::5] = [1, 0]
    packed = ops().pack_columns(dev_u64(rows), t, s)
    dA = to_dev(A)
    rho0, rho1 = rand_ints(2, 33)
    point = rand_ints(t + s - 3, 34)
    eq_c, eq_y = _eq_parts(point, s)
    log(f"gkr leaf round 1 t={t} s={s}")
    h1 = 1 << (t + s - 3)
    if t + s <= 20:
        leaves = O.forest_leaves(rows, t, s, A)
        half = 1 << (t + s - 1)
        l1, r1 = O.fold_halves(leaves[:half], rho0), O.fold_halves(leaves[half:], rho0)
        full_eq = O.eq_table(O.ints_to_gf(point))
        for send_one in (False, True):
            got = ops().gkr_leaf_round1_sums(packed, dA, t, s, rho0, to_dev(eq_c), to_dev(eq_y), send_one)
            assert tuple(got) == O.gkr_round_sums(l1, r1, full_eq, send_one)
        gl, gr = ops().gkr_leaf_fold2(packed, dA, t, s, rho0, rho1)
        assert np.array_equal(to_host(gl), O.fold_halves(l1, rho1))
        assert np.array_equal(to_host(gr), O.fold_halves(r1, rho1))
    else:
        dl, dr = ops().gkr_leaf_fold(packed, dA, t, s, rho0)
        for send_one in (False, True):
            want = ops().gkr_round_sums(dl, dr, h1, to_dev(eq_c), s, to_dev(eq_y), send_one)
            got = ops().gkr_leaf_round1_sums(packed, dA, t, s, rho0, to_dev(eq_c), to_dev(eq_y), send_one)
            assert tuple(got) == tuple(want)
        gl, gr = ops().gkr_leaf_fold2(packed, dA, t, s, rho0, rho1)
        ops().fold_halves_(dl, h1, rho1)
        ops().fold_halves_(dr, h1, rho1)
        assert np.array_equal(to_host(gl), to_host(dl)[:h1])
        assert np.array_equal(to_host(gr), to_host(dr)[:h1])


@pytest.mark.parametrize("log_h,s", [(1, 1), (4, 2), (10, 6), (18, 11)])
@pytest.mark.parametrize("send_one", [False, True])
def test_gkr_carried_column_weights(log_h, s, send_one):
    """premul: r <- eq_c * r in place and the same sums as the plain round;
    rowonly over the weighted halves gives them again; after a row fold the
    weighted round still equals the plain round over the folded halves"""
    h = 1 << log_h
    l, r = rand_gf(2 * h, log_h + 40), rand_gf(2 * h, log_h + 41)
    point = rand_ints(log_h, 42)
    eq_c, eq_y = _eq_parts(point, s)
    full_eq = O.eq_table(O.ints_to_gf(point))
    want = O.gkr_round_sums(l, r, full_eq, send_one)
    log(f"carried column weights h=2^{log_h} s={s} send_one={send_one}")
    dl, dr = to_dev(l), to_dev(r)
    assert tuple(ops().gkr_premul_round_sums_(dl, dr, h, to_dev(eq_c), s, to_dev(eq_y), send_one)) == want
    weights = np.tile(eq_c, (2 * h >> s, 1))
    assert np.array_equal(to_host(dr), O.mul_arrays(r, weights))
    assert tuple(ops().gkr_round_sums_rowonly(dl, dr, h, s, to_dev(eq_y), send_one)) == want
    if log_h > s:
        rho = rand_ints(1, 43)[0]
        fl, fr = O.fold_halves(l, rho), O.fold_halves(r, rho)
        ops().fold_halves_(dl, h, rho)
        ops().fold_halves_(dr, h, rho)
        sub = point[:log_h - 1]
        c2, y2 = _eq_parts(sub, s)
        want2 = O.gkr_round_sums(fl, fr, O.eq_table(O.ints_to_gf(sub)), send_one)
        assert tuple(ops().gkr_round_sums_rowonly(dl, dr, h // 2, s, to_dev(y2), send_one)) == want2


@pytest.mark.parametrize("n", [0, 1, 5, 13])
def test_tensor_table(n):
    a, b = rand_ints(n, 50), rand_ints(n, 51)
    got = O.gf_to_ints(to_host(ops().tensor_table(to_dev(O.ints_to_gf(a or [0])), to_dev(O.ints_to_gf(b or [0])), n)))
    for x in ([0, 1, (1 << n) - 1, (1 << n) // 3] if n else [0]):
        want = 1
        for i in range(n):
            want = O.mul(want, b[i] if (x >> i) & 1 else a[i])
        assert got[x] == want


@pytest.mark.parametrize("t,s", [(7, 3), (10, 8)])
def test_gkr_leaf_fold2_carrying_weights(t, s):
    rows = rand_rows(t, s, 61)
    A = rand_gf(1 << t, 62)
    packed = ops().pack_columns(dev_u64(rows), t, s)
    dA = to_dev(A)
    rho0, rho1 = rand_ints(2, 63)
    eq_c = O.eq_table(O.ints_to_gf(rand_ints(s, 64)))
    pl, pr = ops().gkr_leaf_fold2(packed, dA, t, s, rho0, rho1)
    wl, wr = ops().gkr_leaf_fold2(packed, dA, t, s, rho0, rho1, to_dev(eq_c))
    assert np.array_equal(to_host(wl), to_host(pl))
    assert np.array_equal(to_host(wr), O.mul_arrays(to_host(pr), np.tile(eq_c, (1 << (t - 3), 1))))


@pytest.mark.parametrize("t,s", [(7, 0), (7, 3), (9, 6), (10, 8), (15, 11), (13, 14)])
def test_gkr_level1_layer_from_bits(t, s):
    """level 2 from the bits, and the layer over the implicit level 1 (round-0
    sums and fold, with and without carried column weights): against the
    oracle's product tree on small shapes, against the device's pinned dense
    kernels (level 1, product_level, round sums, fold_halves) at full size"""
    rows = rand_rows(t, s, 71)
    A = rand_gf(1 << t, 72)
    A[::3] = [1, 0]
    packed = ops().pack_columns(dev_u64(rows), t, s)
    dA = to_dev(A)
    rho = rand_ints(1, 73)[0]
    point = rand_ints(t + s - 3, 74)
    eq_c, eq_y = _eq_parts(point, s)
    log(f"gkr level-1 layer from bits t={t} s={s}")
    h = 1 << (t + s - 3)
    if t + s <= 20:
        lvl1 = O.product_level(O.forest_leaves(rows, t, s, A))
        lvl2 = O.product_level(lvl1)
        l, r = lvl1[:2 * h], lvl1[2 * h:]
        full_eq = O.eq_table(O.ints_to_gf(point))
        want_sums = {so: O.gkr_round_sums(l, r, full_eq, so) for so in (False, True)}
        fl, fr = O.fold_halves(l, rho), O.fold_halves(r, rho)
    else:
        d1 = ops().gkr_level1(packed, dA, t, s)
        lvl2 = to_host(ops().product_level(d1))
        dl, dr = d1[:2 * h], d1[2 * h:]
        want_sums = {so: tuple(ops().gkr_round_sums(dl, dr, h, to_dev(eq_c), s, to_dev(eq_y), so))
                     for so in (False, True)}
        ops().fold_halves_(dl, h, rho)
        ops().fold_halves_(dr, h, rho)
        fl, fr = to_host(dl)[:h], to_host(dr)[:h]
    assert np.array_equal(to_host(ops().gkr_level2(packed, dA, t, s)), lvl2)
    lvl3 = O.product_level(lvl2) if t + s <= 20 else to_host(ops().product_level(to_dev(lvl2)))
    assert np.array_equal(to_host(ops().gkr_level3(packed, dA, t, s)), lvl3)
    for so in (False, True):
        got = ops().gkr_level1_round0_sums(packed, dA, t, s, to_dev(eq_c), to_dev(eq_y), so)
        assert tuple(got) == tuple(want_sums[so])
    gl, gr = ops().gkr_level1_fold(packed, dA, t, s, rho)
    assert np.array_equal(to_host(gl), fl)
    assert np.array_equal(to_host(gr), fr)
    wl, wr = ops().gkr_level1_fold(packed, dA, t, s, rho, to_dev(eq_c))
    assert np.array_equal(to_host(wl), fl)
    assert np.array_equal(to_host(wr), O.mul_arrays(fr, np.tile(eq_c, (1 << (t - 3), 1))))


def test_product_level():
    v = rand_gf(1 << 16, 3)
    assert np.array_equal(to_host(ops().product_level(to_dev(v))), O.product_level(v))


@pytest.mark.parametrize("n", [1, 16, 1 << 12, 1 << 19])
def test_ring_switch_folds(n):
    p, suffix = rand_gf(n, 1), rand_gf(n, 2)
    w = rand_gf(128, 3)
    log(f"ring switch n={n}")
    assert np.array_equal(to_host(ops().fold_1b_rows(to_dev(p), to_dev(suffix))), O.fold_1b_rows(p, suffix))
    assert np.array_equal(to_host(ops().fold_b128(to_dev(suffix), to_dev(w))), O.fold_b128_elems(suffix, w))
    eta = rand_ints(1, n)[0]
    acc = rand_gf(n, 4)
    got = to_host(ops().axpy_(to_dev(acc), to_dev(suffix), eta))
    assert np.array_equal(got, O.lig_glue(acc, suffix, eta))


@pytest.mark.parametrize("n", [2, 256, 1 << 19])
def test_ligerito_sumcheck(n):
    f, b = rand_gf(n, 5), rand_gf(n, 6)
    log(f"ligerito sumcheck n={n}")
    u0, u2 = ops().lig_msg(to_dev(f), to_dev(b))
    assert (u0, u2) == O.lig_round_msg(f, b)
    u0e, u2e, y = ops().lig_msg(to_dev(f), to_dev(b), with_eval=True)
    assert (u0e, u2e) == (u0, u2)
    ints_f, ints_b = O.gf_to_ints(f[:64]), O.gf_to_ints(b[:64])
    if n <= 256:
        yy = 0
        for a, c in zip(O.gf_to_ints(f), O.gf_to_ints(b)):
            yy ^= O.mul(a, c)
        assert y == yy
    r = rand_ints(1, n + 1)[0]
    fo, bo = ops().lig_fold(to_dev(f), to_dev(b), r)
    assert np.array_equal(to_host(fo), O.fold_lsb(f, r))
    assert np.array_equal(to_host(bo), O.fold_lsb(b, r))
    del ints_f, ints_b


@pytest.mark.parametrize("log_n,nq", [(0, 3), (4, 5), (9, 37), (13, 148)])
def test_induce_basis(log_n, nq):
    rng = random.Random(log_n)
    queries = sorted(rng.sample(range(1 << (log_n + 2)), nq))
    aw = rand_gf(nq, 8)
    log(f"induce basis log_n={log_n} nq={nq}")
    got = to_host(ops().induce_basis(log_n, queries, to_dev(aw)))
    assert np.array_equal(got, O.induce_basis(log_n, queries, aw))


@pytest.mark.parametrize("bits", [0, 1, 8, 14])
def test_pow_find(bits):
    prefix = b"bitz-pcs-pow-v1\0"
    seed = bytes(range(16, 32))
    log(f"pow bits={bits}")
    got = ops().pow_find(prefix, seed, bits)
    assert got == O.pow_find(prefix, seed, bits)
    assert O.pow_valid(prefix, seed, got, bits)
