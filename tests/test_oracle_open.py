"""Pins the oracle's opening-side functions without the reference binary:
closed forms (eq tables, MLE evaluation, integer folds by bit walk, images by
square-and-multiply), the protocols' own verifier equations (a dense GKR
proved with the oracle's round sums and folds is accepted by a restatement of
gpgkr_verify, src/bitz/gkr.rs:337-383; the inner-sumcheck and Ligerito round
messages satisfy their consistency checks), brute-force MLE evaluation for
the ring-switch claims, the pinned NTT for the induced basis, and exhaustive
search for the proof of work."""
import random

import numpy as np
import pytest

import oracle as O

rng0 = random.Random(2024)


def R(k, rng=rng0):
    return [rng.getrandbits(128) for _ in range(k)]


def gfa(xs):
    return O.ints_to_gf(xs)


def test_eq_table_closed_form():
    r = R(5)
    eq = O.gf_to_ints(O.eq_table(gfa(r)))
    for x in range(32):
        p = 1
        for i in range(5):
            p = O.mul(p, r[i] if (x >> i) & 1 else 1 ^ r[i])
        assert eq[x] == p
    acc = 0
    for v in eq:
        acc ^= v
    assert acc == 1  # sum_x eq(x, r) = 1
    assert O.gf_to_ints(O.eq_table(gfa([])))[0] == 1


def test_mle_eval_is_eq_inner_product():
    v = R(64)
    pt = R(6)
    eq = O.gf_to_ints(O.eq_table(gfa(pt)))
    acc = 0
    for a, b in zip(v, eq):
        acc ^= O.mul(a, b)
    assert O.mle_eval(gfa(v), gfa(pt)) == acc
    for x in (0, 13, 63):  # boolean points pick entries
        bp = [(x >> i) & 1 for i in range(6)]
        assert O.mle_eval(gfa(v), gfa(bp)) == v[x]


def test_fold_columns_bit_walk_and_images():
    t, s = 8, 3
    rng = np.random.default_rng(1)
    rows = rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)
    w = [random.Random(i).getrandbits(128) for i in range(1 << t)]
    w[3] = 2**128 - 1
    got = O.gf_to_ints(O.fold_columns(rows, t, s, gfa(w)))
    words = 1 << (t - 6)
    for c in range(1 << s):
        total = 0
        for b in range(1 << t):
            if (int(rows[c * words + b // 64]) >> (b % 64)) & 1:
                total += w[b]
        assert got[c] == total % (1 << 128)
    g = R(1)[0]
    imgs = O.gf_to_ints(O.pow_many(g, gfa(got[:4])))
    for e, im in zip(got[:4], imgs):
        acc, base = 1, g
        for i in range(128):  # square-and-multiply from the LSB
            if (e >> i) & 1:
                acc = O.mul(acc, base)
            base = O.mul(base, base)
        assert im == acc


# ---- dense GKR: prover from the oracle's primitives, verifier restated ----
def eq_factor(r, z):
    return O.mul(r, z) ^ O.mul(1 ^ r, 1 ^ z)


def gkr_prove(point, layers, chal):
    """gpgkr_prove (gkr.rs:110-125, 200-323) in plain dense form."""
    internal = list(reversed(point))
    msgs = []
    for w in reversed(layers):
        L = len(internal)
        assert w.shape[0] == 2 << L
        l, r = w[: 1 << L].copy(), w[1 << L:].copy()
        factor, nxt = 1, []
        for rnd, z in enumerate(internal):
            rem = L - 1 - rnd
            h = 1 << rem
            ext = list(reversed(internal))
            eq = O.eq_table(gfa(ext[:rem]))
            se, si = O.gkr_round_sums(l, r, eq, z == 0)
            msg = [O.mul(factor, se), O.mul(factor, si)]
            msgs.append(msg)
            rr = chal(msg)
            nxt.append(rr)
            l, r = O.fold_halves(l, rr), O.fold_halves(r, rr)
            factor = O.mul(factor, eq_factor(rr, z))
        last = [O.gf_to_ints(l)[0], O.gf_to_ints(r)[0]]
        msgs.append(last)
        rc = chal(last)
        internal = [rc] + nxt
    return list(reversed(internal)), msgs


def gkr_verify(claim, point, rounds, msgs, chal):
    """gpgkr_verify / verify_layer (gkr.rs:337-383)."""
    internal = list(reversed(point))
    it = iter(msgs)
    for _ in range(rounds):
        prefix, nxt = 1, []
        for z in internal:
            se, si = next(it)
            if z == 0:
                s0, s1 = claim, se
            else:
                s0, s1 = se, O.mul(claim ^ O.mul(1 ^ z, se), O.inv(z))
            rr = chal([se, si])
            nxt.append(rr)
            f = eq_factor(rr, z)
            bracket = (s1 ^ s0) ^ O.mul(rr ^ 1, si)
            claim = O.mul(f, s0 ^ O.mul(rr, bracket))
            prefix = O.mul(prefix, f)
        el, er = next(it)
        if O.mul(prefix, O.mul(el, er)) != claim:
            return None
        rc = chal([el, er])
        internal = [rc] + nxt
        claim = el ^ O.mul(rc, er ^ el)
    return list(reversed(internal)), claim


def _chal_stream(seed):
    state = [seed]

    def chal(msg):
        data = b"".join(m.to_bytes(16, "little") for m in msg) + state[0].to_bytes(32, "little")
        d = O.blake3(data)
        state[0] = int.from_bytes(d, "little")
        return int.from_bytes(d[:16], "little")
    return chal


@pytest.mark.parametrize("t,s,zmode", [(3, 2, "rand"), (4, 3, "rand"), (3, 3, "bool"), (2, 0, "rand")])
def test_dense_gkr_round_trip(t, s, zmode):
    rng = random.Random(t * 10 + s)
    leaves = gfa([rng.getrandbits(128) for _ in range(1 << (t + s))])
    layers = [leaves]
    while layers[-1].shape[0] > (1 << s):
        layers.append(O.product_level(layers[-1]))
    roots = layers.pop()
    zeta = [rng.getrandbits(128) for _ in range(s)]
    if zmode == "bool":
        zeta = [i % 2 for i in range(s)]
    claim0 = O.mle_eval(roots, gfa(zeta)) if s else O.gf_to_ints(roots)[0]
    pt, msgs = gkr_prove(zeta, layers, _chal_stream(7))
    res = gkr_verify(claim0, zeta, t, msgs, _chal_stream(7))
    assert res is not None
    vpt, vclaim = res
    assert vpt == pt
    assert O.mle_eval(leaves, gfa(vpt)) == vclaim  # the leaf claim is true
    bad = [list(m) for m in msgs]
    bad[1][0] ^= 1
    assert gkr_verify(claim0, zeta, t, bad, _chal_stream(7)) is None


def test_forest_leaves_and_level_products():
    t, s = 7, 2
    rng = np.random.default_rng(3)
    rows = rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)
    A = R(1 << t)
    leaves = O.gf_to_ints(O.forest_leaves(rows, t, s, gfa(A)))
    words = 1 << (t - 6)
    for row in range(1 << t):
        for c in range(1 << s):
            bit = (int(rows[c * words + row // 64]) >> (row % 64)) & 1
            assert leaves[row * (1 << s) + c] == (A[row] if bit else 1)
    # roots of the product tree are the column images prod_b (bit ? A_b : 1)
    lv = gfa(leaves)
    while lv.shape[0] > (1 << s):
        lv = O.product_level(lv)
    for c in range(1 << s):
        p = 1
        for row in range(1 << t):
            if (int(rows[c * words + row // 64]) >> (row % 64)) & 1:
                p = O.mul(p, A[row])
        assert O.gf_to_ints(lv)[c] == p


def test_inner_sumcheck_rounds_are_consistent():
    t, s = 7, 3
    rng = np.random.default_rng(9)
    rows = rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)
    w1, w2 = gfa(R(1 << t)), gfa(R(1 << s))
    # target = sum_{b,c} w1[b] w2[c] bit(c,b)
    m = O.xi_combined_rows(rows, t, s, w2)
    target = 0
    for a, b in zip(O.gf_to_ints(m), O.gf_to_ints(w1)):
        target ^= O.mul(a, b)
    chal = _chal_stream(3)
    point = []
    comb, wts = m, w1
    while comb.shape[0] > 1:
        a0, a1, a2 = O.round_polynomial(comb, wts)
        assert a1 ^ a2 == target  # g(0) + g(1) in characteristic two
        r = chal([a0, a1, a2])
        target = a0 ^ O.mul(r, a1 ^ O.mul(r, a2))
        point.append(r)
        comb, wts = O.fold_lsb(comb, r), O.fold_lsb(wts, r)
    folded = O.fold_rows_eq(rows, t, s, O.eq_table(gfa(point)))
    row_scalar = O.gf_to_ints(wts)[0]
    cols = gfa([O.mul(row_scalar, x) for x in O.gf_to_ints(w2)])
    while folded.shape[0] > 1:
        a0, a1, a2 = O.round_polynomial(folded, cols)
        assert a1 ^ a2 == target
        r = chal([a0, a1, a2])
        target = a0 ^ O.mul(r, a1 ^ O.mul(r, a2))
        point.append(r)
        folded, cols = O.fold_lsb(folded, r), O.fold_lsb(cols, r)
    ev = O.gf_to_ints(folded)[0]
    assert target == O.mul(ev, O.gf_to_ints(cols)[0])
    # the evaluation is the bits' MLE at the point (rows before columns)
    bits = []
    words = 1 << (t - 6)
    for c in range(1 << s):
        for b in range(1 << t):
            bits.append((int(rows[c * words + b // 64]) >> (b % 64)) & 1)
    assert O.mle_eval(gfa(bits), gfa(point)) == ev


def test_ring_switch_claims_match_brute_force():
    m = 9  # bits = 2^9, packed 2^2 elements of 128 bits
    rng = random.Random(4)
    p = [rng.getrandbits(128) for _ in range(1 << (m - 7))]
    point = R(m)
    prefix = O.eq_table(gfa(point[:7]))
    suffix = O.eq_table(gfa(point[7:]))
    claims = O.fold_1b_rows(gfa(p), suffix)
    # claim_check = sum_i prefix[i] claims[i] = MLE of the bits at the point,
    # bit i of p[x] being logical bit x*128 + i
    cc = 0
    for a, b in zip(O.gf_to_ints(prefix), O.gf_to_ints(claims)):
        cc ^= O.mul(a, b)
    bits = [(p[x] >> i) & 1 for x in range(len(p)) for i in range(128)]
    assert cc == O.mle_eval(gfa(bits), gfa(point))
    # fold_b128 is the transpose of fold_1b_rows: sum_x p[x].bits . basis
    w = R(128)
    basis = O.gf_to_ints(O.fold_b128_elems(suffix, gfa(w)))
    for x in range(len(p)):
        sx = O.gf_to_ints(suffix)[x]
        acc = 0
        for i in range(128):
            if (sx >> i) & 1:
                acc ^= w[i]
        assert basis[x] == acc
    # batch_claims: <transpose(claims), w>
    cl = O.gf_to_ints(claims)
    tr = [sum((((cl[i] >> b) & 1) << i) for i in range(128)) for b in range(128)]
    acc = 0
    for a, b in zip(tr, w):
        acc ^= O.mul(a, b)
    assert O.batch_claims(claims, gfa(w)) == acc


def test_ligerito_round_message_identities():
    n = 64
    f, b = gfa(R(n)), gfa(R(n))
    u0, u2 = O.lig_round_msg(f, b)
    fi, bi = O.gf_to_ints(f), O.gf_to_ints(b)
    tot = 0
    q1 = 0
    for j in range(n // 2):
        tot ^= O.mul(fi[2 * j], bi[2 * j]) ^ O.mul(fi[2 * j + 1], bi[2 * j + 1])
        q1 ^= O.mul(fi[2 * j + 1], bi[2 * j + 1])
    # q(X) = u0 + u1 X + u2 X^2 with q(0) + q(1) = total and q(1) = sum f1 b1
    u1 = tot ^ u2
    assert u0 ^ u1 ^ u2 == q1
    r = R(1)[0]
    # q(r) equals the next total after folding both by r
    nf, nb = O.fold_lsb(f, r), O.fold_lsb(b, r)
    nt = 0
    for a, c in zip(O.gf_to_ints(nf), O.gf_to_ints(nb)):
        nt ^= O.mul(a, c)
    assert nt == u0 ^ O.mul(r, u1) ^ O.mul(O.mul(r, r), u2)


def test_induced_basis_matches_ntt_evaluations():
    log_n, rate = 4, 2
    queries = [0, 3, 17, 40, 63]
    aw = R(len(queries))
    got = O.gf_to_ints(O.induce_basis(log_n, queries, gfa(aw)))
    log_d = log_n + rate
    for j in range(1 << log_n):
        e = np.zeros((1 << log_d, 2), np.uint64)
        e[j] = [1, 0]
        ev = O.gf_to_ints(O.ntt_forward(e, log_d, 0))  # X_j at every domain point
        acc = 0
        for q, a in zip(queries, aw):
            acc ^= O.mul(a, ev[q])
        assert got[j] == acc


def test_pow_is_smallest_valid_nonce():
    prefix = b"bitz-pcs-pow-v1\0"
    seed = bytes(range(16))
    for bits in (0, 1, 4, 9):
        n = O.pow_find(prefix, seed, bits)
        assert O.pow_valid(prefix, seed, n, bits)
        for k in range(n):
            assert not O.pow_valid(prefix, seed, k, bits)
        # validity is the leading-zero count of the published hash
        d = O.blake3(prefix + seed + n.to_bytes(8, "little"))
        lz = 0
        for byte in d:
            z = 8 - byte.bit_length()
            lz += z
            if z != 8:
                break
        assert lz >= bits


def test_openmp_loops_do_not_depend_on_thread_count():
    """the oracle's parallel loops (the CPU baseline runs them on every host
    core) give the single-thread results, at sizes above their parallel
    thresholds"""
    rs = np.random.default_rng(77)
    n = 1 << 15
    t, s = 12, 6
    rows = rs.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)
    a = gfa(R(n))
    b = gfa(R(n))
    pt = gfa(R(14))
    w128 = gfa(R(128))

    def run():
        eq = O.eq_table(pt)
        return [
            eq,
            O.mle_eval(a, gfa(R(15, random.Random(1)))),
            O.gkr_round_sums(a[: n // 2], b[: n // 2], eq[: n // 4], False),
            O.gkr_round_sums(a[: n // 2], b[: n // 2], eq[: n // 4], True),
            O.fold_halves(a, 12345),
            O.fold_lsb(b, 678),
            O.product_level(a),
            O.round_polynomial(a, b),
            O.xi_combined_rows(rows, t, s, gfa(R(1 << s, random.Random(2)))),
            O.fold_rows_eq(rows, t, s, O.eq_table(gfa(R(t, random.Random(3))))),
            O.fold_1b_rows(a, b),
            O.fold_b128_elems(b, w128),
            O.lig_round_msg(a, b),
            O.ntt_forward(a, 11, 4),
            O.merkle_tree(a.tobytes(), 1 << 10),
        ]

    before = O.max_threads()
    try:
        O.set_threads(1)
        one = run()
        O.set_threads(max(before, 4))
        many = run()
    finally:
        O.set_threads(before)
    for x, y in zip(one, many):
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y)
        else:
            assert x == y
