"""The outer (Blake3Transcript) half of a standalone opening, byte for byte.

Pins, each against the reference rather than against a restatement of
itself:

* the custom Johnson ladder solver (oracle and library) against all 42 of
  the reference's embedded Johnson ladders, m = 22..35 at rates 1/2, 1/4,
  1/8 (vendor/flock-mod/crates/flock-core/configs/ligerito/m*_{fast,slim,
  slim3}.toml, reduced to tests/golden/ligerito_ladders.json by
  tests/golden/make_ligerito_fixture.py): level shapes, queries, query and
  fold grinding, OOD samples, the final block, and the level diagnostics to
  the TOMLs' one decimal;
* the bincode of LigeritoSecurityConfig (pcs/ligerito.rs:541-651), written
  here field by field from the struct declaration;
* the statement frame (src/ligerito_flock.rs:1057-1160, 1245-1276), the
  session label and ladder digest (bitz_opener.rs:589-609) and Round 0's
  header (:1488-1496), written here field by field;
* the claim frame (:1278-1284) and the grinding boundary
  (src/piop/spartan/grinding.rs:145-160, 272-275);
* the prime search's policy words against budgets()' closed form
  (vendor/field/src/prime/sampling.rs:33-56), the sampled prime checked
  prime by an independent Baillie-PSW test and in its interval;
* the Round-0 nonce the smallest valid one by exhaustive scan.

The f64 diagnostics inside the digest are computed with the platform libm
(log2, pow, exp2); they are pinned to the TOMLs' rounding, not to the bit.
"""
import ctypes as C
import json
import math
import os
import struct
import subprocess

import pytest

import bitz_amd._lib as L
import bitz_protocol as P
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
LADDERS = json.load(open(os.path.join(HERE, "golden", "ligerito_ladders.json")))


@pytest.fixture(scope="module")
def oh(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("outer") / "libouter.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-pthread", "-o", so,
                    os.path.join(HERE, "outer_harness.cpp")], check=True)
    lib = C.CDLL(so)
    vp, i64 = C.c_void_p, C.c_int64
    lib.oh_ladder_bincode.argtypes = [C.c_int, C.c_int, C.c_int, vp, i64]
    lib.oh_ladder_bincode.restype = i64
    lib.oh_statement_stream.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, i64]
    lib.oh_statement_stream.restype = i64
    lib.oh_sample_prime.argtypes = [vp, C.c_uint64, C.c_int, vp, vp, i64]
    lib.oh_sample_prime.restype = i64
    lib.oh_grinding_seed.argtypes = [vp, C.c_uint64, C.c_char_p, C.c_uint64, C.c_uint32, vp, vp, i64]
    lib.oh_grinding_seed.restype = i64
    lib.oh_claim_stream.argtypes = [vp, vp, vp, i64]
    lib.oh_claim_stream.restype = i64
    lib.oh_sample_residue.argtypes = [vp, C.c_uint64, vp, vp]
    return lib


def _buf(b: bytes):
    return (C.c_uint8 * max(len(b), 1)).from_buffer_copy(b or b"\0")


def _call_stream(fn, *args, cap=1 << 20):
    out = (C.c_uint8 * cap)()
    n = fn(*args, C.cast(out, C.c_void_p), cap)
    assert 0 <= n <= cap
    return bytes(out[:n])


def u64(v):
    return struct.pack("<Q", v)


def u32(v):
    return struct.pack("<I", v)


def f64(v):
    return struct.pack("<d", v)


def slice_(b: bytes) -> bytes:
    """absorb_slice (src/transcript/traits.rs:371-376)."""
    return b"\x06" + u64(len(b)) + b + b"\x07"


def round1(x: float) -> float:
    """round1 (pcs/ligerito.rs:665-667): f64::round is half away from zero."""
    return math.floor(x * 10.0 + 0.5) / 10.0


# ------------------------------------------------------------------ ladders
def _lib_ladder(m, r, k):
    out = (C.c_int32 * 256)()
    n = L.load().bitz_ligerito_config(m, r, k, 100, C.cast(out, C.c_void_p), 256)
    assert n > 0
    v = list(out[:n])
    keys = ["rate", "msg_cols", "lanes", "k", "queries", "grinding", "fold_grinding", "ood"]
    return v[3], [dict(zip(keys, v[6 + 8 * i:14 + 8 * i])) for i in range(v[5])]


@pytest.mark.parametrize("row", LADDERS, ids=lambda r: f"m{r['m']}_{r['profile']}")
def test_ladder_matches_reference_toml(row):
    assert row["analysis_version"] == P.TEMPLATE["analysis_version"] and row["field"] == P.TEMPLATE["field"]
    cfg = P.custom_johnson_config(row["m"], row["log_inv_rate"], row["initial_k"])
    assert (cfg["m"], cfg["log_n"], cfg["initial_k"], cfg["yr_log_n"]) == (
        row["m"], row["log_n"], row["initial_k"], row["yr_log_n"])
    assert len(cfg["levels"]) == len(row["levels"])
    lib_yr, lib_levels = _lib_ladder(row["m"], row["log_inv_rate"], row["initial_k"])
    assert lib_yr == row["yr_log_n"]
    for mine, lib, ref in zip(cfg["levels"], lib_levels, row["levels"]):
        assert ref["regime"] == "johnson_ood" and ref["eta"] == P.TEMPLATE["eta"]
        assert ref["target_security_bits"] == 100
        want = (ref["log_inv_rate"], ref["log_msg_cols"], ref["log_num_interleaved"], ref["k_recursive"],
                ref["queries"], ref["grinding_bits"], ref["fold_grinding_bits"], ref["ood_samples"])
        for got in (mine, lib):
            assert (got["rate"], got["msg_cols"], got["lanes"], got["k"], got["queries"], got["grinding"],
                    got["fold_grinding"], got["ood"]) == want
        assert round1(mine["eps_pg"]) == ref["expected_eps_pg_bits"]
        assert round1(mine["eps_query"]) == ref["expected_eps_query_bits"]
        assert round1(mine["eps_ood"]) == ref["expected_eps_ood_bits"]


def _expected_bincode(row: dict, eps: list) -> bytes:
    """LigeritoSecurityConfig's fields in declaration order
    (pcs/ligerito.rs:627-651, LigeritoLevelConfig :541-605,
    FinalBlockConfig :607-614) under bincode 1.3's fixint defaults."""
    def s(x):
        return u64(len(x)) + x.encode()
    b = u64(row["m"]) + u64(row["log_n"]) + u64(row["initial_k"]) + u64(row["target_security_bits"])
    b += s(row["analysis_version"]) + s(row["field"]) + s("blake3")
    b += u32(0)  # grinding_step = post_commit_pre_queries, variant 0
    b += u64(len(row["levels"]))
    for lv, (pg, qb, ob) in zip(row["levels"], eps):
        b += u64(lv["log_inv_rate"]) + u64(lv["log_msg_cols"]) + u64(lv["log_num_interleaved"])
        b += u64(lv["k_recursive"])
        b += u32(1)  # regime = johnson_ood, variant 1 of {udr, johnson_ood}
        b += b"\x01" + f64(lv["eta"])  # eta: Some
        b += b"\x00"  # proximity_loss: None
        b += u64(lv["queries"]) + u64(lv["grinding_bits"]) + u64(lv["fold_grinding_bits"])
        b += u64(lv["ood_samples"]) + u64(lv["target_security_bits"])
        b += f64(pg) + f64(qb)
        b += b"\x01" + f64(ob)  # expected_eps_ood_bits: Some
    b += u64(row["yr_log_n"])
    return b


@pytest.mark.parametrize("row", [r for r in LADDERS if r["m"] in (22, 26, 30, 35)],
                         ids=lambda r: f"m{r['m']}_{r['profile']}")
def test_security_config_bincode(oh, row):
    cfg = P.custom_johnson_config(row["m"], row["log_inv_rate"], row["initial_k"])
    eps = [(lv["eps_pg"], lv["eps_query"], lv["eps_ood"]) for lv in cfg["levels"]]
    want = _expected_bincode(row, eps)
    assert P.security_config_bincode(cfg) == want
    lib = _call_stream(oh.oh_ladder_bincode, row["m"], row["log_inv_rate"], row["initial_k"])
    assert lib == want
    assert P.ladder_digest(cfg) == O.blake3(want)


# ---------------------------------------------------------------- statement
def _field(tag, kind, count, payload):
    return b"\x06" + bytes([tag, kind]) + u64(count) + payload + b"\x07"


def _usizes(tag, vs):
    return _field(tag, 3, len(vs), b"".join(u64(v) for v in vs))


def test_statement_frame_bytes(oh):
    """m = 26 (t = 15, s = 11) under custom:2:4, whose ladder is the
    reference's m26_slim.toml: the statement the transcript hashes, written
    from absorb_standalone_mod_q_statement's field list."""
    row = next(r for r in LADDERS if r["m"] == 26 and r["profile"] == "slim")
    lv = row["levels"]
    root = bytes(range(32))
    t, s = 15, 11
    want = slice_(b"bitz/ligerito-flock/statement-frame/v1")
    want += slice_(b"bitz/ligerito-flock/standalone-mod-q-statement/v1")
    want += _field(0x01, 1, 32, root)
    want += _usizes(0x02, [26]) + _usizes(0x03, [2]) + _usizes(0x04, [4])
    want += _field(0x05, 2, 1, b"\x00")  # LigeritoProfile::Fast
    want += _field(0x06, 2, 1, b"\x01")  # HashKind::Blake3
    want += _usizes(0x08, [x["log_inv_rate"] for x in lv])
    want += _usizes(0x09, [len(lv) - 1])
    want += _usizes(0x0a, [lv[0]["log_msg_cols"]]) + _usizes(0x0b, [4]) + _usizes(0x0c, [4])
    want += _usizes(0x0d, [x["log_msg_cols"] for x in lv[1:]])
    want += _usizes(0x0e, [x["k_recursive"] for x in lv[1:]])
    want += _usizes(0x0f, [x["queries"] for x in lv])
    want += _usizes(0x10, [x["grinding_bits"] for x in lv])
    want += _usizes(0x11, [x["fold_grinding_bits"] for x in lv])
    want += _usizes(0x12, [x["ood_samples"] for x in lv])
    want += _field(0x13, 2, 1, b"\x01")
    want += _usizes(0x20, [t]) + _usizes(0x21, [s]) + _usizes(0x22, [1])
    want += _field(0x30, 5, 1, u64(2) + u64(0))  # the generator x (smallest_generator = 2)
    want += _usizes(0x31, [111])  # min(113, 126 - 15)
    # Round 0 at m26 rate 1/4: 128 - log2(C(50, 2)) - log2(2^19 - 1) = 98.7
    # bits, 2 bits of grinding to 100; the word is 1 + 2
    want += _usizes(0x32, [3])
    want += slice_(b"bitz/bitz-opener/standalone/v1")
    cfg = P.custom_johnson_config(26, 2, 4)
    eps = [(x["eps_pg"], x["eps_query"], x["eps_ood"]) for x in cfg["levels"]]
    want += slice_(O.blake3(_expected_bincode(row, eps)))
    # Round 0's header: its domain, m_p as u64, the grinding bits as u32
    want += slice_(b"bitz/core/ood-round/v1") + slice_(u64(19)) + slice_(u32(2))
    tr = P.outer_statement(root, t, s, cfg)
    assert bytes(tr.buf) == want
    lib = _call_stream(oh.oh_statement_stream, _buf(root), t, s, 2, 4, _buf(u64(2) + u64(0)))
    assert lib == want


@pytest.mark.parametrize("t,s,r", [(12, 8, 2), (17, 13, 2), (16, 12, 1), (14, 10, 3)])
def test_statement_stream_library_equals_oracle(oh, t, s, r):
    root = O.blake3(bytes([t, s, r]))
    cfg = P.custom_johnson_config(t + s, r, 4)
    lib = _call_stream(oh.oh_statement_stream, _buf(root), t, s, r, 4, _buf(u64(2) + u64(0)))
    assert lib == bytes(P.outer_statement(root, t, s, cfg).buf)


def test_claim_frame_bytes(oh):
    q, claimed = (1 << 110) + 12345, (1 << 109) + 777
    want = slice_(b"bitz/ligerito-flock/statement-frame/v1")
    want += slice_(b"bitz/ligerito-flock/standalone-mod-q-claim/v1")
    want += _field(0x30, 4, 2, q.to_bytes(16, "little") + claimed.to_bytes(16, "little"))
    tr = P.Transcript()
    P.absorb_standalone_claim(tr, q, claimed)
    assert bytes(tr.buf) == want
    lib = _call_stream(oh.oh_claim_stream, _buf(q.to_bytes(16, "little")), _buf(claimed.to_bytes(16, "little")))
    assert lib == want


# ------------------------------------------------------------ prime search
def test_prime_policy_closed_form():
    """budgets() by hand for the standalone widths: security = max(128 + 16,
    128) = 144; a b-bit interval has 64 b candidates; rounds =
    ceil((144 + bits(64 b - 1)) / 2); draws = 144 + bits(64 b - 1) +
    bits(rounds)."""
    assert P.prime_budgets(128, 64 * 111) == (79, 164)  # 7103 has 13 bits; 79 has 7
    assert P.prime_budgets(128, 64 * 113) == (79, 164)
    assert P.prime_budgets(128, 64 * 32) == (78, 162)  # 2047: 11 bits -> ceil(155/2) = 78, 7 bits
    assert P.prime_budgets(128, 64 * 120) == (79, 164)
    # a 100-bit target: security max(116, 128) = 128; one candidate: 0 bits,
    # ceil(128/2) = 64 rounds (7 bits)
    assert P.prime_budgets(100, 1) == (64, 135)


def _is_square(n):
    r = math.isqrt(n)
    return r * r == n


def _bpsw(n: int) -> bool:
    """Baillie-PSW: a strong base-2 Fermat test and a strong Lucas test with
    Selfridge's parameters (no known counterexample)."""
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(2, d, n)
    if x not in (1, n - 1):
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    if _is_square(n):
        return False
    D = 5
    while True:
        j = _jacobi(D, n)
        if j == -1:
            break
        if j == 0 and abs(D) != n:
            return False
        D = -D - 2 if D > 0 else -D + 2
    P_, Q = 1, (1 - D) // 4
    d, s = n + 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    U, V, Qk = 1, P_, Q % n
    inv2 = (n + 1) // 2
    for bit in bin(d)[3:]:
        U, V = U * V % n, (V * V - 2 * Qk) % n
        Qk = Qk * Qk % n
        if bit == "1":
            U, V = (P_ * U + V) * inv2 % n, (D * U + P_ * V) * inv2 % n
            Qk = Qk * Q % n
    if U == 0 or V == 0:
        return True
    for _ in range(s - 1):
        V = (V * V - 2 * Qk) % n
        Qk = Qk * Qk % n
        if V == 0:
            return True
    return False


def _jacobi(a: int, n: int) -> int:
    a %= n
    r = 1
    while a:
        while a % 2 == 0:
            a //= 2
            if n % 8 in (3, 5):
                r = -r
        a, n = n, a
        if a % 4 == 3 and n % 4 == 3:
            r = -r
        a %= n
    return r if n == 1 else 0


def test_bpsw_sanity():
    assert [n for n in range(2, 60) if _bpsw(n)] == [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59]
    assert _bpsw((1 << 127) - 1) and not _bpsw((1 << 127) + 1)
    assert not _bpsw(3825123056546413051)  # a strong pseudoprime to bases 2..23


@pytest.mark.parametrize("bits", [32, 64, 111, 113, 120])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sampled_prime(oh, bits, seed):
    pre = bytes([seed, bits]) * 7
    tr = P.Transcript()
    tr.absorb_slice(pre)
    start = len(tr.buf)
    q = P.sample_prime(tr, bits)
    assert (1 << (bits - 1)) <= q < (1 << bits) and _bpsw(q)
    # the policy words come first, then the sampling reads
    top = 1 << (bits - 1)
    rounds, draws = P.prime_budgets(128, 64 * bits)
    head = (slice_(b"bitz/shared-prime-sampling/v1") + slice_(top.to_bytes(16, "little")) +
            slice_((top | (top - 1)).to_bytes(16, "little")) + slice_(u32(128)) + slice_(u64(64 * bits)) +
            slice_(u64(draws)))
    assert bytes(tr.buf[start:start + len(head)]) == head
    qb = (C.c_uint8 * 16)()
    lib = _call_stream(oh.oh_sample_prime, _buf(pre), len(pre), bits, C.cast(qb, C.c_void_p))
    assert int.from_bytes(bytes(qb), "little") == q
    assert lib == bytes(tr.buf[start:])


def test_sample_below_reads_whole_words():
    """sample_below: one 8-byte sampling read per word, masked to the
    bound's width; a bound of 1 reads nothing."""
    tr = P.Transcript()
    n0 = len(tr.buf)
    assert P.sample_below(tr, 1, 10) == 0 and len(tr.buf) == n0
    v = P.sample_below(tr, 1 << 70, 10)
    assert v < 1 << 70
    frame = len(b"bitz/sampling-read/v1") + 8 + 8
    assert len(tr.buf) - n0 == 2 * frame


def test_sample_residue_library_equals_oracle(oh):
    q = (1 << 110) + 31
    for k in range(4):
        pre = bytes([k]) * 9
        tr = P.Transcript()
        tr.absorb_slice(pre)
        want = P.sample_residue(tr, q)
        out = (C.c_uint8 * 16)()
        oh.oh_sample_residue(_buf(pre), len(pre), _buf(q.to_bytes(16, "little")), C.cast(out, C.c_void_p))
        assert int.from_bytes(bytes(out), "little") == want < q


# ------------------------------------------------------------ grinding
def test_grinding_boundary_bytes(oh):
    pre = b"round zero"
    tr = P.Transcript()
    tr.absorb_slice(pre)
    start = len(tr.buf)
    seed = P.grinding_seed(tr, P.OOD_GRIND_DOMAIN, 0, 6)
    want = (slice_(b"bitz/spartan/fiat-shamir-grinding/v1") + slice_(b"bitz/core/ood-round-grinding/v1") +
            slice_(u64(0)) + slice_(u32(6)) + b"\x12" + seed + b"\x34")
    assert bytes(tr.buf[start:]) == want and len(seed) == 32
    sb = (C.c_uint8 * 32)()
    lib = _call_stream(oh.oh_grinding_seed, _buf(pre), len(pre), P.OOD_GRIND_DOMAIN, 0, 6, C.cast(sb, C.c_void_p))
    assert lib == want and bytes(sb) == seed


def _lz_bits(h: bytes) -> int:
    n = 0
    for b in h:
        if b:
            return n + 8 - b.bit_length()
        n += 8
    return n


@pytest.mark.parametrize("bits", [1, 4, 8, 11])
def test_round0_nonce_is_smallest(bits):
    """find_grinding_nonce: the smallest nonce with BLAKE3(seed || nonce_le)
    having >= bits leading zero bits (blake3x4.rs:42-62), by scanning every
    nonce below it with the pinned BLAKE3."""
    tr = P.Transcript()
    tr.absorb_slice(bytes([bits]) * 5)
    seed = P.grinding_seed(tr, P.OOD_GRIND_DOMAIN, 0, bits)
    nonce = O.pow_find(seed[:16], seed[16:], bits)
    assert _lz_bits(O.blake3(seed + u64(nonce))) >= bits
    assert all(_lz_bits(O.blake3(seed + u64(n))) < bits for n in range(nonce))
