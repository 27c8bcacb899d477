"""ctypes wrapper over the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline. The
product package never imports it.

Gf128 arrays are numpy uint64 arrays of shape (n, 2) holding (lo, hi); their
bytes are the reference's 16-byte wire form lo_le || hi_le.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class GF(C.Structure):
    _fields_ = [("lo", C.c_uint64), ("hi", C.c_uint64)]


def _load():
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"oracle library not built: run `make -C {_HERE}`")
    lib = C.CDLL(_LIB_PATH)
    P = C.c_void_p
    sig = {
        "or_gf_mul": (GF, [GF, GF]),
        "or_gf_square": (GF, [GF]),
        "or_gf_inv": (GF, [GF]),
        "or_gf_pow": (GF, [GF, C.c_uint64, C.c_uint64]),
        "or_gf_mul_arrays": (None, [P, P, P, C.c_size_t]),
        "or_gf_uses_pclmul": (C.c_int, []),
        "or_gf_force_portable": (None, [C.c_int]),
        "or_blake3_hash": (None, [P, C.c_size_t, P, C.c_size_t]),
        "or_blake3_leaf_cv": (None, [P, C.c_size_t, P]),
        "or_blake3_parent_cv": (None, [P, P, P]),
        "or_blake3_compress": (None, [P, P, C.c_uint64, C.c_uint32, C.c_uint32, P]),
        "or_b3_state_size": (C.c_size_t, []),
        "or_b3_init": (None, [P]),
        "or_b3_update": (None, [P, P, C.c_size_t]),
        "or_b3_finalize": (None, [P, P, C.c_size_t]),
        "or_ntt_evals": (None, [C.c_int, P]),
        "or_ntt_twiddle": (GF, [C.c_int, P, C.c_int, C.c_uint64]),
        "or_ntt_forward_interleaved": (None, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
        "or_ntt_inverse_interleaved": (None, [P, C.c_int, C.c_int, C.c_int]),
        "or_merkle_tree": (None, [P, C.c_uint64, C.c_uint64, P]),
        "or_merkle_multi_proof": (C.c_size_t, [P, C.c_uint64, P, C.c_size_t, P]),
        "or_merkle_verify_multi": (C.c_int, [P, C.c_uint64, P, P, C.c_size_t, P, C.c_size_t]),
        "or_commit": (C.c_int, [P, C.c_int, C.c_int, C.c_int, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def gf(x) -> GF:
    if isinstance(x, GF):
        return x
    if isinstance(x, int):
        return GF(x & (2**64 - 1), x >> 64)
    lo, hi = x
    return GF(int(lo), int(hi))


def gf_int(g: GF) -> int:
    return g.lo | (g.hi << 64)


def mul(a: int, b: int) -> int:
    return gf_int(lib.or_gf_mul(gf(a), gf(b)))


def square(a: int) -> int:
    return gf_int(lib.or_gf_square(gf(a)))


def inv(a: int) -> int:
    return gf_int(lib.or_gf_inv(gf(a)))


def power(a: int, e: int) -> int:
    return gf_int(lib.or_gf_pow(gf(a), e & (2**64 - 1), e >> 64))


def mul_arrays(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.empty_like(a)
    lib.or_gf_mul_arrays(_ptr(a), _ptr(b), _ptr(out), a.shape[0])
    return out


def blake3(data: bytes, out_len: int = 32) -> bytes:
    buf = np.frombuffer(bytes(data), dtype=np.uint8).copy() if data else np.zeros(1, np.uint8)
    out = np.zeros(out_len, np.uint8)
    lib.or_blake3_hash(_ptr(buf), len(data), _ptr(out), out_len)
    return out.tobytes()


class Blake3:
    """incremental BLAKE3; digest() leaves the state open for more input"""

    def __init__(self):
        self._st = (C.c_uint8 * lib.or_b3_state_size())()
        lib.or_b3_init(self._st)

    def update(self, data: bytes) -> None:
        if data:
            b = (C.c_uint8 * len(data)).from_buffer_copy(bytes(data))
            lib.or_b3_update(self._st, b, len(data))

    def digest(self, out_len: int = 32) -> bytes:
        out = (C.c_uint8 * max(out_len, 1))()
        lib.or_b3_finalize(self._st, out, out_len)
        return bytes(out)[:out_len]


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's heavy loops (results do not depend on it)"""
    lib.or_set_threads(n)


def max_threads() -> int:
    return int(lib.or_max_threads())


def leaf_cv(data: bytes) -> bytes:
    buf = np.frombuffer(bytes(data), dtype=np.uint8).copy()
    out = np.zeros(32, np.uint8)
    lib.or_blake3_leaf_cv(_ptr(buf), len(data), _ptr(out))
    return out.tobytes()


def parent_cv(left: bytes, right: bytes) -> bytes:
    lb = np.frombuffer(left, dtype=np.uint8).copy()
    rb = np.frombuffer(right, dtype=np.uint8).copy()
    out = np.zeros(32, np.uint8)
    lib.or_blake3_parent_cv(_ptr(lb), _ptr(rb), _ptr(out))
    return out.tobytes()


def ntt_evals(ell: int) -> list[list[int]]:
    n = ell * (ell + 1) // 2
    buf = np.zeros((max(n, 1), 2), np.uint64)
    lib.or_ntt_evals(ell, _ptr(buf))
    rows, off = [], 0
    for i in range(ell):
        rows.append([int(buf[off + k, 0]) | (int(buf[off + k, 1]) << 64) for k in range(ell - i)])
        off += ell - i
    return rows


def ntt_forward(data: np.ndarray, log_d: int, log_lanes: int, start_layer: int = 0,
                ell: int | None = None) -> np.ndarray:
    out = np.ascontiguousarray(data, dtype=np.uint64).copy()
    assert out.shape == ((1 << (log_d + log_lanes)), 2)
    lib.or_ntt_forward_interleaved(_ptr(out), log_d, log_lanes, start_layer,
                                   log_d if ell is None else ell)
    return out


def ntt_inverse(data: np.ndarray, log_d: int, log_lanes: int, ell: int | None = None) -> np.ndarray:
    out = np.ascontiguousarray(data, dtype=np.uint64).copy()
    lib.or_ntt_inverse_interleaved(_ptr(out), log_d, log_lanes, log_d if ell is None else ell)
    return out


def merkle_tree(data: bytes | np.ndarray, num_leaves: int) -> np.ndarray:
    buf = np.frombuffer(bytes(data), dtype=np.uint8) if isinstance(data, (bytes, bytearray)) \
        else np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    buf = np.ascontiguousarray(buf)
    assert buf.size % num_leaves == 0
    tree = np.zeros((2 * num_leaves - 1, 32), np.uint8)
    lib.or_merkle_tree(_ptr(buf), num_leaves, buf.size // num_leaves, _ptr(tree))
    return tree


def merkle_multi_proof(tree: np.ndarray, num_leaves: int, positions) -> np.ndarray:
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.uint64))
    depth = max(num_leaves.bit_length() - 1, 1)
    out = np.zeros((max(len(pos) * depth, 1), 32), np.uint8)
    n = lib.or_merkle_multi_proof(_ptr(np.ascontiguousarray(tree)), num_leaves, _ptr(pos),
                                  len(pos), _ptr(out))
    return out[:n].copy()


def merkle_verify_multi(root: bytes, num_leaves: int, sorted_positions, leaf_hashes: np.ndarray,
                        proof: np.ndarray) -> bool:
    pos = np.ascontiguousarray(np.asarray(sorted_positions, dtype=np.uint64))
    rb = np.frombuffer(root, dtype=np.uint8).copy()
    lh = np.ascontiguousarray(leaf_hashes, dtype=np.uint8)
    pf = np.ascontiguousarray(proof, dtype=np.uint8) if len(proof) else np.zeros((1, 32), np.uint8)
    return bool(lib.or_merkle_verify_multi(_ptr(rb), num_leaves, _ptr(pos), _ptr(lh), len(pos),
                                           _ptr(pf), len(proof)))


def commit(packed: np.ndarray, log_msg_len: int, log_batch: int, log_inv_rate: int):
    """Returns (root bytes, codeword (n,2) u64, tree (2N-1,32) u8)."""
    packed = np.ascontiguousarray(packed, dtype=np.uint64)
    assert packed.shape == (1 << log_msg_len, 2)
    cw = np.zeros((1 << (log_msg_len + log_inv_rate), 2), np.uint64)
    n_leaves = 1 << (log_msg_len - log_batch + log_inv_rate)
    tree = np.zeros((2 * n_leaves - 1, 32), np.uint8)
    root = np.zeros(32, np.uint8)
    rc = lib.or_commit(_ptr(packed), log_msg_len, log_batch, log_inv_rate, _ptr(cw), _ptr(tree),
                       _ptr(root))
    if rc != 0:
        raise ValueError(f"or_commit rejected the geometry (rc={rc})")
    return root.tobytes(), cw, tree


# ---------------------------------------------------------------- opening side
def _sig_open():
    P = C.c_void_p
    sig = {
        "or_eq_table": (None, [P, C.c_int, P]),
        "or_mle_eval": (GF, [P, P, C.c_int]),
        "or_fold_columns": (None, [P, C.c_int, C.c_int, P, P]),
        "or_pow_many": (None, [GF, P, C.c_size_t, P]),
        "or_forest_leaves": (None, [P, C.c_int, C.c_int, P, P]),
        "or_product_level": (None, [P, C.c_uint64, P]),
        "or_gkr_round_sums": (None, [P, P, C.c_uint64, P, C.c_int, P]),
        "or_fold_halves": (None, [P, C.c_uint64, GF]),
        "or_fold_halves_out": (None, [P, C.c_uint64, GF, P]),
        "or_xi_combined_rows": (None, [P, C.c_int, C.c_int, P, P]),
        "or_fold_rows_eq": (None, [P, C.c_int, C.c_int, P, P]),
        "or_round_polynomial": (None, [P, P, C.c_uint64, P]),
        "or_fold_lsb": (None, [P, C.c_uint64, GF]),
        "or_fold_lsb_out": (None, [P, C.c_uint64, GF, P]),
        "or_set_threads": (None, [C.c_int]),
        "or_max_threads": (C.c_int, []),
        "or_fold_1b_rows": (None, [P, P, C.c_uint64, P]),
        "or_fold_b128_elems": (None, [P, C.c_uint64, P, P]),
        "or_batch_claims": (GF, [P, P]),
        "or_lig_round_msg": (None, [P, P, C.c_uint64, P]),
        "or_lig_glue": (None, [P, P, C.c_uint64, GF]),
        "or_eval_sk_at_vks": (None, [C.c_int, P]),
        "or_induce_basis": (None, [C.c_int, P, P, P, C.c_size_t, P]),
        "or_pow_valid": (C.c_int, [P, P, C.c_uint64, C.c_uint32]),
        "or_pow_find": (C.c_uint64, [P, P, C.c_uint32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


_sig_open()


def _gfa(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64).reshape(-1, 2))


def ints_to_gf(xs) -> np.ndarray:
    return np.array([[x & (2**64 - 1), x >> 64] for x in xs], np.uint64).reshape(-1, 2)


def gf_to_ints(a) -> list[int]:
    a = np.asarray(a, dtype=np.uint64).reshape(-1, 2)
    return [int(x[0]) | (int(x[1]) << 64) for x in a]


def eq_table(r) -> np.ndarray:
    r = _gfa(r)
    out = np.zeros((1 << len(r), 2), np.uint64)
    lib.or_eq_table(_ptr(r), len(r), _ptr(out))
    return out


def mle_eval(v, point) -> int:
    v, point = _gfa(v), _gfa(point)
    assert v.shape[0] == 1 << point.shape[0]
    return gf_int(lib.or_mle_eval(_ptr(v), _ptr(point), point.shape[0]))


def fold_columns(rows: np.ndarray, t: int, s: int, weights) -> np.ndarray:
    rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
    w = _gfa(weights)
    out = np.zeros((1 << s, 2), np.uint64)
    lib.or_fold_columns(_ptr(rows), t, s, _ptr(w), _ptr(out))
    return out


def pow_many(g: int, exps) -> np.ndarray:
    e = _gfa(exps)
    out = np.zeros_like(e)
    lib.or_pow_many(gf(g), _ptr(e), e.shape[0], _ptr(out))
    return out


def forest_leaves(rows, t, s, images) -> np.ndarray:
    rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
    im = _gfa(images)
    out = np.zeros((1 << (t + s), 2), np.uint64)
    lib.or_forest_leaves(_ptr(rows), t, s, _ptr(im), _ptr(out))
    return out


def product_level(v) -> np.ndarray:
    v = _gfa(v)
    out = np.zeros((v.shape[0] // 2, 2), np.uint64)
    lib.or_product_level(_ptr(v), v.shape[0], _ptr(out))
    return out


def gkr_round_sums(l, r, eq, send_one: bool) -> tuple[int, int]:
    l, r, eq = _gfa(l), _gfa(r), _gfa(eq)
    h = eq.shape[0]
    out = np.zeros((2, 2), np.uint64)
    lib.or_gkr_round_sums(_ptr(l), _ptr(r), h, _ptr(eq), 1 if send_one else 0, _ptr(out))
    return tuple(gf_to_ints(out))


def fold_halves(v, rho: int) -> np.ndarray:
    v = _gfa(v)
    h = v.shape[0] // 2
    out = np.empty((h, 2), np.uint64)
    lib.or_fold_halves_out(_ptr(v), h, gf(rho), _ptr(out))
    return out


def xi_combined_rows(rows, t, s, w2) -> np.ndarray:
    rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
    w2 = _gfa(w2)
    out = np.zeros((1 << t, 2), np.uint64)
    lib.or_xi_combined_rows(_ptr(rows), t, s, _ptr(w2), _ptr(out))
    return out


def fold_rows_eq(rows, t, s, eq) -> np.ndarray:
    rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
    eq = _gfa(eq)
    out = np.zeros((1 << s, 2), np.uint64)
    lib.or_fold_rows_eq(_ptr(rows), t, s, _ptr(eq), _ptr(out))
    return out


def round_polynomial(m, w) -> list[int]:
    m, w = _gfa(m), _gfa(w)
    out = np.zeros((3, 2), np.uint64)
    lib.or_round_polynomial(_ptr(m), _ptr(w), m.shape[0], _ptr(out))
    return gf_to_ints(out)


def fold_lsb(v, r: int) -> np.ndarray:
    v = _gfa(v)
    out = np.empty((v.shape[0] // 2, 2), np.uint64)
    lib.or_fold_lsb_out(_ptr(v), v.shape[0], gf(r), _ptr(out))
    return out


def fold_1b_rows(p, suffix) -> np.ndarray:
    p, suffix = _gfa(p), _gfa(suffix)
    out = np.zeros((128, 2), np.uint64)
    lib.or_fold_1b_rows(_ptr(p), _ptr(suffix), p.shape[0], _ptr(out))
    return out


def fold_b128_elems(suffix, w) -> np.ndarray:
    suffix, w = _gfa(suffix), _gfa(w)
    out = np.zeros_like(suffix)
    lib.or_fold_b128_elems(_ptr(suffix), suffix.shape[0], _ptr(w), _ptr(out))
    return out


def batch_claims(claims, w) -> int:
    claims, w = _gfa(claims), _gfa(w)
    return gf_int(lib.or_batch_claims(_ptr(claims), _ptr(w)))


def lig_round_msg(f, b) -> tuple[int, int]:
    f, b = _gfa(f), _gfa(b)
    out = np.zeros((2, 2), np.uint64)
    lib.or_lig_round_msg(_ptr(f), _ptr(b), f.shape[0], _ptr(out))
    return tuple(gf_to_ints(out))


def lig_glue(acc, bnew, alpha: int) -> np.ndarray:
    acc, bnew = _gfa(acc).copy(), _gfa(bnew)
    lib.or_lig_glue(_ptr(acc), _ptr(bnew), acc.shape[0], gf(alpha))
    return acc


def eval_sk_at_vks(log_n: int) -> np.ndarray:
    out = np.zeros((log_n + 1, 2), np.uint64)
    lib.or_eval_sk_at_vks(log_n, _ptr(out))
    return out


def induce_basis(log_n: int, queries, alpha_w) -> np.ndarray:
    sks = eval_sk_at_vks(log_n)
    q = np.ascontiguousarray(np.asarray(queries, dtype=np.uint64))
    aw = _gfa(alpha_w)
    out = np.zeros((1 << log_n, 2), np.uint64)
    lib.or_induce_basis(log_n, _ptr(sks), _ptr(q), _ptr(aw), len(q), _ptr(out))
    return out


def pow_find(prefix: bytes, seed: bytes, bits: int) -> int:
    pb = np.frombuffer(prefix, np.uint8).copy()
    sb = np.frombuffer(seed, np.uint8).copy()
    return int(lib.or_pow_find(_ptr(pb), _ptr(sb), bits))


def pow_valid(prefix: bytes, seed: bytes, nonce: int, bits: int) -> bool:
    pb = np.frombuffer(prefix, np.uint8).copy()
    sb = np.frombuffer(seed, np.uint8).copy()
    return bool(lib.or_pow_valid(_ptr(pb), _ptr(sb), nonce, bits))
