"""Device-side API over libbitz_hip: a context bound to one GPU and to
torch's current HIP stream, with entry points taking torch tensors that live
in HBM. Gf128 tensors are int64 tensors of shape (n, 2) holding (lo, hi)
words; their bytes are the reference's 16-byte wire form.

PyTorch supplies device memory and the stream; every computation runs in the
HIP kernels of libbitz_hip.so.
"""
from __future__ import annotations

import ctypes as C

from . import _lib


def _torch():
    import torch
    return torch


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


class Context:
    """One device + stream. Not thread-safe."""

    def __init__(self, device: int = 0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("bitz_amd.Context needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = device
        self.torch = torch
        lib = _lib.load()
        p = C.c_void_p()
        _lib.check(lib.bitz_ctx_create(device, C.byref(p)), "bitz_ctx_create")
        self._ptr = p
        self.bind_current_stream()

    @property
    def ptr(self) -> C.c_void_p:
        return self._ptr

    def bind_current_stream(self) -> None:
        torch = self.torch
        with torch.cuda.device(self.device):
            s = torch.cuda.current_stream().cuda_stream
        _lib.call("bitz_ctx_set_stream", self._ptr, C.c_void_p(s))

    def synchronize(self) -> None:
        _lib.call("bitz_ctx_synchronize", self._ptr)

    def close(self) -> None:
        if self._ptr:
            _lib.load().bitz_ctx_destroy(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- helpers ----
    def empty_gf(self, n: int):
        return self.torch.empty((n, 2), dtype=self.torch.int64, device=f"cuda:{self.device}")

    def _check_gf(self, t, n: int | None = None, name: str = "tensor"):
        torch = self.torch
        if t.device.type != "cuda" or t.dtype != torch.int64 or t.dim() != 2 or t.shape[1] != 2 \
                or not t.is_contiguous():
            raise ValueError(f"{name}: expected a contiguous cuda int64 tensor of shape (n, 2)")
        if n is not None and t.shape[0] != n:
            raise ValueError(f"{name}: expected {n} elements, got {t.shape[0]}")

    # ---- arithmetic ----
    def gf_mul(self, a, b):
        self._check_gf(a, name="a")
        self._check_gf(b, a.shape[0], "b")
        out = self.empty_gf(a.shape[0])
        _lib.call("bitz_gf_mul_dev", self._ptr, _ptr(a), _ptr(b), _ptr(out), a.shape[0])
        return out

    def gf_mul_const(self, a, c: int):
        self._check_gf(a, name="a")
        out = self.empty_gf(a.shape[0])
        cb = (C.c_uint8 * 16).from_buffer_copy(int(c).to_bytes(16, "little"))
        _lib.call("bitz_gf_mul_const_dev", self._ptr, _ptr(a), C.cast(cb, C.c_void_p), _ptr(out),
                  a.shape[0])
        return out

    def blake3_leaves(self, data_u8, n: int, leaf_size: int):
        out = self.torch.empty((n, 32), dtype=self.torch.uint8, device=data_u8.device)
        _lib.call("bitz_blake3_leaves_dev", self._ptr, _ptr(data_u8), n, leaf_size, _ptr(out))
        return out

    # ---- RS encode / Merkle / commit ----
    def ntt_forward_(self, data, log_d: int, log_lanes: int, start_layer: int = 0,
                     ell: int | None = None):
        """In place (returns `data`)."""
        self._check_gf(data, 1 << (log_d + log_lanes), "data")
        _lib.call("bitz_ntt_forward_dev", self._ptr, _ptr(data), log_d, log_lanes, start_layer,
                  log_d if ell is None else ell)
        return data

    def merkle_tree(self, data, num_leaves: int):
        nbytes = data.numel() * data.element_size()
        if nbytes % num_leaves:
            raise ValueError("data length must be a multiple of num_leaves")
        tree = self.torch.empty((2 * num_leaves - 1, 32), dtype=self.torch.uint8,
                                device=data.device)
        _lib.call("bitz_merkle_tree_dev", self._ptr, _ptr(data), num_leaves, nbytes // num_leaves,
                  _ptr(tree))
        return tree

    def commit(self, packed, log_msg_len: int, log_batch: int, log_inv_rate: int,
               codeword=None, tree=None, want_root: bool = True):
        """PCS commit of a device-resident packed message. Returns
        (root bytes or None, codeword tensor, tree tensor)."""
        self._check_gf(packed, 1 << log_msg_len, "packed")
        torch = self.torch
        n_leaves = 1 << (log_msg_len - log_batch + log_inv_rate)
        if codeword is None:
            codeword = self.empty_gf(1 << (log_msg_len + log_inv_rate))
        if tree is None:
            tree = torch.empty((2 * n_leaves - 1, 32), dtype=torch.uint8, device=packed.device)
        root = (C.c_uint8 * 32)() if want_root else None
        _lib.call("bitz_commit_dev", self._ptr, _ptr(packed), log_msg_len, log_batch, log_inv_rate,
                  _ptr(codeword), _ptr(tree), C.cast(root, C.c_void_p) if want_root else None)
        return (bytes(root) if want_root else None), codeword, tree


class CommitItem(C.Structure):
    """bitz_commit_item (include/bitz_hip.h)"""
    _fields_ = [("packed", C.c_void_p), ("t", C.c_uint32), ("s", C.c_uint32), ("log_batch", C.c_uint32),
                ("log_inv_rate", C.c_uint32), ("codeword", C.c_void_p), ("tree", C.c_void_p),
                ("root", C.c_uint8 * 32)]


class CommitBatch:
    """Resident buffers of a batch of commitments, committed in one graph
    launch (bitz_commit_batch). specs: (t, s, log_batch, log_inv_rate) per
    item; `packed[i]` is item i's device message (2^(t+s-7) Gf128: the rows),
    filled by the caller; commit() returns the roots. The same object
    committed again replays the captured graph."""

    def __init__(self, ctx: "Context", specs):
        torch = ctx.torch
        self.ctx = ctx
        self.specs = list(specs)
        self.packed, self.codeword, self.tree = [], [], []
        self.items = (CommitItem * len(self.specs))()
        for i, (t, s, k, r) in enumerate(self.specs):
            lm = t + s - 7
            leaves = 1 << (lm - k + r)
            self.packed.append(ctx.empty_gf(1 << lm))
            self.codeword.append(ctx.empty_gf(1 << (lm + r)))
            self.tree.append(torch.empty((2 * leaves - 1, 32), dtype=torch.uint8, device="cuda"))
            it = self.items[i]
            it.packed, it.codeword, it.tree = _ptr(self.packed[i]), _ptr(self.codeword[i]), _ptr(self.tree[i])
            it.t, it.s, it.log_batch, it.log_inv_rate = t, s, k, r

    def commit(self) -> list[bytes]:
        _lib.call("bitz_commit_batch", self.ctx._ptr, len(self.specs), C.cast(self.items, C.c_void_p))
        return [bytes(it.root) for it in self.items]


# ----------------------------------------------------------------- opening
def _b16(x: int):
    return (C.c_uint8 * 16).from_buffer_copy(int(x).to_bytes(16, "little"))


def _vp(x):
    return C.cast(x, C.c_void_p)


def _host_gf(vals):
    buf = b"".join(int(v).to_bytes(16, "little") for v in vals)
    return (C.c_uint8 * max(len(buf), 1)).from_buffer_copy(buf or b"\0")


def _ints(buf, k: int):
    raw = bytes(buf)
    return [int.from_bytes(raw[16 * i:16 * i + 16], "little") for i in range(k)]


def kernel_stats(ctx: Context) -> dict:
    """{kernel: (launches, total_ms, algorithmic_bytes, algorithmic_int_ops)}"""
    names = C.create_string_buffer(1 << 16)
    vals = (C.c_double * 4096)()
    n = _lib.load().bitz_ctx_kernel_stats(ctx.ptr, names, 1 << 16, C.cast(vals, C.c_void_p), 4096)
    if n < 0:
        _lib.check(n, "bitz_ctx_kernel_stats")
    keys = names.value.decode().split("\n")[:n]
    return {k: tuple(vals[4 * i:4 * i + 4]) for i, k in enumerate(keys)}


def set_timing(ctx: Context, on: bool) -> None:
    _lib.call("bitz_ctx_set_timing", ctx.ptr, 1 if on else 0)


def set_timing_only(ctx: Context, kernel: str) -> None:
    """Time only the launches of `kernel` (fewer events in a timed region)."""
    _lib.call("bitz_ctx_set_timing_only", ctx.ptr, kernel.encode())


TUNE_KEYS = ("round_r_max", "round_min_chunks", "round_block_cap", "host_tail_log", "ntt_radix_log")


def get_tune(ctx: Context) -> dict:
    """The context's launch-shape choices, plus `tuned` (set by autotune)."""
    out = (C.c_uint32 * 6)()
    _lib.call("bitz_ctx_get_tune", ctx.ptr, C.cast(out, C.c_void_p))
    d = {k: int(out[i]) for i, k in enumerate(TUNE_KEYS)}
    d["tuned"] = bool(out[5])
    return d


def set_tune(ctx: Context, **kw) -> None:
    """Override some of the launch-shape choices (the rest keep their values)."""
    cur = get_tune(ctx)
    unknown = set(kw) - set(TUNE_KEYS)
    if unknown:
        raise ValueError(f"unknown tune keys {sorted(unknown)}")
    cur.update(kw)
    _lib.call("bitz_ctx_set_tune", ctx.ptr, *(int(cur[k]) for k in TUNE_KEYS))


def autotune(ctx: Context, t: int, s: int, log_batch: int, log_inv_rate: int) -> dict:
    """Time candidate launch shapes for jobs of this shape on the device and
    keep the fastest in the context; returns get_tune() plus `tune_ms`."""
    ms = C.c_double(0.0)
    _lib.call("bitz_ctx_autotune", ctx.ptr, t, s, log_batch, log_inv_rate, C.byref(ms))
    d = get_tune(ctx)
    d["tune_ms"] = ms.value
    return d


class OpenOps:
    """Thin wrappers over the opening kernels (device tensors in, device
    tensors or host ints out)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.torch = ctx.torch

    def _gf(self, n):
        return self.ctx.empty_gf(n)

    def eq_table(self, r: list[int]):
        out = self._gf(1 << len(r))
        _lib.call("bitz_eq_table_dev", self.ctx.ptr, _vp(_host_gf(r)), len(r), _ptr(out))
        return out

    def fold_columns(self, rows, t, s, w):
        out = self._gf(1 << s)
        _lib.call("bitz_fold_columns_dev", self.ctx.ptr, _ptr(rows), t, s, _ptr(w), _ptr(out))
        return out

    def pow_images(self, g: int, exps):
        out = self._gf(exps.shape[0])
        _lib.call("bitz_pow_images_dev", self.ctx.ptr, _vp(_b16(g)), _ptr(exps), exps.shape[0], _ptr(out))
        return out

    def pack_columns(self, rows, t, s):
        groups = ((1 << s) + 63) // 64
        out = self.torch.empty(groups << t, dtype=self.torch.int64, device=rows.device)
        _lib.call("bitz_pack_columns_dev", self.ctx.ptr, _ptr(rows), t, s, _ptr(out))
        return out

    def gkr_round_sums(self, l, r, h, eq_c, s_c, eq_y, send_one):
        buf = (C.c_uint8 * 32)()
        _lib.call("bitz_gkr_round_sums_dev", self.ctx.ptr, _ptr(l), _ptr(r), h, _ptr(eq_c), s_c, _ptr(eq_y),
                  1 if send_one else 0, _vp(buf))
        return _ints(buf, 2)

    def gkr_leaf_round_sums(self, packed, A, t, s, eq_c, s_c, eq_y, send_one):
        buf = (C.c_uint8 * 32)()
        _lib.call("bitz_gkr_leaf_round_sums_dev", self.ctx.ptr, _ptr(packed), _ptr(A), t, s, _ptr(eq_c), s_c,
                  _ptr(eq_y), 1 if send_one else 0, _vp(buf))
        return _ints(buf, 2)

    def gkr_fold_round_sums_(self, l, r, h, rho, eq_c, s_c, eq_y, send_one):
        """folds l and r (4h entries) in place to 2h and returns the round sums"""
        buf = (C.c_uint8 * 32)()
        _lib.call("bitz_gkr_fold_round_sums_dev", self.ctx.ptr, _ptr(l), _ptr(r), h, _vp(_b16(rho)),
                  _ptr(eq_c) if eq_c is not None else None,
                  s_c, _ptr(eq_y), 1 if send_one else 0, _vp(buf))
        return _ints(buf, 2)

    def gkr_leaf_round1_sums(self, packed, A, t, s, rho0, eq_c, eq_y, send_one):
        buf = (C.c_uint8 * 32)()
        _lib.call("bitz_gkr_leaf_round1_sums_dev", self.ctx.ptr, _ptr(packed), _ptr(A), t, s, _vp(_b16(rho0)),
                  _ptr(eq_c), _ptr(eq_y), 1 if send_one else 0, _vp(buf))
        return _ints(buf, 2)

    def gkr_leaf_fold2(self, packed, A, t, s, rho0, rho1, eq_c=None):
        n = 1 << (t + s - 3)
        l, r = self._gf(n), self._gf(n)
        _lib.call("bitz_gkr_leaf_fold2_dev", self.ctx.ptr, _ptr(packed), _ptr(A), t, s, _vp(_b16(rho0)),
                  _vp(_b16(rho1)), _ptr(eq_c) if eq_c is not None else None, _ptr(l), _ptr(r))
        return l, r

    def gkr_premul_round_sums_(self, l, r, h, eq_c, s, eq_y, send_one):
        """multiplies eq_c into r (2h entries) in place; returns the round sums"""
        buf = (C.c_uint8 * 32)()
        _lib.call("bitz_gkr_premul_round_sums_dev", self.ctx.ptr, _ptr(l), _ptr(r), h, _ptr(eq_c), s, _ptr(eq_y),
                  1 if send_one else 0, _vp(buf))
        return _ints(buf, 2)

    def gkr_round_sums_rowonly(self, l, r, h, s, eq_y, send_one):
        buf = (C.c_uint8 * 32)()
        _lib.call("bitz_gkr_round_sums_rowonly_dev", self.ctx.ptr, _ptr(l), _ptr(r), h, s, _ptr(eq_y),
                  1 if send_one else 0, _vp(buf))
        return _ints(buf, 2)

    def gkr_level2(self, packed, A, t, s):
        out = self._gf(1 << (t + s - 2))
        _lib.call("bitz_gkr_level2_dev", self.ctx.ptr, _ptr(packed), _ptr(A), t, s, _ptr(out))
        return out

    def gkr_level3(self, packed, A, t, s):
        out = self._gf(1 << (t + s - 3))
        _lib.call("bitz_gkr_level3_dev", self.ctx.ptr, _ptr(packed), _ptr(A), t, s, _ptr(out))
        return out

    def gkr_level1_round0_sums(self, packed, A, t, s, eq_c, eq_y, send_one):
        buf = (C.c_uint8 * 32)()
        _lib.call("bitz_gkr_level1_round0_sums_dev", self.ctx.ptr, _ptr(packed), _ptr(A), t, s, _ptr(eq_c),
                  _ptr(eq_y), 1 if send_one else 0, _vp(buf))
        return _ints(buf, 2)

    def gkr_level1_fold(self, packed, A, t, s, rho, eq_c=None):
        n = 1 << (t + s - 3)
        l, r = self._gf(n), self._gf(n)
        _lib.call("bitz_gkr_level1_fold_dev", self.ctx.ptr, _ptr(packed), _ptr(A), t, s, _vp(_b16(rho)),
                  _ptr(eq_c) if eq_c is not None else None, _ptr(l), _ptr(r))
        return l, r

    def tensor_table(self, a, b, n):
        out = self._gf(1 << n)
        _lib.call("bitz_tensor_table_dev", self.ctx.ptr, _ptr(a), _ptr(b), n, _ptr(out))
        return out

    def fold_halves_(self, v, h, rho):
        _lib.call("bitz_fold_halves_dev", self.ctx.ptr, _ptr(v), h, _vp(_b16(rho)))
        return v

    def gkr_leaf_fold(self, packed, A, t, s, rho):
        n = 1 << (t + s - 2)
        l, r = self._gf(n), self._gf(n)
        _lib.call("bitz_gkr_leaf_fold_dev", self.ctx.ptr, _ptr(packed), _ptr(A), t, s, _vp(_b16(rho)), _ptr(l),
                  _ptr(r))
        return l, r

    def gkr_level1(self, packed, A, t, s):
        out = self._gf(1 << (t + s - 1))
        _lib.call("bitz_gkr_level1_dev", self.ctx.ptr, _ptr(packed), _ptr(A), t, s, _ptr(out))
        return out

    def product_level(self, v):
        out = self._gf(v.shape[0] // 2)
        _lib.call("bitz_product_level_dev", self.ctx.ptr, _ptr(v), v.shape[0], _ptr(out))
        return out

    def xi_combined(self, packed, t, s, w2):
        out = self._gf(1 << t)
        _lib.call("bitz_xi_combined_dev", self.ctx.ptr, _ptr(packed), t, s, _ptr(w2), _ptr(out))
        return out

    def fold_rows_eq(self, rows, t, s, eq):
        out = self._gf(1 << s)
        _lib.call("bitz_fold_rows_eq_dev", self.ctx.ptr, _ptr(rows), t, s, _ptr(eq), _ptr(out))
        return out

    def round_poly(self, m, w):
        buf = (C.c_uint8 * 48)()
        _lib.call("bitz_round_poly_dev", self.ctx.ptr, _ptr(m), _ptr(w), m.shape[0], _vp(buf))
        return _ints(buf, 3)

    def fold_lsb(self, v, r):
        out = self._gf(v.shape[0] // 2)
        _lib.call("bitz_fold_lsb_dev", self.ctx.ptr, _ptr(v), v.shape[0], _vp(_b16(r)), _ptr(out))
        return out

    def fold_1b_rows(self, p, suffix):
        out = self._gf(128)
        _lib.call("bitz_fold_1b_rows_dev", self.ctx.ptr, _ptr(p), _ptr(suffix), p.shape[0], _ptr(out))
        return out

    def fold_b128(self, suffix, w):
        out = self._gf(suffix.shape[0])
        _lib.call("bitz_fold_b128_dev", self.ctx.ptr, _ptr(suffix), suffix.shape[0], _ptr(w), _ptr(out))
        return out

    def axpy_(self, acc, e, eta):
        _lib.call("bitz_axpy_dev", self.ctx.ptr, _ptr(acc), _ptr(e), acc.shape[0], _vp(_b16(eta)))
        return acc

    def lig_msg(self, f, b, with_eval=False):
        buf = (C.c_uint8 * 48)()
        _lib.call("bitz_lig_msg_dev", self.ctx.ptr, _ptr(f), _ptr(b), f.shape[0], 1 if with_eval else 0, _vp(buf))
        return _ints(buf, 3 if with_eval else 2)

    def lig_fold(self, f, b, r):
        fo, bo = self._gf(f.shape[0] // 2), self._gf(b.shape[0] // 2)
        _lib.call("bitz_lig_fold_dev", self.ctx.ptr, _ptr(f), _ptr(b), f.shape[0], _vp(_b16(r)), _ptr(fo), _ptr(bo))
        return fo, bo

    def induce_basis(self, log_n, queries, alpha_w):
        q = (C.c_uint64 * max(len(queries), 1))(*queries)
        out = self._gf(1 << log_n)
        _lib.call("bitz_induce_basis_dev", self.ctx.ptr, _vp(q), len(queries), log_n, _ptr(alpha_w), _ptr(out))
        return out

    def pow_find(self, prefix: bytes, seed: bytes, bits: int) -> int:
        p = (C.c_uint8 * 16).from_buffer_copy(prefix)
        s = (C.c_uint8 * 16).from_buffer_copy(seed)
        n = C.c_uint64()
        _lib.call("bitz_pow_find_dev", self.ctx.ptr, _vp(p), _vp(s), bits, C.byref(n))
        return n.value


# ------------------------------------------------------- standalone opening
class Commitment:
    """A commitment of host bit rows kept on the device (bitz_commit handle):
    Pcs::commit (src/bitz/pcs.rs:270-284)."""

    def __init__(self, ctx: Context, rows, t: int, s: int, log_batch: int, log_inv_rate: int):
        """rows: a host array (numpy u64) or a cuda tensor of 2^s * 2^(t-6) words."""
        import numpy as np
        self.ctx, self.t, self.s = ctx, t, s
        self.log_batch, self.log_inv_rate = log_batch, log_inv_rate
        words = (1 << s) << (t - 6)
        root = (C.c_uint8 * 32)()
        h = C.c_void_p()
        if hasattr(rows, "is_cuda") and rows.is_cuda:
            if rows.numel() * rows.element_size() != words * 8 or not rows.is_contiguous():
                raise ValueError("rows must hold 2^s columns of 2^(t-6) words")
            _lib.call("bitz_commit_device_rows", ctx.ptr, _ptr(rows), t, s, log_batch, log_inv_rate, _vp(root),
                      C.byref(h))
        else:
            rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
            if rows.size != words:
                raise ValueError("rows must hold 2^s columns of 2^(t-6) words")
            _lib.call("bitz_commit", ctx.ptr, rows.ctypes.data_as(C.c_void_p), t, s, log_batch, log_inv_rate,
                      _vp(root), C.byref(h))
        self._h = h
        self.root = bytes(root)

    def close(self):
        if self._h:
            _lib.load().bitz_commit_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def evaluation(self) -> int:
        out = (C.c_uint8 * 16)()
        _lib.call("bitz_standalone_evaluation", self.ctx.ptr, self._h, _vp(out))
        return int.from_bytes(bytes(out), "little")

    def prove(self, claimed: int, capacity: int = 1 << 22) -> bytes:
        buf = bytearray(capacity)
        n = C.c_uint64()
        _lib.call("bitz_prove_standalone", self.ctx.ptr, self._h, _vp(_b16(claimed)),
                  _vp((C.c_uint8 * capacity).from_buffer(buf)), capacity, C.byref(n))
        return bytes(memoryview(buf)[: n.value])


def verify_standalone(ctx: Context, root: bytes, t: int, s: int, log_inv_rate: int, initial_k: int,
                      claimed: int, proof: bytes) -> None:
    """verify_standalone (src/piop/spartan/protocol/bitz_opener.rs:709-747):
    returns when the proof is accepted, raises BitzError (code -5, "VERIFY",
    with the reason) when it is rejected."""
    if len(root) != 32:
        raise ValueError("root must be 32 bytes")
    rb = (C.c_uint8 * 32).from_buffer_copy(root)
    pb = (C.c_uint8 * max(len(proof), 1)).from_buffer_copy(proof if proof else b"\0")
    _lib.call("bitz_verify_standalone", ctx.ptr, _vp(rb), t, s, log_inv_rate, initial_k, _vp(_b16(claimed)),
              _vp(pb), len(proof))
