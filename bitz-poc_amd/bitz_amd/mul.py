"""Witnesses of the native multiplication relations on the device: the
mirror of MulLayout / MulWitness (src/piop/spartan/mul.rs:62-330).

A batch of n multiplications x * y = lo + hi * 2^B (B = 32, 64 or 128) is
committed as the bits of its four limbs in a 2^t x 2^s grid; bitz_bit_rows
returns that grid as the column-major u64 words Commitment takes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib


class MulError(ValueError):
    """EmptyBatch / DomainTooLarge / InvalidRowShape (mul.rs:50-60)."""


@dataclass(frozen=True)
class MulLayout:
    """MulLayout::new + with_split_shift (mul.rs:70-136)."""
    word_bits: int
    multiplications: int
    gate_vars: int
    col_vars: int
    row_vars: int
    split_shift: int = 0

    @classmethod
    def new(cls, word_bits: int, multiplications: int, split_shift: int = 0) -> "MulLayout":
        out = (C.c_uint32 * 4)()
        rc = _lib.load().bitz_mul_layout(word_bits, multiplications, split_shift, C.cast(out, C.c_void_p))
        if rc == -4:
            _lib.check(rc, "bitz_mul_layout")
        if rc != 0:
            msg = _lib.load().bitz_last_error()
            raise MulError(msg.decode() if msg else "bitz_mul_layout failed")
        return cls(word_bits, multiplications, out[0], out[1], out[2], split_shift)

    @property
    def capacity(self) -> int:
        return 1 << self.gate_vars

    @property
    def t(self) -> int:
        return self.row_vars

    @property
    def s(self) -> int:
        return self.col_vars

    @property
    def rows_words(self) -> int:
        return 1 << (self.row_vars + self.col_vars - 6)

    def bitz_cell(self, bit_slot: int, gate: int):
        """(row bit, column) of slot bit_slot of gate (mul.rs:139-150), or None."""
        if bit_slot >= 4 * self.word_bits or gate >= self.capacity:
            return None
        high = self.gate_vars - self.col_vars
        return (bit_slot << high) | (gate >> self.col_vars), gate & ((1 << self.col_vars) - 1)


class MulWitness:
    """Limb blocks [x | y | lo | hi] of a batch in device memory.

    x, y (and lo, hi) are cuda tensors: int64 (n,) for B = 64, int32 (n,) for
    B = 32, int64 (n, 2) (low, high words) for B = 128. Without lo/hi the
    limbs are the exact product (MulWitness::from_inputs); with them the
    supplied limbs are kept (MulWitness::from_rows)."""

    def __init__(self, ctx, word_bits: int, x, y, lo=None, hi=None, split_shift: int = 0):
        if (lo is None) != (hi is None):
            raise ValueError("supply both product limbs or neither")
        n = x.shape[0]
        want = {32: 4, 64: 8, 128: 16}.get(word_bits)
        if want is None:
            raise ValueError("word_bits must be 32, 64 or 128")
        for t in (x, y) + ((lo, hi) if lo is not None else ()):
            if not t.is_cuda or not t.is_contiguous() or t.shape[0] != n or t.numel() * t.element_size() != n * want:
                raise ValueError(f"limb tensors must be contiguous cuda tensors of {n} x {want} bytes")
        self.ctx = ctx
        self.layout = MulLayout.new(word_bits, n, split_shift)
        self._limbs = (x, y, lo, hi)

    def write_bitz_bit_rows(self, rows) -> None:
        """Overwrite every word of rows (a cuda tensor of rows_words u64)."""
        x, y, lo, hi = self._limbs
        lay = self.layout
        if not rows.is_cuda or not rows.is_contiguous() or rows.numel() * rows.element_size() != lay.rows_words * 8:
            raise MulError("packed rows do not match the committed layout")

        def p(t):
            return C.c_void_p(t.data_ptr()) if t is not None else None

        _lib.call("bitz_mul_bit_rows_dev", self.ctx.ptr, lay.word_bits, lay.multiplications, lay.split_shift,
                  p(x), p(y), p(lo), p(hi), p(rows), lay.rows_words)

    def bitz_bit_rows(self):
        import torch
        rows = torch.empty(self.layout.rows_words, dtype=torch.int64, device=self._limbs[0].device)
        self.write_bitz_bit_rows(rows)
        return rows


def prove_u64(ctx, x, y, split_shift: int = 0, lo=None, hi=None):
    """protocol::commit + protocol::prove of the u64 multiplication relation
    (src/piop/spartan/protocol/mod.rs:1117-1170): x, y (and optionally the
    supplied product limbs lo, hi) are contiguous int64 cuda tensors of n
    values. Returns (root, outer transcript state digest, proof bytes)."""
    if (lo is None) != (hi is None):
        raise ValueError("supply both product limbs or neither")
    n = x.shape[0]
    for t in (x, y) + ((lo, hi) if lo is not None else ()):
        if not t.is_cuda or not t.is_contiguous() or t.shape[0] != n or t.element_size() * t.numel() != 8 * n:
            raise ValueError(f"operand tensors must be contiguous cuda tensors of {n} u64")

    def p(t):
        return C.c_void_p(t.data_ptr()) if t is not None else None

    root = (C.c_uint8 * 32)()
    state = (C.c_uint8 * 32)()
    plen = C.c_uint64(0)
    lib = _lib.load()
    # one proof into a buffer larger than any supported shape's proof; a
    # second run only if the library reports a longer one
    cap = 1 << 22
    while True:
        buf = (C.c_uint8 * cap)()
        rc = lib.bitz_prove_mul_u64(ctx.ptr, p(x), p(y), p(lo), p(hi), n, split_shift, root, state, buf, cap,
                                    C.byref(plen))
        if rc == _lib.BITZ_ERR_ARG and plen.value > cap:
            cap = plen.value
            continue
        _lib.check(rc, "bitz_prove_mul_u64")
        return bytes(root), bytes(state), bytes(buf)[:plen.value]


def verify_u64(ctx, root: bytes, n: int, proof: bytes, split_shift: int = 0) -> bytes:
    """protocol::verify of a u64 multiplication proof; raises BitzError when
    rejected, returns the outer transcript state digest when accepted."""
    state = (C.c_uint8 * 32)()
    pb = (C.c_uint8 * max(len(proof), 1)).from_buffer_copy(proof or b"\0")
    _lib.call("bitz_verify_mul_u64", ctx.ptr, (C.c_uint8 * 32).from_buffer_copy(root), n, split_shift, pb,
              len(proof), state)
    return bytes(state)
