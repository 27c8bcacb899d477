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
