"""CPU restatement of the committed bit grid of the native multiplication
relations (MulLayout / MulWitness, src/piop/spartan/mul.rs).

TEST INFRASTRUCTURE ONLY (see oracle.py): the checker of the device packer.

Pins (tests/test_oracle_mul.py): the packing cases of the reference's own
test (mul/tests.rs:7-83: fixed rows, n in {37, 257, 2053, 4097}, shifts
-1..2, expected words from its slot-major index formula), the layout bounds
test (mul/tests.rs:85-100) and the shape of the u64 multiplication workload at
2^20 multiplications (t = 16, s = 12); and, for random batches, decoding the
grid back through bitz_cell recovers x, y, lo, hi with x * y = lo + hi 2^B.
"""
from __future__ import annotations

import numpy as np

PACK_BITS = 7  # src/bitz/params.rs:12


def reference_log_rows(n: int) -> int:
    """src/bitz/params.rs:30-35: ceil(3n/5) - 1, at least PACK_BITS, at most n - 1."""
    return min(max(-(-3 * n // 5) - 1, PACK_BITS), max(n - 1, 0))


class LayoutError(ValueError):
    pass


def layout(word_bits: int, multiplications: int, split_shift: int = 0) -> dict:
    """MulLayout::new (mul.rs:71-93), default_col_vars (:119-124),
    with_split_shift (:125-133), bitz_params (:134-139)."""
    if multiplications == 0:
        raise LayoutError("EmptyBatch")
    cap = 1
    while cap < max(multiplications, 256):
        cap <<= 1
    if cap * 4 * word_bits >= 1 << 64:
        raise LayoutError("DomainTooLarge")
    g = cap.bit_length() - 1
    slot_log = (4 * word_bits).bit_length() - 1
    n = g + slot_log
    default = min(max(n - reference_log_rows(n), 0), max(g - 6, 0))
    s = default + split_shift
    if s < 0 or s > g:
        raise LayoutError("DomainTooLarge")
    return {"gate_vars": g, "col_vars": s, "row_vars": g - s + slot_log, "capacity": cap, "word_bits": word_bits}


def bit_rows(lay: dict, x, y, lo, hi) -> np.ndarray:
    """write_bit_rows_bitwise_at (mul.rs:310-323): every set bit of limb l of
    gate i goes to bitz_cell(l B + b, i) (mul.rs:140-150). Limbs are Python
    int sequences of length n (< capacity; padding gates are zero). Returns
    the grid as 2^s columns of 2^(t-6) u64 words, flattened column-major."""
    B = lay["word_bits"]
    g, s, t = lay["gate_vars"], lay["col_vars"], lay["row_vars"]
    high = g - s
    words = 1 << (t - 6)
    rows = np.zeros((1 << s) * words, dtype=np.uint64)
    n = len(x)
    gates = np.arange(n, dtype=np.uint64)
    col = gates & np.uint64((1 << s) - 1)
    ghi = gates >> np.uint64(s)
    for limb, vals in enumerate((x, y, lo, hi)):
        # the limb's bits as an (n, B) 0/1 matrix
        parts = [np.array([(int(v) >> (64 * k)) & ((1 << 64) - 1) for v in vals], dtype=np.uint64)
                 for k in range((B + 63) // 64)]
        for b in range(B):
            bit = (parts[b // 64] >> np.uint64(b % 64)) & np.uint64(1)
            sel = bit.astype(bool)
            if not sel.any():
                continue
            slot = limb * B + b
            rowbit = (np.uint64(slot) << np.uint64(high)) | ghi[sel]
            idx = col[sel] * np.uint64(words) + (rowbit >> np.uint64(6))
            np.bitwise_or.at(rows, idx.astype(np.int64), np.uint64(1) << (rowbit & np.uint64(63)))
    return rows


def decode(lay: dict, rows: np.ndarray, gates) -> list:
    """Read limbs [x, y, lo, hi] of the given gates back out of the grid
    through bitz_cell (the inverse of bit_rows)."""
    B = lay["word_bits"]
    g, s, t = lay["gate_vars"], lay["col_vars"], lay["row_vars"]
    high = g - s
    words = 1 << (t - 6)
    out = []
    for i in gates:
        c, gh = i & ((1 << s) - 1), i >> s
        limbs = []
        for limb in range(4):
            v = 0
            for b in range(B):
                rb = ((limb * B + b) << high) | gh
                w = int(rows[c * words + (rb >> 6)])
                v |= ((w >> (rb & 63)) & 1) << b
            limbs.append(v)
        out.append(limbs)
    return out


def products(word_bits: int, x, y):
    """MulWord::multiply (mul.rs:350-353, 376-379, 402-404): lo, hi of x * y."""
    mask = (1 << word_bits) - 1
    p = [int(a) * int(b) for a, b in zip(x, y)]
    return [v & mask for v in p], [v >> word_bits for v in p]
