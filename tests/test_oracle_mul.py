"""The CPU restatement of the multiplication relations' committed bit grid
(oracle/mul_witness.py) against the reference's own packing test, its layout
bounds, the shape of the u64 workload at 2^20 multiplications, and decoding; and the library's host-side layout
(bitz_mul_layout, no GPU needed) against the restatement."""
import random

import pytest

import mul_witness as MW

# the fixed rows of native_packing_matches_reference_and_overwrites_reused_rows
# (src/piop/spartan/mul/tests.rs:44-83): (x, y, lo, hi) per word width
REF_ROWS = {
    32: (0xFFFFFFFF, 0x80102041, 0xFEFE1313, 0xFEEDFACE),
    64: (0xFFFFFFFFFFFFFFFF, 0x8010204180808081, 0xFEFE131341341234, 0xFEEDFACE87654321),
    128: ((1 << 128) - 1, (1 << 127) | 0x123456789ABCDEF, 0xFEFE131341341234, (1 << 120) | 0xFEEDFACE87654321),
}


def expected_slot_major(lay, n, row):
    """check_packing's expected grid (mul/tests.rs:7-24): bit `bit` of limb
    `limb` of gate `gate` at index (limb B + bit) * high + gate / cols of
    column gate % cols."""
    B = lay["word_bits"]
    cols = 1 << lay["col_vars"]
    high = 1 << (lay["gate_vars"] - lay["col_vars"])
    words = (1 << lay["row_vars"]) // 64
    out = [[0] * words for _ in range(cols)]
    for gate in range(n):
        for limb, value in enumerate(row):
            for bit in range(B):
                index = (limb * B + bit) * high + gate // cols
                out[gate % cols][index // 64] |= ((value >> bit) & 1) << (index % 64)
    return [w for col in out for w in col]


@pytest.mark.parametrize("B", [32, 64, 128])
@pytest.mark.parametrize("n", [37, 257, 2053, 4097])
@pytest.mark.parametrize("shift", [-1, 0, 1, 2])
def test_reference_packing_cases(B, n, shift):
    lay = MW.layout(B, n, shift)
    row = REF_ROWS[B]
    got = MW.bit_rows(lay, *([v] * n for v in row))
    assert [int(w) for w in got] == expected_slot_major(lay, n, row)


def test_layout_bounds_and_config_shapes():
    # packing_bounds_reject_overflow_and_invalid_splits (mul/tests.rs:85-100)
    with pytest.raises(MW.LayoutError):
        MW.layout(64, 0)
    with pytest.raises(MW.LayoutError):
        MW.layout(64, (1 << 64) - 1)
    with pytest.raises(MW.LayoutError):
        MW.layout(64, 1 << 24, 30)
    with pytest.raises(MW.LayoutError):
        MW.layout(64, 1 << 24, -30)
    # u64 multiplication at 2^20: n = 28, t = 16, s = 12 (gate_vars 20,
    # col_vars min(28 - 16, 20 - 6))
    lay = MW.layout(64, 1 << 20)
    assert (lay["gate_vars"], lay["col_vars"], lay["row_vars"]) == (20, 12, 16)
    # the 256-gate minimum capacity
    assert MW.layout(64, 1)["capacity"] == 256
    assert MW.layout(32, 257)["capacity"] == 512


@pytest.mark.parametrize("B", [32, 64, 128])
def test_decode_recovers_products(B):
    rng = random.Random(B)
    n = 300
    x = [rng.getrandbits(B) for _ in range(n)]
    y = [rng.getrandbits(B) for _ in range(n)]
    x[0], y[0] = (1 << B) - 1, (1 << B) - 1  # the largest product
    lo, hi = MW.products(B, x, y)
    for shift in (0, 1):
        lay = MW.layout(B, n, shift)
        rows = MW.bit_rows(lay, x, y, lo, hi)
        gates = [0, 1, 63, 64, 299] + rng.sample(range(n), 20)
        for i, (a, b, zl, zh) in zip(gates, MW.decode(lay, rows, gates)):
            assert (a, b) == (x[i], y[i])
            assert a * b == zl + (zh << B)
        # padding gates decode to zero
        assert MW.decode(lay, rows, [lay["capacity"] - 1]) == [[0, 0, 0, 0]]


def test_planted_index_mistake_is_caught():
    """a transposed (column, high-gate) split in the restatement fails the
    reference's expected grid"""
    lay = MW.layout(64, 2053)
    row = REF_ROWS[64]
    want = expected_slot_major(lay, 2053, row)
    bad = dict(lay, col_vars=lay["col_vars"] + 1, gate_vars=lay["gate_vars"])
    bad["row_vars"] = lay["row_vars"] - 1
    got = MW.bit_rows(bad, *([v] * 2053 for v in row))
    assert [int(w) for w in got] != want


def test_library_layout_matches_restatement():
    from bitz_amd.mul import MulError, MulLayout
    for B in (32, 64, 128):
        for n in (1, 37, 255, 256, 257, 4097, 1 << 20, (1 << 20) + 1, 1 << 30):
            for shift in (-3, -1, 0, 1, 2, 5):
                try:
                    want = MW.layout(B, n, shift)
                except MW.LayoutError:
                    with pytest.raises(MulError):
                        MulLayout.new(B, n, shift)
                    continue
                got = MulLayout.new(B, n, shift)
                assert (got.gate_vars, got.col_vars, got.row_vars) == (
                    want["gate_vars"], want["col_vars"], want["row_vars"])
    with pytest.raises(MulError):
        MulLayout.new(64, 0)
