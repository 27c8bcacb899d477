"""The device packer of the multiplication relations' committed bit grid
(bitz_mul_bit_rows_dev) against the CPU restatement (oracle/mul_witness.py):
the reference's own packing cases with supplied limbs, random batches with
device-computed products (padding, shifted splits, the few-high-gates
gather path), the u64 workload at 2^20 multiplications through decoding
and bit counts, and a commitment + standalone opening of a packed witness
equal to the oracle prover's bytes."""
import random

import numpy as np
import pytest

import mul_witness as MW
from gpu_util import ctx, log

pytestmark = pytest.mark.gpu

REF_ROWS = {
    32: (0xFFFFFFFF, 0x80102041, 0xFEFE1313, 0xFEEDFACE),
    64: (0xFFFFFFFFFFFFFFFF, 0x8010204180808081, 0xFEFE131341341234, 0xFEEDFACE87654321),
    128: ((1 << 128) - 1, (1 << 127) | 0x123456789ABCDEF, 0xFEFE131341341234, (1 << 120) | 0xFEEDFACE87654321),
}


def _limb_tensor(B, vals):
    import torch
    if B == 32:
        a = np.array(vals, dtype=np.uint32).view(np.int32)
    elif B == 64:
        a = np.array(vals, dtype=np.uint64).view(np.int64)
    else:
        a = np.array([[v & (2**64 - 1), v >> 64] for v in vals], dtype=np.uint64).view(np.int64)
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _device_rows(B, x, y, lo=None, hi=None, shift=0):
    from bitz_amd.mul import MulWitness
    args = [_limb_tensor(B, v) for v in (x, y)]
    kw = {}
    if lo is not None:
        kw = {"lo": _limb_tensor(B, lo), "hi": _limb_tensor(B, hi)}
    w = MulWitness(ctx(), B, *args, split_shift=shift, **kw)
    rows = w.bitz_bit_rows()
    return w.layout, rows.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("B", [32, 64, 128])
@pytest.mark.parametrize("n", [37, 257, 2053, 4097])
@pytest.mark.parametrize("shift", [-1, 0, 1, 2])
def test_reference_packing_cases(B, n, shift):
    log(f"mul rows: reference row B={B} n={n} shift={shift}")
    row = REF_ROWS[B]
    lay, got = _device_rows(B, *([v] * n for v in row), shift=shift)
    want = MW.bit_rows(MW.layout(B, n, shift), *([v] * n for v in row))
    assert lay.rows_words == want.size
    assert np.array_equal(got, want)


@pytest.mark.parametrize("B", [32, 64, 128])
@pytest.mark.parametrize("n,shift", [(5000, 0), (5000, 2), (300, 0), (300, 3), (1 << 14, -2)])
def test_random_products(B, n, shift):
    log(f"mul rows: random B={B} n={n} shift={shift}")
    rng = random.Random(1000 * B + n + shift)
    x = [rng.getrandbits(B) for _ in range(n)]
    y = [rng.getrandbits(B) for _ in range(n)]
    x[-1] = y[-1] = (1 << B) - 1
    lay, got = _device_rows(B, x, y, shift=shift)
    lo, hi = MW.products(B, x, y)
    want = MW.bit_rows(MW.layout(B, n, shift), x, y, lo, hi)
    assert np.array_equal(got, want)


def _decode_all(rows, t, s, B=64):
    """every gate's four limbs from the grid (vectorised bitz_cell inverse):
    bit (slot << (g - s)) | (gate >> s) of column gate mod 2^s"""
    cols = 1 << s
    bits = np.unpackbits(rows.view(np.uint8).reshape(cols, -1), axis=1, bitorder="little")
    high = (1 << t) // (4 * B)
    bits = bits.reshape(cols, 4 * B, high)
    out = []
    for limb in range(4):
        blk = bits[:, limb * B:(limb + 1) * B, :].transpose(2, 0, 1).reshape(high * cols, B)
        out.append(np.packbits(blk, axis=1, bitorder="little").view(np.uint64).reshape(-1))
    return out


def test_u64_at_2p20_multiplications():
    """the u64 multiplication workload's grid (t = 16, s = 12): every gate
    decodes to its limbs, lo = x * y mod 2^64 and hi the high word (from
    32-bit halves on the host); sampled gates also through the oracle's
    bitz_cell decoder with x * y = lo + hi 2^64"""
    import torch
    from bitz_amd.mul import MulWitness
    n = 1 << 20
    log("mul rows: u64 n=2^20")
    g = torch.Generator(device="cuda").manual_seed(20)
    x = torch.randint(-(2**63), 2**63 - 1, (n,), dtype=torch.int64, device="cuda", generator=g)
    y = torch.randint(-(2**63), 2**63 - 1, (n,), dtype=torch.int64, device="cuda", generator=g)
    w = MulWitness(ctx(), 64, x, y)
    assert (w.layout.t, w.layout.s) == (16, 12)
    rows_dev = w.bitz_bit_rows()
    torch.cuda.synchronize()
    rows = rows_dev.cpu().numpy().view(np.uint64)
    again = w.bitz_bit_rows()
    torch.cuda.synchronize()
    assert torch.equal(again, rows_dev), "two packings of one witness differ"
    xs, ys = x.cpu().numpy().view(np.uint64), y.cpu().numpy().view(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    xl, xh, yl, yh = xs & m, xs >> np.uint64(32), ys & m, ys >> np.uint64(32)
    ll, lh, hl, hh = xl * yl, xl * yh, xh * yl, xh * yh
    mid = (ll >> np.uint64(32)) + (lh & m) + (hl & m)
    lo = (ll & m) | ((mid & m) << np.uint64(32))
    hi = hh + (lh >> np.uint64(32)) + (hl >> np.uint64(32)) + (mid >> np.uint64(32))
    got = _decode_all(rows, 16, 12)
    for name, a, b in zip("x y lo hi".split(), got, (xs, ys, lo, hi)):
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"limb {name}: {bad.size} gates differ, first {bad[:8].tolist()}"
    lay = MW.layout(64, n)
    rng = random.Random(7)
    gates = [0, 1, 4095, 4096, n - 1] + rng.sample(range(n), 20)
    for i, (a, b, zl, zh) in zip(gates, MW.decode(lay, rows, gates)):
        assert (a, b) == (int(xs[i]), int(ys[i]))
        assert a * b == zl + (zh << 64)


def test_packed_witness_commit_and_open_equal_oracle():
    """2^12 u64 multiplications (t = 14, s = 6): commitment root, honest claim
    and proof bytes equal the oracle prover's, and the proof verifies"""
    import bitz_protocol as P
    from bitz_amd.device import Commitment
    from bitz_amd.mul import MulWitness
    n = 1 << 12
    rng = random.Random(12)
    x = [rng.getrandbits(64) for _ in range(n)]
    y = [rng.getrandbits(64) for _ in range(n)]
    w = MulWitness(ctx(), 64, _limb_tensor(64, x), _limb_tensor(64, y))
    t, s = w.layout.t, w.layout.s
    log(f"mul witness commit+open t={t} s={s}")
    assert (t, s) == (14, 6)
    rows = w.bitz_bit_rows()
    lo, hi = MW.products(64, x, y)
    host_rows = MW.bit_rows(MW.layout(64, n), x, y, lo, hi)
    assert np.array_equal(rows.cpu().numpy().view(np.uint64), host_rows)
    com = Commitment(ctx(), rows, t, s, 4, 2)
    pr = P.OracleProver(host_rows, t, s, 2, 4)
    assert com.root == pr.root
    claimed = com.evaluation()
    assert claimed == pr.evaluation()
    proof = com.prove(claimed)
    assert proof == pr.prove(claimed)
    P.verify_standalone(com.root, t, s, 2, 4, claimed, proof)
    com.close()
