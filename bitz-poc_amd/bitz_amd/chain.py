"""The SHA-256 chain proof on the device: the mirror of
prove_sha256_chain (src/piop/spartan/sha256/chain.rs:817-845).

N = 2^k blocks M_0..M_{N-1} (16 u32 words each, the circuit's word order)
are chained from an initial state H_0 (FIPS 180-4's by default):
H_{i+1} = Compress(H_i, M_i). The library computes the chain, the witness,
the commitment of the source bits, the linear PIOP over a runtime prime and
the virtual BitZ opening."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def prove_chain(ctx, blocks, initial_state=None):
    """blocks: (2^k, 16) uint32 array-like. Returns (root, digest (8 u32),
    outer transcript state digest, proof bytes)."""
    b = np.ascontiguousarray(np.asarray(blocks, dtype=np.uint32))
    if b.ndim != 2 or b.shape[1] != 16:
        raise ValueError("blocks must be (n, 16) u32 words")
    n = b.shape[0]
    k = n.bit_length() - 1
    if n != 1 << k:
        raise ValueError("the chain needs a power-of-two number of blocks")
    iv = None
    if initial_state is not None:
        iv = np.ascontiguousarray(np.asarray(initial_state, dtype=np.uint32))
        if iv.shape != (8,):
            raise ValueError("initial_state must be 8 u32 words")
    root = (C.c_uint8 * 32)()
    digest = (C.c_uint32 * 8)()
    state = (C.c_uint8 * 32)()
    plen = C.c_uint64(0)
    cap = 8 << 20
    while True:
        buf = (C.c_uint8 * cap)()
        rc = _lib.load().bitz_prove_sha256_chain(
            ctx.ptr, C.c_void_p(b.ctypes.data), k, C.c_void_p(iv.ctypes.data) if iv is not None else None, root,
            digest, state, buf, cap, C.byref(plen))
        if rc == _lib.BITZ_ERR_ARG and plen.value > cap:
            cap = plen.value
            continue
        _lib.check(rc, "bitz_prove_sha256_chain")
        return bytes(root), [int(x) for x in digest], bytes(state), bytes(buf)[:plen.value]


def verify_chain(ctx, blocks, digest, root: bytes, proof: bytes, initial_state=None) -> bytes:
    """verify_sha256_chain (chain.rs:850-872): the statement (blocks, digest,
    initial state), the root and the proof; raises BitzError when rejected,
    returns the verifier's outer transcript state digest."""
    b = np.ascontiguousarray(np.asarray(blocks, dtype=np.uint32))
    n = b.shape[0]
    k = n.bit_length() - 1
    if b.ndim != 2 or b.shape[1] != 16 or n != 1 << k:
        raise ValueError("blocks must be (2^k, 16) u32 words")
    d = np.ascontiguousarray(np.asarray(digest, dtype=np.uint32))
    iv = None if initial_state is None else np.ascontiguousarray(np.asarray(initial_state, dtype=np.uint32))
    state = (C.c_uint8 * 32)()
    pb = (C.c_uint8 * max(len(proof), 1)).from_buffer_copy(proof or b"\0")
    _lib.call("bitz_verify_sha256_chain", ctx.ptr, C.c_void_p(b.ctypes.data), k,
              C.c_void_p(iv.ctypes.data) if iv is not None else None, C.c_void_p(d.ctypes.data), bytes(root), pb,
              len(proof), state)
    return bytes(state)
