"""The host transcript's long-message path (bitz-poc_amd/csrc/host_proto.cpp,
Transcript::absorb_slice_runs): a message absorbed as its lead bytes, the
chaining values of aligned power-of-two runs of its whole chunks, and its tail
gives the same transcript as absorbing the bytes (absorb_slice, whose hash is
pinned against the oracle BLAKE3 by test_host_protocol.py), at every offset
of the message within a chunk. The prover hashes the runs on the device."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def hp(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("hostproto") / "libhostproto.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-pthread", "-o", so,
                    os.path.join(HERE, "host_proto_harness.cpp")], check=True)
    lib = C.CDLL(so)
    lib.hp_transcript.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_int, C.c_void_p]
    lib.hp_runs.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_int), C.c_int]
    lib.hp_runs.restype = C.c_int
    return lib


def _run(hp, pre, msg, runs):
    out = (C.c_uint8 * 32)()
    pb = (C.c_uint8 * max(len(pre), 1)).from_buffer_copy(pre or b"\0")
    mb = (C.c_uint8 * max(len(msg), 1)).from_buffer_copy(msg or b"\0")
    hp.hp_transcript(C.cast(pb, C.c_void_p), len(pre), C.cast(mb, C.c_void_p), len(msg), runs,
                     C.cast(out, C.c_void_p))
    return bytes(out)


@pytest.mark.parametrize("pre_len", [0, 1, 6, 500, 1014, 1015, 1016, 3000, 5 * 1024 - 9])
@pytest.mark.parametrize("msg_len", [0, 100, 1024, 1033, 5000, 64 * 1024 + 77, 300 * 1024 + 5])
def test_runs_equal_bytes(hp, pre_len, msg_len):
    rng = np.random.default_rng(pre_len * 7 + msg_len)
    pre = rng.integers(0, 256, pre_len, dtype=np.uint8).tobytes()
    msg = rng.integers(0, 256, msg_len, dtype=np.uint8).tobytes()
    assert _run(hp, pre, msg, 1) == _run(hp, pre, msg, 0)


@pytest.mark.parametrize("c0,m", [(0, 1), (0, 8), (3, 13), (5, 1000), (1024, 1024), (7, 0)])
def test_runs_tile_aligned(hp, c0, m):
    ks = (C.c_int * 128)()
    n = hp.hp_runs(c0, m, ks, 128)
    c = c0
    for k in ks[:n]:
        assert c % (1 << k) == 0
        c += 1 << k
    assert c == c0 + m
    assert n <= 2 * max(m, 1).bit_length()


def test_compress_paths_agree(hp):
    """the SSE4.1 row form of the BLAKE3 compression (used where the CPU has
    it) equals the portable one on random inputs and every flag combination;
    the portable form's digests are pinned against the oracle BLAKE3
    (test_host_protocol.py)"""
    hp.hp_has_sse41.restype = C.c_int
    hp.hp_compress.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p]
    paths = [0, 1, 2] if hp.hp_has_sse41() else [0, 2]
    rng = np.random.default_rng(3)
    for i in range(300):
        cv = rng.integers(0, 2**32, size=8, dtype=np.uint64).astype(np.uint32)
        block = rng.integers(0, 256, size=64, dtype=np.uint8)
        counter = int(rng.integers(0, 2**63))
        block_len = int(rng.integers(0, 65))
        flags = i % 16
        outs = []
        for p in paths:
            out = np.zeros(16, np.uint32)
            hp.hp_compress(p, cv.ctypes.data, block.ctypes.data, counter, block_len, flags, out.ctypes.data)
            outs.append(out)
        for o in outs[1:]:
            assert (o == outs[0]).all()
