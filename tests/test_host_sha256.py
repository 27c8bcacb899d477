"""The library's host SHA-256 compression (bitz-poc_amd/csrc/sha256_host.cpp:
the x86 SHA extensions where the CPU has them, portable code otherwise),
which computes the chain's chaining values, against hashlib: a chain of
compressions from H_0 over random blocks followed by the padding block's is
SHA-256 of the message, for each path this CPU has, and the paths agree on
random states and blocks."""
import ctypes as C
import hashlib
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
IV = [0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19]


@pytest.fixture(scope="module")
def sh(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("sha256") / "libsha256.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so,
                    os.path.join(HERE, "sha256_harness.cpp")], check=True)
    lib = C.CDLL(so)
    lib.sh_has_shani.restype = C.c_int
    lib.sh_compress.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def _paths(sh):
    return [0, 1, 2] if sh.sh_has_shani() else [0, 2]


def _compress(sh, path, state, block):
    st = np.ascontiguousarray(state, dtype=np.uint32)
    bl = np.ascontiguousarray(block, dtype=np.uint32)
    out = np.zeros(8, np.uint32)
    sh.sh_compress(path, st.ctypes.data, bl.ctypes.data, out.ctypes.data)
    return out


@pytest.mark.parametrize("nblocks", [1, 3, 17])
def test_chain_then_padding_is_sha256(sh, nblocks):
    rng = np.random.default_rng(nblocks)
    msg = rng.integers(0, 256, size=64 * nblocks, dtype=np.uint8).tobytes()
    pad = b"\x80" + bytes(55) + (8 * len(msg)).to_bytes(8, "big")
    want = np.frombuffer(hashlib.sha256(msg).digest(), ">u4").astype(np.uint32)
    for path in _paths(sh):
        st = np.array(IV, np.uint32)
        for i in range(nblocks):
            st = _compress(sh, path, st, np.frombuffer(msg[64 * i: 64 * i + 64], ">u4"))
        st = _compress(sh, path, st, np.frombuffer(pad, ">u4"))
        assert (st == want).all(), path


def test_paths_agree_on_random_states(sh):
    rng = np.random.default_rng(7)
    for _ in range(200):
        st = rng.integers(0, 2**32, size=8, dtype=np.uint64).astype(np.uint32)
        bl = rng.integers(0, 2**32, size=16, dtype=np.uint64).astype(np.uint32)
        outs = [_compress(sh, p, st, bl) for p in _paths(sh)]
        for o in outs[1:]:
            assert (o == outs[0]).all()
