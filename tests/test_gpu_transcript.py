"""The device copy of the Fiat-Shamir transcript (used by the GKR rounds)
against the oracle's Blake3Transcript restatement: a host transcript absorbs
a prefix, the device continues with message absorbs and 16-byte challenges.
Prefix lengths put the hasher at every kind of chunk offset and stack depth;
the message runs cross many chunk boundaries."""
import ctypes as C

import numpy as np
import pytest

import bitz_protocol as P
from gpu_util import ctx, log

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prefix_len", [0, 1, 54, 1013, 1014, 5000, 70000])
@pytest.mark.parametrize("msg_len,n", [(32, 300), (16, 70), (48, 5), (0, 3), (1500, 4)])
def test_device_transcript_matches_oracle(prefix_len, msg_len, n):
    from bitz_amd import _lib
    rng = np.random.default_rng(prefix_len * 7 + msg_len)
    prefix = rng.integers(0, 256, size=prefix_len, dtype=np.uint8).tobytes()
    msgs = rng.integers(0, 256, size=n * msg_len, dtype=np.uint8).tobytes()
    log(f"device transcript prefix={prefix_len} msg_len={msg_len} n={n}")
    out = (C.c_uint8 * (16 * n))()
    pbuf = (C.c_uint8 * max(1, prefix_len)).from_buffer_copy(prefix or b"\0")
    mbuf = (C.c_uint8 * max(1, len(msgs))).from_buffer_copy(msgs or b"\0")
    _lib.call("bitz_transcript_challenges_dev", ctx().ptr, C.cast(pbuf, C.c_void_p), prefix_len,
              C.cast(mbuf, C.c_void_p), n, msg_len, C.cast(out, C.c_void_p))
    tr = P.Transcript()
    tr.absorb_slice(prefix)
    want = b""
    for i in range(n):
        tr.absorb_slice(msgs[i * msg_len:(i + 1) * msg_len])
        want += tr.challenge(16)
    assert bytes(out) == want
