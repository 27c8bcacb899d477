"""Launch-shape choices change no result: the proof bytes under non-default
round-kernel shapes, host-tail thresholds and NTT radices equal the default
context's (which the oracle parity tests pin), and bitz_ctx_autotune leaves
valid choices that prove the same bytes. Bad settings are rejected."""
import numpy as np
import pytest

from gpu_util import log

pytestmark = pytest.mark.gpu

T, S, RATE, K0 = 15, 11, 2, 4  # 2^20 u64 coefficients


def _rows(seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**64, size=(1 << S) << (T - 6), dtype=np.uint64)


def _prove(c, rows):
    from bitz_amd.device import Commitment
    com = Commitment(c, rows, T, S, K0, RATE)
    claimed = com.evaluation()
    proof = com.prove(claimed)
    root = com.root
    com.close()
    return root, claimed, proof


@pytest.fixture(scope="module")
def reference_proof():
    from bitz_amd.device import Context
    rows = _rows(77)
    c = Context(0)
    log("tune: default-context proof t=15 s=11")
    return rows, _prove(c, rows)


TUNES = [
    dict(round_r_max=1, round_min_chunks=1, round_block_cap=64, host_tail_log=4, ntt_radix_log=1),
    dict(round_r_max=8, round_min_chunks=4096, round_block_cap=2048, host_tail_log=12, ntt_radix_log=3),
    dict(round_r_max=32, round_min_chunks=1, round_block_cap=65536, host_tail_log=14, ntt_radix_log=2),
]


@pytest.mark.parametrize("device_fs", [1, 0])
@pytest.mark.parametrize("tune", TUNES, ids=["small", "mid", "wide"])
def test_proof_bytes_independent_of_tune(reference_proof, tune, device_fs):
    from bitz_amd import _lib
    from bitz_amd.device import Context, get_tune, set_tune
    rows, want = reference_proof
    log(f"tune: {tune} device_fs={device_fs}")
    c = Context(0)
    _lib.call("bitz_ctx_set_device_transcript", c.ptr, device_fs)
    set_tune(c, **tune)
    got = get_tune(c)
    assert {k: got[k] for k in tune} == tune and not got["tuned"]
    assert _prove(c, rows) == want


def test_set_tune_rejects_bad_values():
    from bitz_amd.device import Context, set_tune
    c = Context(0)
    for bad in (dict(round_r_max=3), dict(round_r_max=64), dict(round_block_cap=8), dict(host_tail_log=2),
                dict(host_tail_log=15), dict(ntt_radix_log=0), dict(ntt_radix_log=5), dict(round_min_chunks=0)):
        with pytest.raises(RuntimeError):
            set_tune(c, **bad)


def test_autotune_choices_prove_same_bytes(reference_proof):
    from bitz_amd.device import Context, autotune
    rows, want = reference_proof
    c = Context(0)
    log("tune: autotune t=15 s=11")
    tn = autotune(c, T, S, K0, RATE)
    log(f"tune: chose {tn}")
    assert tn["tuned"]
    assert tn["round_r_max"] in (8, 16, 32)
    assert tn["round_min_chunks"] in (1024, 2048, 4096)
    assert tn["round_block_cap"] in (2048, 4096, 8192)
    assert 6 <= tn["host_tail_log"] <= 13
    assert tn["ntt_radix_log"] in (3, 4)
    assert 0 < tn["tune_ms"] < 60000
    assert _prove(c, rows) == want
