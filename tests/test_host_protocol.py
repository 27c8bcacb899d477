"""CPU checks of the product library's host-side protocol pieces: its
incremental BLAKE3 (the transcript hash) against the pinned oracle BLAKE3,
and its Ligerito ladder solver against the CPU restatement and against the
reference's own embedded slim ladder for m = 26
(vendor/flock-mod/crates/flock-core/configs/ligerito/m26_slim.toml: L0 queries
90, fold grinding 14; L1 queries 60, fold grinding 10; L2 queries 45, fold
grinding 7)."""
import ctypes as C

import numpy as np
import pytest

import bitz_amd._lib as L
import bitz_protocol as P
import oracle as O


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1023, 1024, 1025, 2048, 3000, 8193, 70000])
def test_host_blake3_matches_oracle(n):
    data = bytes((i * 7 + 3) % 251 for i in range(n))
    buf = (C.c_uint8 * max(n, 1)).from_buffer_copy(data or b"\\0")
    out = (C.c_uint8 * 32)()
    L.load().bitz_host_blake3(C.cast(buf, C.c_void_p), n, C.cast(out, C.c_void_p))
    assert bytes(out) == O.blake3(data)


@pytest.mark.parametrize("n", [65 * 1024, 65 * 1024 + 1, 200 * 1024 + 7, (1 << 20) + 5])
@pytest.mark.parametrize("split", [0, 1, 100, 1023, 1024, 1025, 5000])
def test_host_blake3_bulk_updates(n, split):
    """long updates take the multi-threaded whole-chunk path; a first update
    leaves the hasher at every kind of chunk offset"""
    rng = np.random.default_rng(n + split)
    data = rng.integers(0, 256, size=n, dtype=np.uint8).tobytes()
    buf = (C.c_uint8 * n).from_buffer_copy(data)
    out = (C.c_uint8 * 32)()
    L.load().bitz_host_blake3_parts(C.cast(buf, C.c_void_p), n, split, C.cast(out, C.c_void_p))
    want = O.blake3(data)
    assert bytes(out) == want
    L.load().bitz_host_blake3(C.cast(buf, C.c_void_p), n, C.cast(out, C.c_void_p))
    assert bytes(out) == want


def _lib_config(m, r, k):
    out = (C.c_int32 * 256)()
    n = L.load().bitz_ligerito_config(m, r, k, 100, C.cast(out, C.c_void_p), 256)
    assert n > 0, L.load().bitz_last_error()
    v = list(out[:n])
    levels = [dict(zip(["rate", "msg_cols", "lanes", "k", "queries", "grinding", "fold_grinding", "ood"],
                       v[6 + 8 * i:14 + 8 * i])) for i in range(v[5])]
    return dict(m=v[0], log_n=v[1], initial_k=v[2], yr_log_n=v[3], round0_grinding=v[4], levels=levels)


@pytest.mark.parametrize("m,r,k", [(20, 2, 4), (22, 1, 4), (26, 2, 4), (28, 2, 4), (30, 2, 4), (26, 3, 4),
                                   (35, 1, 4)])
def test_ladder_matches_restatement(m, r, k):
    a = _lib_config(m, r, k)
    b = P.custom_johnson_config(m, r, k)
    for key in ("m", "log_n", "initial_k", "yr_log_n", "round0_grinding"):
        assert a[key] == b[key], key
    keys = ("rate", "msg_cols", "lanes", "k", "queries", "grinding", "fold_grinding", "ood")
    assert a["levels"] == [{k: lv[k] for k in keys} for lv in b["levels"]]


def test_ladder_matches_reference_slim_toml_m26():
    cfg = P.custom_johnson_config(26, 2, 4)
    lv = cfg["levels"]
    assert [(x["rate"], x["msg_cols"], x["lanes"]) for x in lv[:3]] == [(2, 15, 4), (3, 12, 3), (4, 9, 3)]
    assert [x["queries"] for x in lv[:3]] == [90, 60, 45]
    assert [x["fold_grinding"] for x in lv[:3]] == [14, 10, 7]
    assert cfg["yr_log_n"] == 3 and len(lv) == 5


def test_ladder_rejections():
    with pytest.raises(ValueError):
        P.custom_johnson_config(19, 2, 4)
    out = (C.c_int32 * 256)()
    assert L.load().bitz_ligerito_config(19, 2, 4, 100, C.cast(out, C.c_void_p), 256) == -1
