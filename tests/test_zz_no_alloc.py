"""The entry point that allocates nothing (bitz_fold_1b_rows_noalloc: the
ring-switch bit fold with its block partials in the library's fixed device
array) against bitz_fold_1b_rows_dev, whose partials come from the context's
pool, and against the oracle: bit-exact on two calls in a row, the second
smaller, so the second call runs over the first call's leftover partials."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gpu_util import ctx, log, rand_gf, to_dev, to_host

pytestmark = pytest.mark.gpu


def _fold(name, p, suffix):
    import torch
    from bitz_amd import _lib
    out = torch.zeros((128, 2), dtype=torch.int64, device="cuda")
    _lib.call(name, ctx().ptr, C.c_void_p(p.data_ptr()), C.c_void_p(suffix.data_ptr()), p.shape[0],
              C.c_void_p(out.data_ptr()))
    _lib.call("bitz_ctx_synchronize", ctx().ptr)
    return to_host(out)


@pytest.mark.parametrize("sizes", [(1 << 20, 3000), (1 << 16, 1), (4097, 2048)])
def test_noalloc_equals_pool_entry_point(sizes):
    for i, n in enumerate(sizes):
        log(f"fold_1b_rows noalloc call {i} n={n}")
        p, suffix = rand_gf(n, 10 + n), rand_gf(n, 20 + n)
        dp, ds = to_dev(p), to_dev(suffix)
        got = _fold("bitz_fold_1b_rows_noalloc", dp, ds)
        want = _fold("bitz_fold_1b_rows_dev", dp, ds)
        assert np.array_equal(got, want)
        if n <= 1 << 16:
            assert np.array_equal(got, O.fold_1b_rows(p, suffix))
