"""Helpers shared by the GPU parity tests: seeded inputs and host<->device
conversion of Gf128 arrays (numpy uint64 (n, 2) <-> torch int64 (n, 2))."""
import functools

import numpy as np


def rand_gf(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**64, size=(n, 2), dtype=np.uint64)


def to_dev(a: np.ndarray):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to("cuda")


def to_host(t) -> np.ndarray:
    return t.cpu().numpy().view(np.uint64)


@functools.lru_cache(maxsize=1)
def ctx():
    from bitz_amd.device import Context
    return Context(0)


def log(msg: str) -> None:
    print(msg, flush=True)
