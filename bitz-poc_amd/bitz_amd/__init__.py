"""bitz_amd -- MI355X (gfx950) BitZ prover hot path.

The compute lives in libbitz_hip.so (HIP kernels behind the C ABI of
include/bitz_hip.h); this package is its Python binding and the host-side
mirror of the reference's commitment interface.
"""
from ._lib import BitzError, load  # noqa: F401

__all__ = ["BitzError", "load"]
