"""ctypes binding of libbitz_hip.so (the C ABI declared in include/bitz_hip.h).

The library is the product: there is no CPU fallback. If it is missing or
fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libbitz_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG_DIR), "include", "bitz_hip.h")

BITZ_OK = 0
BITZ_ERR_ARG = -4
BITZ_ERR_VERIFY = -5
ERR_NAMES = {-1: "SHAPE", -2: "ALLOC", -3: "DEVICE", -4: "ARG", -5: "VERIFY"}


class BitzError(RuntimeError):
    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} failed ({ERR_NAMES.get(code, code)}): {msg}")
        self.code = code


_P = C.c_void_p
_I = C.c_int
_U32 = C.c_uint32
_U64 = C.c_uint64

# name -> (restype, argtypes); every symbol of include/bitz_hip.h
SIGNATURES = {
    "bitz_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "bitz_ctx_destroy": (None, [_P]),
    "bitz_ctx_set_stream": (_I, [_P, _P]),
    "bitz_ctx_stream": (_P, [_P]),
    "bitz_ctx_synchronize": (_I, [_P]),
    "bitz_ctx_trim": (_I, [_P]),
    "bitz_ctx_set_device_transcript": (_I, [_P, _I]),
    "bitz_transcript_challenges_dev": (_I, [_P, _P, _U64, _P, C.c_uint32, C.c_uint32, _P]),
    "bitz_dev_alloc": (_I, [_P, _U64, _P]),
    "bitz_dev_free": (_I, [_P, _P]),
    "bitz_dev_upload": (_I, [_P, _P, _P, _U64]),
    "bitz_dev_download": (_I, [_P, _P, _P, _U64]),
    "bitz_ctx_cached_bytes": (_U64, [_P]),
    "bitz_last_error": (C.c_char_p, []),
    "bitz_version": (C.c_char_p, []),
    "bitz_ctx_set_timing": (_I, [_P, _I]),
    "bitz_ctx_set_timing_only": (_I, [_P, C.c_char_p]),
    "bitz_ctx_kernel_stats": (_I, [_P, _P, _U64, _P, _U64]),
    "bitz_ctx_set_tune": (_I, [_P, _U32, _U32, _U32, _U32, _U32]),
    "bitz_ctx_get_tune": (_I, [_P, _P]),
    "bitz_ctx_autotune": (_I, [_P, _U32, _U32, _U32, _U32, _P]),
    "bitz_mul_layout": (_I, [_U32, _U64, C.c_int32, _P]),
    "bitz_mul_bit_rows_dev": (_I, [_P, _U32, _U64, C.c_int32, _P, _P, _P, _P, _P, _U64]),
    "bitz_commit_batch": (_I, [_P, _U32, _P]),
    "bitz_prove_mul_u64": (_I, [_P, _P, _P, _P, _P, _U64, C.c_int32, _P, _P, _P, _U64, _P]),
    "bitz_verify_mul_u64": (_I, [_P, _P, _U64, C.c_int32, _P, _U64, _P]),
    "bitz_prove_sha256_chain": (_I, [_P, _P, _U32, _P, _P, _P, _P, _P, _U64, _P]),
    "bitz_verify_sha256_chain": (_I, [_P, _P, _U32, _P, _P, _P, _P, _U64, _P]),
    "bitz_gf_mul_dev": (_I, [_P, _P, _P, _P, _U64]),
    "bitz_gf_mul_const_dev": (_I, [_P, _P, _P, _P, _U64]),
    "bitz_blake3_leaves_dev": (_I, [_P, _P, _U64, _U64, _P]),
    "bitz_ntt_forward_dev": (_I, [_P, _P, _I, _I, _I, _I]),
    "bitz_merkle_tree_dev": (_I, [_P, _P, _U64, _U64, _P]),
    "bitz_commit_dev": (_I, [_P, _P, _I, _I, _I, _P, _P, _P]),
    "bitz_eq_table_dev": (_I, [_P, _P, _I, _P]),
    "bitz_fold_columns_dev": (_I, [_P, _P, _I, _I, _P, _P]),
    "bitz_pow_images_dev": (_I, [_P, _P, _P, _U64, _P]),
    "bitz_pack_columns_dev": (_I, [_P, _P, _I, _I, _P]),
    "bitz_gkr_round_sums_dev": (_I, [_P, _P, _P, _U64, _P, _I, _P, _I, _P]),
    "bitz_gkr_leaf_round_sums_dev": (_I, [_P, _P, _P, _I, _I, _P, _I, _P, _I, _P]),
    "bitz_fold_halves_dev": (_I, [_P, _P, _U64, _P]),
    "bitz_gkr_fold_round_sums_dev": (_I, [_P, _P, _P, _U64, _P, _P, _I, _P, _I, _P]),
    "bitz_gkr_leaf_round1_sums_dev": (_I, [_P, _P, _P, _I, _I, _P, _P, _P, _I, _P]),
    "bitz_gkr_leaf_fold2_dev": (_I, [_P, _P, _P, _I, _I, _P, _P, _P, _P, _P]),
    "bitz_gkr_premul_round_sums_dev": (_I, [_P, _P, _P, _U64, _P, _I, _P, _I, _P]),
    "bitz_gkr_round_sums_rowonly_dev": (_I, [_P, _P, _P, _U64, _I, _P, _I, _P]),
    "bitz_tensor_table_dev": (_I, [_P, _P, _P, _I, _P]),
    "bitz_gkr_level2_dev": (_I, [_P, _P, _P, _I, _I, _P]),
    "bitz_gkr_level3_dev": (_I, [_P, _P, _P, _I, _I, _P]),
    "bitz_gkr_level1_round0_sums_dev": (_I, [_P, _P, _P, _I, _I, _P, _P, _I, _P]),
    "bitz_gkr_level1_fold_dev": (_I, [_P, _P, _P, _I, _I, _P, _P, _P, _P]),
    "bitz_gkr_leaf_fold_dev": (_I, [_P, _P, _P, _I, _I, _P, _P, _P]),
    "bitz_gkr_level1_dev": (_I, [_P, _P, _P, _I, _I, _P]),
    "bitz_product_level_dev": (_I, [_P, _P, _U64, _P]),
    "bitz_xi_combined_dev": (_I, [_P, _P, _I, _I, _P, _P]),
    "bitz_fold_rows_eq_dev": (_I, [_P, _P, _I, _I, _P, _P]),
    "bitz_round_poly_dev": (_I, [_P, _P, _P, _U64, _P]),
    "bitz_fold_lsb_dev": (_I, [_P, _P, _U64, _P, _P]),
    "bitz_fold_1b_rows_dev": (_I, [_P, _P, _P, _U64, _P]),
    "bitz_gkr_nibble_rows_dev": (_I, [_P, _P, _I, _I, _P]),
    "bitz_gkr_nibble_round_sums_dev": (_I, [_P, _I, _P, _P, _I, _I, _P, _P, _P, _I, _P]),
    "bitz_gkr_nibble_fold_dev": (_I, [_P, _I, _P, _P, _I, _I, _P, _P, _P, _P]),
    "bitz_gkr_jit_round_sums_dev": (_I, [_P, _I, _P, _P, _I, _I, _P, _P, _P, _I, _P]),
    "bitz_gkr_jit_fold_dev": (_I, [_P, _I, _P, _P, _I, _I, _P, _P, _P, _P]),
    "bitz_gkr_level4_dev": (_I, [_P, _P, _P, _I, _I, _P]),
    "bitz_fold_1b_rows_noalloc": (_I, [_P, _P, _P, _U64, _P]),
    "bitz_fold_b128_dev": (_I, [_P, _P, _U64, _P, _P]),
    "bitz_axpy_dev": (_I, [_P, _P, _P, _U64, _P]),
    "bitz_lig_msg_dev": (_I, [_P, _P, _P, _U64, _I, _P]),
    "bitz_lig_fold_dev": (_I, [_P, _P, _P, _U64, _P, _P, _P]),
    "bitz_induce_basis_dev": (_I, [_P, _P, _U32, _I, _P, _P]),
    "bitz_pow_find_dev": (_I, [_P, _P, _P, _U32, C.POINTER(_U64)]),
    "bitz_commit": (_I, [_P, _P, _U32, _U32, _U32, _U32, _P, C.POINTER(_P)]),
    "bitz_commit_device_rows": (_I, [_P, _P, _U32, _U32, _U32, _U32, _P, C.POINTER(_P)]),
    "bitz_commit_free": (None, [_P]),
    "bitz_commit_packed": (_P, [_P]),
    "bitz_commit_codeword": (_P, [_P]),
    "bitz_commit_tree": (_P, [_P]),
    "bitz_commit_read_tree": (_I, [_P, _P, _U64]),
    "bitz_standalone_evaluation": (_I, [_P, _P, _P]),
    "bitz_prove_standalone": (_I, [_P, _P, _P, _P, _U64, C.POINTER(_U64)]),
    "bitz_verify_standalone": (_I, [_P, _P, _U32, _U32, _U32, _U32, _P, _P, _U64]),
    "bitz_ligerito_config": (_I, [_U32, _U32, _U32, _U32, _P, _U32]),
    "bitz_host_blake3": (None, [_P, _U64, _P]),
    "bitz_job_footprint": (_U64, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "bitz_server_create": (_I, [_I, C.c_uint32, _U64, _P]),
    "bitz_server_submit": (_I, [_P, _P]),
    "bitz_server_done": (_I, [_P, _P]),
    "bitz_server_wait_all": (_I, [_P]),
    "bitz_server_clock_ms": (C.c_double, [_P]),
    "bitz_server_destroy": (None, [_P]),
    "bitz_batch_run": (_I, [_I, C.c_uint32, _U64, _P, C.c_uint32]),
    "bitz_host_blake3_parts": (None, [_P, _U64, _U64, _P]),
}

_lib = None


def load() -> C.CDLL:
    """Load libbitz_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libbitz_hip.so not found at {LIB_PATH}; build it with "
            f"`make -C {_PKG_DIR}` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, fn: str) -> None:
    if rc != BITZ_OK:
        msg = load().bitz_last_error()
        raise BitzError(rc, fn, msg.decode() if msg else "")


def call(name: str, *args) -> int:
    lib = load()
    rc = getattr(lib, name)(*args)
    check(rc, name)
    return rc
