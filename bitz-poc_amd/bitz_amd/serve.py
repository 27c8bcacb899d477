"""Batched and served commit+open jobs (bitz_server / bitz_batch_run of the
C ABI): each job is one commitment and its standalone opening, the same
calls as Pcs::commit + standalone_evaluation / prove_standalone
(src/bitz/pcs.rs:270-284, src/piop/spartan/protocol/bitz_opener.rs:643-707),
run on one of several per-thread contexts so that jobs overlap on the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class _JobStruct(C.Structure):
    _fields_ = [
        ("rows", C.c_void_p),
        ("t", C.c_uint32), ("s", C.c_uint32), ("log_batch", C.c_uint32), ("log_inv_rate", C.c_uint32),
        ("rows_on_device", C.c_uint32), ("have_claim", C.c_uint32),
        ("claimed", C.c_uint8 * 16),
        ("proof", C.c_void_p), ("proof_capacity", C.c_uint64),
        ("root", C.c_uint8 * 32),
        ("proof_len", C.c_uint64),
        ("status", C.c_int32), ("done", C.c_int32),
        ("submit_ms", C.c_double), ("start_ms", C.c_double), ("end_ms", C.c_double),
        ("error", C.c_char * 128),
        ("kind", C.c_uint32), ("verify", C.c_uint32),
        ("mul_x", C.c_void_p), ("mul_y", C.c_void_p), ("mul_n", C.c_uint64),
        ("verified", C.c_int32), ("reserved", C.c_int32),
        ("chain_blocks", C.c_void_p), ("chain_log_n", C.c_uint32), ("chain_digest", C.c_uint32 * 8),
        ("reserved2", C.c_uint32),
    ]

JOB_ROWS, JOB_MUL_U64, JOB_MUL_U64_PROOF, JOB_SHA256_CHAIN = 0, 1, 2, 3


class Job:
    """One commit+open job over host bit rows (numpy u64, column-major: 2^s
    columns of 2^(t-6) words). claimed=None: the job computes the honest
    claim (standalone_evaluation) first."""

    def __init__(self, rows: np.ndarray, t: int, s: int, log_batch: int, log_inv_rate: int,
                 claimed: int | None = None, proof_capacity: int = 1 << 21, verify: bool = False):
        rows = np.ascontiguousarray(rows, dtype=np.uint64)
        if rows.size != (1 << s) << (t - 6):
            raise ValueError("rows must hold 2^s * 2^(t-6) u64 words")
        self._rows = rows
        self._proof = (C.c_uint8 * proof_capacity)()
        j = _JobStruct()
        j.kind = JOB_ROWS
        j.verify = 1 if verify else 0
        j.rows = rows.ctypes.data
        j.t, j.s, j.log_batch, j.log_inv_rate = t, s, log_batch, log_inv_rate
        j.have_claim = 0 if claimed is None else 1
        if claimed is not None:
            C.memmove(j.claimed, int(claimed).to_bytes(16, "little"), 16)
        j.proof = C.cast(self._proof, C.c_void_p)
        j.proof_capacity = proof_capacity
        self.s_ = j

    @property
    def footprint(self) -> int:
        j = self.s_
        return int(_lib.load().bitz_job_footprint(j.t, j.s, j.log_batch, j.log_inv_rate))

    def check(self) -> None:
        if self.s_.status != 0:
            raise _lib.BitzError(self.s_.status, "bitz_job", self.s_.error.decode(errors="replace"))

    @property
    def root(self) -> bytes:
        return bytes(self.s_.root)

    @property
    def claimed(self) -> int:
        return int.from_bytes(bytes(self.s_.claimed), "little")

    @property
    def proof(self) -> bytes:
        self.check()
        return C.string_at(self._proof, self.s_.proof_len)

    @property
    def latency_ms(self) -> float:
        return self.s_.end_ms - self.s_.submit_ms

    @property
    def t(self) -> int:
        return int(self.s_.t)

    @property
    def s(self) -> int:
        return int(self.s_.s)

    @property
    def verified(self) -> bool:
        return bool(self.s_.verified)


class MulJob(Job):
    """A commit+open job over the witness grid of n u64 multiplications
    x[i] * y[i] (MulWitness::bitz_bit_rows, src/piop/spartan/mul.rs:268-330),
    packed on the device by the worker; t and s follow MulLayout."""

    def __init__(self, x: np.ndarray, y: np.ndarray, log_batch: int, log_inv_rate: int,
                 claimed: int | None = None, proof_capacity: int = 1 << 21, verify: bool = False):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        y = np.ascontiguousarray(y, dtype=np.uint64)
        if x.shape != y.shape or x.ndim != 1 or x.size == 0:
            raise ValueError("x and y must be equal-length non-empty u64 vectors")
        self._x, self._y = x, y
        self._proof = (C.c_uint8 * proof_capacity)()
        j = _JobStruct()
        j.kind = JOB_MUL_U64
        j.verify = 1 if verify else 0
        j.mul_x, j.mul_y, j.mul_n = x.ctypes.data, y.ctypes.data, x.size
        out = (C.c_uint32 * 4)()
        _lib.call("bitz_mul_layout", 64, x.size, 0, C.cast(out, C.c_void_p))
        j.t, j.s = out[2], out[1]
        j.log_batch, j.log_inv_rate = log_batch, log_inv_rate
        j.have_claim = 0 if claimed is None else 1
        if claimed is not None:
            C.memmove(j.claimed, int(claimed).to_bytes(16, "little"), 16)
        j.proof = C.cast(self._proof, C.c_void_p)
        j.proof_capacity = proof_capacity
        self.s_ = j


class MulProofJob(MulJob):
    """The full u64 multiplication proof of x[i] * y[i] (bitz_prove_mul_u64:
    commit, Spartan PIOP, BitZ opening; protocol::prove for MulLayout<u64>,
    src/piop/spartan/protocol/mod.rs:1138-1170), verified by its worker
    (bitz_verify_mul_u64) when verify is set. `claimed` is the first half of
    the outer transcript state digest."""

    def __init__(self, x: np.ndarray, y: np.ndarray, proof_capacity: int = 1 << 21, verify: bool = False):
        super().__init__(x, y, 4, 1, None, proof_capacity, verify)
        self.s_.kind = JOB_MUL_U64_PROOF


class ChainJob(Job):
    """The SHA-256 chain proof of 2^k blocks (16 u32 words each) from FIPS
    180-4's initial state (bitz_prove_sha256_chain; prove_sha256_chain,
    src/piop/spartan/sha256/chain.rs:817-845), verified by its worker
    (bitz_verify_sha256_chain) when verify is set. `claimed` is the first
    half of the outer transcript state digest, `digest` the final chaining
    value."""

    def __init__(self, blocks: np.ndarray, proof_capacity: int = 1 << 21, verify: bool = False):
        b = np.ascontiguousarray(np.asarray(blocks, dtype=np.uint32))
        n = b.shape[0]
        k = n.bit_length() - 1
        if b.ndim != 2 or b.shape[1] != 16 or n != 1 << k or not 7 <= k <= 16:
            raise ValueError("blocks must be (2^k, 16) u32 words with 7 <= k <= 16")
        self._blocks = b
        self._proof = (C.c_uint8 * proof_capacity)()
        j = _JobStruct()
        j.kind = JOB_SHA256_CHAIN
        j.verify = 1 if verify else 0
        j.chain_blocks = b.ctypes.data
        j.chain_log_n = k
        j.t = min(k, 13)
        j.s = 15 + k - j.t
        j.log_batch, j.log_inv_rate = 4, 1
        j.proof = C.cast(self._proof, C.c_void_p)
        j.proof_capacity = proof_capacity
        self.s_ = j

    @property
    def digest(self) -> list[int]:
        return [int(x) for x in self.s_.chain_digest]


class Server:
    """bitz_server: `workers` contexts on one device; jobs admitted in order
    while the footprint of the jobs in flight fits hbm_budget bytes."""

    def __init__(self, device: int = 0, workers: int = 4, hbm_budget: int = 200 << 30):
        self._p = C.c_void_p()
        _lib.call("bitz_server_create", device, workers, hbm_budget, C.byref(self._p))
        self._jobs: list[Job] = []

    def submit(self, job: Job) -> None:
        self._jobs.append(job)  # the struct must outlive the job
        _lib.call("bitz_server_submit", self._p, C.byref(job.s_))

    def done(self, job: Job) -> bool:
        return _lib.load().bitz_server_done(self._p, C.byref(job.s_)) == 1

    def wait_all(self) -> None:
        _lib.call("bitz_server_wait_all", self._p)

    def clock_ms(self) -> float:
        return float(_lib.load().bitz_server_clock_ms(self._p))

    def close(self) -> None:
        if self._p:
            _lib.load().bitz_server_destroy(self._p)
            self._p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_batch(jobs: list[Job], workers: int, hbm_budget: int = 200 << 30, device: int = 0) -> None:
    """bitz_batch_run: all jobs, largest first, on a fresh server."""
    arr = (_JobStruct * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = j.s_
    rc = _lib.load().bitz_batch_run(device, workers, hbm_budget, C.cast(arr, C.c_void_p), len(jobs))
    for i, j in enumerate(jobs):
        j.s_ = arr[i]
    if rc != 0:
        bad = [j for j in jobs if j.s_.status != 0]
        msg = bad[0].s_.error.decode(errors="replace") if bad else _lib.load().bitz_last_error().decode()
        raise _lib.BitzError(rc, "bitz_batch_run", msg)
