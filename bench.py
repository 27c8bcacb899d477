"""BitZ PCS commit+open throughput on one MI355X.

A step is one pass of the hot path over one batch of synthetic input: the
BitZ PCS commit of 2^24 uniform u64 coefficients (n = 30 committed bits,
t = 17 row variables, s = 13 column variables, rate 1/4 Reed-Solomon code,
16 interleaved lanes, BLAKE3 Merkle tree) followed by the standalone opening
(statement, Round 0, fold round, batched grand-product GKR, inner-product
sumcheck, ring switch, recursive Ligerito opening) at a transcript-derived
point. The coefficients are already resident in HBM when the timed region
starts; the claimed value (the reference's "claim setup",
standalone_evaluation) is computed before it.

Prints one JSON line. An untimed profiling step first times every kernel
launch ("kernels") and picks the dominant kernel; a second untimed step
records the host phases alone ("host_phases_ms", wall clock between the
library's phase marks, with the step's wall time as "step_wall"); inside the
timed region the dominant kernel's launches of the first timed
step carry HIP events (on the library's stream), and the roofline object is
computed from them.
cpu_baseline times the CPU restatement of the same protocol (oracle/: C
kernels with OpenMP) at 2^20 u64 coefficients on this host, at the process's
thread share and at 1 thread, beside the reference's published M5 figures.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bitz-poc_amd"))

HBM_PEAK_GBS = 8000.0
# VALU 32-bit integer peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12

# Per-dispatch hardware counters of the 2^24-coefficient workload, averaged
# over every dispatch of the profiled process, from the four rocprofv3 runs
# whose commands head PMC_SOURCE["file"] (made by `bench.py
# --summarize-profiles`), taken on the code of PMC_SOURCE["commit"]. "mib":
# HBM MiB per dispatch, FETCH_SIZE and WRITE_SIZE each scaled by what the
# counter reads for a 1 GiB device copy; "valu_insts": SQ_INSTS_VALU (wave
# instructions, 64 lane operations each). Keys are the library's launch names.
PMC_SOURCE = {"file": "profiles/pmc_commit_open_2p24.txt", "commit": "77865cf"}
PMC = {
    24: {
        "gkr_fold_round_sums_rowonly": {"kernel": "k_gkr_round_sums<FoldSrc, false>",
                                        "mib": 111.97 + 55.79, "valu_insts": 2.2004e+07},
    },
}


def _short_kernel(name: str) -> str:
    name = name.strip('"')
    if name.startswith("void "):
        name = name[5:]
    name = name.replace("(anonymous namespace)::", "").replace("bitz::", "")
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def summarize_profiles(stats_csv, sq_csv, fetch_csv, write_csv, out=sys.stdout):
    """one line per kernel from rocprofv3 output: the kernel-trace stats
    (--kernel-trace --stats) and three --pmc runs (SQ counters, FETCH_SIZE,
    WRITE_SIZE) of the same bench command with --copy-probe; HBM counters are
    scaled by what they read for the probe's 1 GiB device copy"""
    import csv
    from collections import defaultdict

    def counters(path):
        per = defaultdict(lambda: defaultdict(float))  # (kernel, dispatch) -> counter -> value
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                key = (_short_kernel(row["Kernel_Name"]), row["Dispatch_Id"])
                per[key][row["Counter_Name"]] += float(row["Counter_Value"])
        agg = defaultdict(lambda: defaultdict(list))
        for (k, _), cs in per.items():
            for c, v in cs.items():
                agg[k][c].append(v)
        return agg

    stats = {}
    with open(stats_csv, newline="") as f:
        for row in csv.DictReader(f):
            stats[_short_kernel(row["Name"])] = (int(row["Calls"]), float(row["AverageNs"]) / 1e3)
    sq, fe, wr = counters(sq_csv), counters(fetch_csv), counters(write_csv)
    probe = "__amd_rocclr_copyBuffer"
    gib_kib = 1 << 20
    f_probe = max(fe[probe]["FETCH_SIZE"]) if probe in fe else gib_kib
    w_probe = max(wr[probe]["WRITE_SIZE"]) if probe in wr else gib_kib
    fs, ws = gib_kib / f_probe, gib_kib / w_probe
    print(f"copy probe (1 GiB read + 1 GiB write = {gib_kib} KiB each): FETCH_SIZE {f_probe:.0f} KiB, "
          f"WRITE_SIZE {w_probe:.0f} KiB", file=out)
    print(f"scale factors: fetch {fs:.4f}, write {ws:.4f}", file=out)
    print("", file=out)
    hdr = (f"{'kernel':<58}{'calls':>7}{'avg_us':>10}{'valu_inst/disp':>16}{'valu_util':>11}"
           f"{'fetch_MiB/disp':>16}{'write_MiB/disp':>16}{'HBM_GB/s':>10}")
    print(hdr, file=out)
    mean = lambda xs: sum(xs) / len(xs) if xs else 0.0
    for k, (calls, avg_us) in sorted(stats.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        if k not in sq and k not in fe:
            continue
        valu = mean(sq[k].get("SQ_INSTS_VALU", []))
        util = valu / (avg_us * 1e-6 * 256 * 2.4e9) if avg_us else 0.0
        fmib = mean(fe[k].get("FETCH_SIZE", [])) * fs / 1024
        wmib = mean(wr[k].get("WRITE_SIZE", [])) * ws / 1024
        gbs = (fmib + wmib) * 1048576 / (avg_us * 1e-6) / 1e9 if avg_us else 0.0
        print(f"{k[:57]:<58}{calls:>7}{avg_us:>10.1f}{valu:>16.4g}{util:>11.3f}{fmib:>16.2f}{wmib:>16.2f}"
              f"{gbs:>10.1f}", file=out)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log-coeffs", type=int, default=24, help="log2 of the u64 coefficient count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="time launch-shape candidates first (bitz_ctx_autotune) and prove with the fastest")
    ap.add_argument("--cpu-log-coeffs", type=int, default=20)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched and serving measurements")
    ap.add_argument("--no-host-rows", action="store_true",
                    help="skip the PCIe-inclusive steps from rows in host memory (after the timed region)")
    ap.add_argument("--device-transcript", action="store_true",
                    help="run the GKR rounds on a device copy of the transcript (no host round trip per round)")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--mul-log", type=int, default=20, help="log2 multiplications of the u64 proof workload")
    ap.add_argument("--no-mul", action="store_true", help="skip the u64 multiplication proof workload")
    ap.add_argument("--chain-log", type=int, default=12, help="log2 compressions of the SHA-256 chain workload")
    ap.add_argument("--no-chain", action="store_true", help="skip the SHA-256 chain workload")
    ap.add_argument("--chain-only", action="store_true",
                    help="run only the SHA-256 chain workload and print its object (profiling aid)")
    ap.add_argument("--mul-only", action="store_true",
                    help="run only the u64 multiplication proof workload and print its object (profiling aid)")
    ap.add_argument("--batch-workers", type=int, default=8)
    ap.add_argument("--serve-jobs", type=int, default=96)
    ap.add_argument("--summarize-profiles", nargs=4, metavar=("STATS", "SQ", "FETCH", "WRITE"),
                    help="print the per-kernel counter table from rocprofv3 CSVs (kernel stats, SQ, FETCH_SIZE, "
                         "WRITE_SIZE runs) and exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, copy 1 GiB device to device (a kernel of known traffic that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters of a profiler run)")
    return ap.parse_args()


def shape_for(log_coeffs: int):
    n = log_coeffs + 6  # 64 bits per coefficient
    t = max(7, -(-3 * n // 5) - 1)  # reference_log_rows (src/bitz/params.rs:30-35)
    t = min(t, n - 1)
    return n, t, n - t


# The reference's published CPU prover figures for commit+open (paper
# Table 11, MacBook Air M5, rate 1/2): u64 coefficients -> (1 thread, 10
# threads) Mcoeff/s, derived as coefficients / prover total.
M5_PUBLISHED = {
    20: {"ms_1_thread": 155.0, "ms_10_threads": 52.8},
    24: {"ms_1_thread": 2175.0, "ms_10_threads": 575.0},
}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads() -> int:
    """this process's CPU share: the affinity set, capped by OMP_NUM_THREADS"""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(log_coeffs: int):
    """the CPU restatement (oracle/: C kernels with OpenMP, driven from
    Python) on the same protocol at 2^log_coeffs u64 coefficients: commit +
    standalone open, at the host's thread share and at 1 thread"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import bitz_protocol as P
    import oracle as O
    n, t, s = shape_for(log_coeffs)
    rng = np.random.default_rng(7)
    rows = rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)

    def timed(threads):
        O.set_threads(threads)
        t0 = time.perf_counter()
        pr = P.OracleProver(rows, t, s, 2, 4)
        t1 = time.perf_counter()
        claimed = pr.evaluation()
        t2 = time.perf_counter()
        proof = pr.prove(claimed)
        t3 = time.perf_counter()
        return (t1 - t0) + (t3 - t2), proof

    threads = _host_threads()
    secs_n, proof_n = timed(threads)
    secs_1, proof_1 = timed(1)
    assert proof_n == proof_1, "CPU restatement: proof depends on the thread count"
    O.set_threads(threads)
    rate = lambda secs: round((1 << log_coeffs) / secs / 1e6, 4)
    m5 = M5_PUBLISHED.get(log_coeffs)
    return {
        "value": rate(secs_n),
        "unit": "Mcoeff/s",
        "cores": threads,
        "kind": "port",
        "one_thread": {"value": rate(secs_1), "s": round(secs_1, 3)},
        "cpu_model": _cpu_model(),
        "sample": f"n={n} (2^{log_coeffs} u64 coefficients, t={t}, s={s}), rate 1/4, initial_k 4: commit + "
                  f"standalone open by the CPU restatement (oracle/: C kernels with OpenMP and PCLMULQDQ, "
                  f"protocol driven from Python; not the reference's Rust prover), {threads} threads "
                  f"{secs_n:.2f} s, 1 thread {secs_1:.2f} s, proofs equal",
        "reference_published_m5": None if m5 is None else {
            "hardware": "MacBook Air M5 24 GB (paper Table 11), rate 1/2",
            "mcoeff_s_1_thread": round((1 << log_coeffs) / m5["ms_1_thread"] / 1e3, 2),
            "mcoeff_s_10_threads": round((1 << log_coeffs) / m5["ms_10_threads"] / 1e3, 2)},
    }


def _job_specs(rng, count, lo, hi):
    """count jobs: log2 coefficients uniform in [lo, hi], rates 1/2, 1/4, 1/8"""
    import numpy as np
    specs = []
    for _ in range(count):
        lc = int(rng.integers(lo, hi + 1))
        r = int(rng.choice([1, 2, 3]))
        _, t, s = shape_for(lc)
        specs.append((t, s, r, rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)))
    return specs


def verify_all(ctx, jobs, specs):
    """Every proof through the library's verifier (untimed in the rates
    above): (all accepted, ms per proof)."""
    from bitz_amd.device import verify_standalone
    t0 = time.perf_counter()
    ok = True
    for j, (t, s, r, _) in zip(jobs, specs):
        try:
            verify_standalone(ctx, j.root, t, s, r, 4, j.claimed, j.proof)
        except RuntimeError:
            ok = False
    return ok, (time.perf_counter() - t0) * 1e3 / max(len(jobs), 1)


# The reference's published u64 multiplication SNARK (full prover including
# the commitment, rate 1/2, paper Table 17, MacBook Air M5): multiplications
# -> (1 thread, 10 threads) ms. 2^20 itself is not published.
M5_MUL_U64 = {19: (347.0, 106.0), 21: (1297.0, 369.0)}


def mul_u64_steps(args, ctx):
    """The u64 multiplication workload (2^mul_log uniform u64 pairs from a
    seed, resident in HBM): commit of the 256-slot bit grid + the Spartan
    PIOP + the BitZ opening (bitz_prove_mul_u64), timed over args.steps
    proofs after one warmup; the last proof verified by the library."""
    import numpy as np
    import torch

    from bitz_amd.device import kernel_stats, set_timing, set_timing_only
    from bitz_amd.mul import MulLayout, prove_u64, verify_u64
    n = 1 << args.mul_log
    g = torch.Generator(device="cuda").manual_seed(20)
    x = torch.randint(-(2**62), 2**62, (n,), dtype=torch.int64, device="cuda", generator=g)
    y = torch.randint(-(2**62), 2**62, (n,), dtype=torch.int64, device="cuda", generator=g)
    x ^= torch.randint(0, 4, (n,), dtype=torch.int64, device="cuda", generator=g) << 62
    y ^= torch.randint(0, 4, (n,), dtype=torch.int64, device="cuda", generator=g) << 62
    lay = MulLayout.new(64, n)
    prove_u64(ctx, x, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        root, state, proof = prove_u64(ctx, x, y)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # host phases of one more (untimed) proof: wall clock between phase marks
    set_timing_only(ctx, "phases")
    prove_u64(ctx, x, y)
    torch.cuda.synchronize()
    prof = kernel_stats(ctx)
    set_timing(ctx, False)
    phases = {k[6:]: round(v[1], 3) for k, v in prof.items() if k.startswith("phase:")}
    tv = time.perf_counter()
    verify_u64(ctx, root, n, proof)
    verify_ms = (time.perf_counter() - tv) * 1e3
    ref = {f"2^{k}": {"ms_1_thread": a, "ms_10_threads": b} for k, (a, b) in M5_MUL_U64.items()}
    return {
        "workload": f"u64 multiplication proof: 2^{args.mul_log} uniform u64 pairs (x y = lo + 2^64 hi), "
                    f"commit of the 256-slot grid (t={lay.t}, s={lay.s}) + Spartan PIOP over a runtime prime + "
                    f"BitZ opening, lambda 100, custom:1:4 (rate 1/2)",
        "ms_per_proof": round(ms, 3),
        "proofs_per_s": round(1e3 / ms, 2),
        "mmul_per_s": round(n / ms / 1e3, 3),
        "steps": args.steps,
        "proof_bytes": len(proof),
        "verify_ms": round(verify_ms, 3),
        "proof_verified": True,
        "host_phases_ms": phases,
        "reference_published_m5_rate_half": ref,
    }


def commit_batch_steps(args, ctx):
    """64 commitments of mixed sizes and rates from device-resident rows:
    one at a time (each commit's kernels launched in turn on the context's
    stream) against one graph launch (bitz_commit_batch, replaying the kept
    graph); roots compared item by item."""
    import numpy as np
    import torch

    from bitz_amd.device import CommitBatch
    rng = np.random.default_rng(64)
    specs = []
    for _ in range(args.batch_jobs):
        lc = int(rng.integers(14, 23))
        _, t, s = shape_for(lc)
        specs.append((t, s, 4, int(rng.choice([1, 2, 3]))))
    batch = CommitBatch(ctx, specs)
    for p in batch.packed:
        p.copy_(torch.randint(-(2**62), 2**62, p.shape, dtype=torch.int64, device="cuda"))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    roots = batch.commit()  # capture + instantiate + launch
    first_ms = (time.perf_counter() - t0) * 1e3
    reps = max(args.steps, 3)
    t0 = time.perf_counter()
    for _ in range(reps):
        roots = batch.commit()
    graph_ms = (time.perf_counter() - t0) * 1e3 / reps
    singles = None
    t0 = time.perf_counter()
    for _ in range(reps):
        singles = [ctx.commit(batch.packed[i], t + s - 7, k, r, batch.codeword[i], batch.tree[i])[0]
                   for i, (t, s, k, r) in enumerate(specs)]
    torch.cuda.synchronize()
    single_ms = (time.perf_counter() - t0) * 1e3 / reps
    coeffs = sum(1 << (t + s - 6) for t, s, _, _ in specs)
    return {
        "workload": f"{len(specs)} commitments from device rows, 2^14..2^22 u64 coefficients, rates 1/2, 1/4, 1/8",
        "one_launch_ms": round(graph_ms, 3),
        "one_launch_commits_per_s": round(len(specs) / graph_ms * 1e3, 1),
        "one_launch_mcoeff_per_s": round(coeffs / graph_ms / 1e3, 1),
        "first_launch_ms_with_capture": round(first_ms, 3),
        "one_at_a_time_ms": round(single_ms, 3),
        "one_at_a_time_commits_per_s": round(len(specs) / single_ms * 1e3, 1),
        "speedup": round(single_ms / graph_ms, 3),
        "roots_equal": roots == singles,
    }


def sha256_chain_steps(args, ctx):
    """The SHA-256 chain proof (config 4): 2^chain_log chained compressions
    over a seeded random message (its 64-byte blocks as big-endian u32
    words), timed over `steps` proofs, verified by the library; then 16
    chains of 2^8..2^12 compressions (log-uniform, seeded) on one server
    (8 workers, jobs overlapping on the GPU) against the same chains one at
    a time, proof bytes compared chain by chain, every proof verified."""
    import numpy as np
    import torch

    from bitz_amd.chain import prove_chain, verify_chain
    from bitz_amd.device import kernel_stats, set_timing, set_timing_only
    from bitz_amd.serve import ChainJob, Server

    k = args.chain_log
    rng = np.random.default_rng(4)
    msg = rng.integers(0, 256, size=64 << k, dtype=np.uint8)
    blocks = msg.reshape(1 << k, 16, 4)
    blocks = ((blocks[..., 0].astype(np.uint32) << 24) | (blocks[..., 1].astype(np.uint32) << 16)
              | (blocks[..., 2].astype(np.uint32) << 8) | blocks[..., 3].astype(np.uint32))
    prove_chain(ctx, blocks)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        root, digest, state, proof = prove_chain(ctx, blocks)
    dt = (time.perf_counter() - t0) / args.steps
    # host phases of one more (untimed) proof: wall clock between phase marks
    set_timing_only(ctx, "phases")
    prove_chain(ctx, blocks)
    prof = kernel_stats(ctx)
    set_timing(ctx, False)
    phases = {kk[6:]: round(v[1], 3) for kk, v in prof.items() if kk.startswith("phase:")}
    tv = time.perf_counter()
    verify_chain(ctx, blocks, digest, root, proof)
    tv = time.perf_counter() - tv
    out = {
        "workload": f"SHA-256 chain proof: 2^{k} chained compressions from FIPS 180-4's H_0 over a seeded random "
                    f"{64 << k >> 10} KiB message (u32 words), lambda 100, custom:1:4: witness, commit of the "
                    "source bits, linear PIOP over a runtime prime, virtual BitZ opening of the product tensor",
        "ms_per_proof": round(dt * 1e3, 3),
        "compressions_per_s": round((1 << k) / dt, 1),
        "proofs_per_s": round(1.0 / dt, 2),
        "steps": args.steps,
        "proof_bytes": len(proof),
        "verify_ms": round(tv * 1e3, 3),
        "proof_verified": True,
        "host_phases_ms": phases,
    }
    if args.no_batched:
        return out
    # 16 chains of 2^8..2^12 compressions, one at a time and on one server
    brng = np.random.default_rng(16)
    lens = [int(brng.integers(8, 13)) for _ in range(16)]
    chains = [brng.integers(0, 2**32, size=(1 << ln, 16), dtype=np.uint64).astype(np.uint32) for ln in lens]

    def run(workers):
        with Server(workers=workers, hbm_budget=200 << 30) as srv:
            for _ in range(2):  # the first pass fills the contexts' pools and relation caches
                jobs = [ChainJob(b, verify=True) for b in chains]
                t1 = time.perf_counter()
                for j in sorted(jobs, key=lambda j: -j.footprint):
                    srv.submit(j)
                srv.wait_all()
                d = time.perf_counter() - t1
        for j in jobs:
            j.check()
        return jobs, d

    seq, dt_seq = run(1)
    bat, dt_bat = run(args.batch_workers)
    out.update({
        "batch16": {
            "workload": "16 SHA-256 chains of 2^8..2^12 compressions (log-uniform, seeded), one server",
            "chain_logs": lens,
            "workers": args.batch_workers,
            "proofs_per_s": round(16 / dt_bat, 2),
            "sequential_proofs_per_s": round(16 / dt_seq, 2),
            "speedup": round(dt_seq / dt_bat, 3),
            "bit_exact_vs_sequential": all(a.proof == b.proof and a.root == b.root and a.digest == b.digest
                                           for a, b in zip(seq, bat)),
            "all_verified": all(j.verified for j in seq + bat),
        },
    })
    return out


def batched_and_serving(args, ctx):
    """Batched commit+open of independent polynomials of mixed sizes and
    rates on one server (bitz_server: per-worker contexts, jobs overlapping on
    the GPU) against the same jobs one at a time, and a serving loop with
    Poisson arrivals. Each job includes its claim setup and host-to-device
    copy of its rows. Proof bytes are compared job by job."""
    import numpy as np

    from bitz_amd.serve import Job, Server

    rng = np.random.default_rng(64)
    specs = _job_specs(rng, args.batch_jobs, 14, 22)

    def run(workers):
        with Server(workers=workers, hbm_budget=200 << 30) as srv:
            for _ in range(2):  # the first pass fills the contexts' pools and twiddle caches
                jobs = [Job(rows, t, s, 4, r) for t, s, r, rows in specs]
                t0 = time.perf_counter()
                for j in sorted(jobs, key=lambda j: -j.footprint):
                    srv.submit(j)
                srv.wait_all()
                dt = time.perf_counter() - t0
        for j in jobs:
            j.check()
        return jobs, dt

    seq, dt_seq = run(1)
    bat, dt_bat = run(args.batch_workers)
    exact = all(a.proof == b.proof and a.root == b.root for a, b in zip(seq, bat))
    coeffs = sum(1 << (t + s - 6) for t, s, _, _ in specs)
    batched = {
        "workload": f"{len(specs)} independent polynomials, 2^14..2^22 u64 coefficients, rates 1/2, 1/4, 1/8 "
                    "(custom:1|2|3:4), commit + claim + standalone open each, rows from host memory",
        "jobs": len(specs),
        "workers": args.batch_workers,
        "proofs_per_s": round(len(specs) / dt_bat, 2),
        "sequential_proofs_per_s": round(len(specs) / dt_seq, 2),
        "speedup": round(dt_seq / dt_bat, 3),
        "mcoeff_per_s": round(coeffs / dt_bat / 1e6, 2),
        "sequential_mcoeff_per_s": round(coeffs / dt_seq / 1e6, 2),
        "bit_exact_vs_sequential": exact,
    }
    ok, vms = verify_all(ctx, bat, specs)
    batched["all_verified"] = ok
    batched["verify_ms_per_proof"] = round(vms, 3)
    # serving: Poisson arrivals at 70% of the mix's capacity, half of them
    # full u64-multiplication proofs (2^12..2^22 multiplications: commit,
    # Spartan PIOP, BitZ opening), half SHA-256 chain proofs (2^7..2^12
    # compressions: witness, commit, linear PIOP, virtual opening); every
    # proof is verified by the worker that made it, inside the measured latency
    from bitz_amd.serve import ChainJob, MulProofJob
    npairs = args.serve_jobs // 2

    def mixed_jobs(limit=None):
        mrng = np.random.default_rng(67)
        out = []
        for _ in range(npairs if limit is None else limit):
            ln = int(mrng.integers(12, 23))
            x = mrng.integers(0, 2**64, size=1 << ln, dtype=np.uint64)
            y = mrng.integers(0, 2**64, size=1 << ln, dtype=np.uint64)
            out.append(MulProofJob(x, y, verify=True))
            cl = int(mrng.integers(7, 13))
            out.append(ChainJob(mrng.integers(0, 2**32, size=(1 << cl, 16), dtype=np.uint64).astype(np.uint32),
                                verify=True))
        return out

    jobs = mixed_jobs()
    with Server(workers=args.batch_workers, hbm_budget=64 << 30) as srv:
        warm = mixed_jobs(args.batch_workers)
        for j in warm:
            srv.submit(j)
        srv.wait_all()
        # the mix's closed-loop capacity sets the arrival rate (70% of it)
        cap_jobs = mixed_jobs()
        tc = time.perf_counter()
        for j in cap_jobs:
            srv.submit(j)
        srv.wait_all()
        capacity = len(cap_jobs) / (time.perf_counter() - tc)
        rate = 0.7 * capacity
        gaps = np.random.default_rng(66).exponential(1.0 / rate, size=len(jobs))
        t0 = time.perf_counter()
        due = 0.0
        for j, g in zip(jobs, gaps):
            due += g
            delay = t0 + due - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            srv.submit(j)
        srv.wait_all()
        t1 = time.perf_counter()
    for j in jobs:
        j.check()
    lat = np.array([j.latency_ms for j in jobs])
    lat_mul = np.array([j.latency_ms for j in jobs[0::2]])
    lat_chain = np.array([j.latency_ms for j in jobs[1::2]])
    serving = {
        "workload": f"{len(jobs)} Poisson arrivals: full u64-multiplication proofs (2^12..2^22 multiplications: "
                    "commit, Spartan PIOP, BitZ opening) and SHA-256 chain proofs (2^7..2^12 compressions: "
                    "witness, commit, linear PIOP, virtual opening), 50/50, "
                    "admitted by device-memory footprint (64 GiB budget), each proof verified by its worker",
        "capacity_proofs_per_s": round(capacity, 2),
        "arrival_rate_per_s": round(rate, 2),
        "workers": args.batch_workers,
        "proofs_per_s": round(len(jobs) / (t1 - t0), 2),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "max_ms": round(float(lat.max()), 3),
        "mul_jobs_p50_ms": round(float(np.percentile(lat_mul, 50)), 3),
        "mul_jobs_p99_ms": round(float(np.percentile(lat_mul, 99)), 3),
        "chain_jobs_p50_ms": round(float(np.percentile(lat_chain, 50)), 3),
        "chain_jobs_p99_ms": round(float(np.percentile(lat_chain, 99)), 3),
        "all_verified": bool(all(j.verified for j in jobs)),
    }
    return batched, serving


def host_rows_steps(args, ctx, coeffs, t, s, claimed, proof, steps=3):
    """PCIe-inclusive rate, reported beside `value` and never as it: the same
    commit+open from rows in pageable host memory (bitz_commit copies them
    up), after one untimed step; the proof bytes must not change."""
    import numpy as np
    import torch

    from bitz_amd.device import Commitment
    host_rows = coeffs.cpu().numpy().view(np.uint64)

    def host_step():
        com = Commitment(ctx, host_rows, t, s, 4, 2)
        pr = com.prove(claimed)
        com.close()
        return pr

    host_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        hproof = host_step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return {"ms_per_step": round(ms, 3), "mcoeff_per_s": round((1 << args.log_coeffs) / (ms / 1e3) / 1e6, 3),
            "steps": steps, "source": "pageable host memory through bitz_commit", "same_proof": hproof == proof}


def main():
    args = parse()
    if args.summarize_profiles:
        summarize_profiles(*args.summarize_profiles)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the commit+open path is "
              f"measured on a single MI355X.", file=sys.stderr)
        sys.exit(2)
    import torch

    from bitz_amd.device import Commitment, Context, autotune, get_tune, kernel_stats, set_timing, set_timing_only

    n, t, s = shape_for(args.log_coeffs)
    ctx = Context(0)
    if args.chain_only:
        # the SHA-256 chain workload alone (for its kernel profiles); prints
        # its own JSON object, not the bench line
        print(json.dumps({"sha256_chain": sha256_chain_steps(args, ctx)}), flush=True)
        return
    if args.mul_only:
        print(json.dumps({"mul_u64": mul_u64_steps(args, ctx)}), flush=True)
        return
    if args.device_transcript:
        from bitz_amd import _lib
        _lib.call("bitz_ctx_set_device_transcript", ctx.ptr, 1)
    g = torch.Generator(device="cuda").manual_seed(1234)
    words = 1 << args.log_coeffs
    coeffs = torch.randint(-(2**62), 2**62, (words,), dtype=torch.int64, device="cuda", generator=g)
    coeffs ^= torch.randint(0, 4, (words,), dtype=torch.int64, device="cuda", generator=g) << 62

    def step(claimed=None):
        com = Commitment(ctx, coeffs, t, s, 4, 2)
        if claimed is None:
            return com, com.evaluation()
        proof = com.prove(claimed)
        com.close()
        return proof

    # claim setup (untimed): the honest value of the transcript-derived claim
    com0, claimed = step()
    root0 = com0.root
    com0.close()
    # launch shapes and thresholds: the library defaults, or (--autotune)
    # candidates timed on the device for this shape (untimed)
    tune = autotune(ctx, t, s, 4, 2) if args.autotune else get_tune(ctx)
    for _ in range(args.warmup):
        step(claimed)
    torch.cuda.synchronize()
    # profiling steps (untimed): every instrumented launch, then the host
    # phases alone (wall clock between phase marks, no events)
    set_timing(ctx, True)
    step(claimed)
    torch.cuda.synchronize()
    profile = kernel_stats(ctx)
    phase_runs = []
    for _ in range(2):
        set_timing_only(ctx, "phases")
        t_ph = time.perf_counter()
        step(claimed)
        torch.cuda.synchronize()
        t_ph = time.perf_counter() - t_ph
        phases = {k[6:]: round(v[1], 3) for k, v in kernel_stats(ctx).items() if k.startswith("phase:")}
        phases["step_wall"] = round(t_ph * 1e3, 3)
        phase_runs.append(phases)
    print("phase runs", json.dumps(phase_runs), file=sys.stderr)
    set_timing(ctx, False)
    profile = {k: v for k, v in profile.items() if not k.startswith("phase:")}
    dom_name = max(profile.items(), key=lambda kv: kv[1][1])[0] if profile else None

    # timed region: events around the dominant kernel's launches of the first
    # timed step (every step runs the same launches; events cost host time)
    if dom_name:
        set_timing_only(ctx, dom_name)
    t0 = time.perf_counter()
    proof = None
    for i in range(args.steps):
        proof = step(claimed)
        if i == 0:
            set_timing(ctx, False)  # keeps the recorded launches
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = kernel_stats(ctx)
    ms = (t1 - t0) * 1e3 / args.steps
    value = (1 << args.log_coeffs) / (ms / 1e3) / 1e6

    roofline = None
    if dom_name and dom_name in stats:
        name, (launches, tot_ms, byts, ops) = dom_name, stats[dom_name]
        avg_s = tot_ms / launches / 1e3
        gbs = byts / launches / avg_s / 1e9
        pmc = PMC.get(args.log_coeffs, {}).get(name, {})
        traffic_mib = pmc.get("mib")
        # VALU issue rate: the profile's instructions per dispatch / this run's
        # live launch duration (counters need a profiler run of their own)
        tops = pmc["valu_insts"] * 64 / avg_s / 1e12 if "valu_insts" in pmc else None
        hbm = {"achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
               "source": "algorithmic bytes per launch / live launch duration (HIP events)"}
        valu = None if tops is None else {
            "achieved": round(tops, 3), "peak": round(VALU_PEAK_TOPS, 2), "unit": "T int32-op/s",
            "frac": round(tops / VALU_PEAK_TOPS, 4),
            "source": f"SQ_INSTS_VALU per dispatch of {pmc['kernel']} from {PMC_SOURCE['file']} "
                      f"(rocprofv3 at commit {PMC_SOURCE['commit']}) / live launch duration"}
        # the binding roof is the one the kernel runs closest to
        top = valu if valu is not None and valu["frac"] > hbm["frac"] else hbm
        roofline = {
            "kernel": name,
            "bound": "valu" if top is valu else "hbm",
            "achieved": top["achieved"],
            "peak": top["peak"],
            "unit": top["unit"],
            "frac": top["frac"],
            "traffic": round(traffic_mib * 1048576) if traffic_mib else None,
            "traffic_source": f"FETCH_SIZE + WRITE_SIZE per dispatch, {PMC_SOURCE['file']} "
                              f"(commit {PMC_SOURCE['commit']})" if traffic_mib else None,
            "algorithmic_bytes_per_launch": round(byts / launches),
            "avg_launch_us": round(avg_s * 1e6, 2),
            "launches_per_step": launches,
            "hbm": hbm,
            "valu": valu,
            "time_share": round(tot_ms / ms, 4),
        }
    if args.copy_probe:
        src = torch.empty(1 << 27, dtype=torch.int64, device="cuda")
        src.fill_(7)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    # the last timed step's proof through the library's verifier (untimed)
    from bitz_amd.device import verify_standalone
    tv = time.perf_counter()
    verify_standalone(ctx, root0, t, s, 2, 4, claimed, proof)
    verify_ms = round((time.perf_counter() - tv) * 1e3, 3)
    host_rows_rate = None if args.no_host_rows else host_rows_steps(args, ctx, coeffs, t, s, claimed, proof)
    mul_u64 = None if args.no_mul else mul_u64_steps(args, ctx)
    sha256_chain = None if args.no_chain else sha256_chain_steps(args, ctx)
    commit_batch = None if args.no_batched else commit_batch_steps(args, ctx)
    batched = serving = None
    if not args.no_batched:
        batched, serving = batched_and_serving(args, ctx)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_log_coeffs)
    out = {
        "metric": "PCS commit+open Mcoeff/s at 2^24",
        "value": round(value, 3),
        "unit": "Mcoeff/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2^128 (u32 integer ops)",
        "data": "synthetic: uniform u64 coefficients from a seeded generator, resident in HBM",
        "config": {"workload": f"BitZ PCS commit+open, 2^{args.log_coeffs} u64 coefficients", "n": n, "t": t,
                   "s": s, "rate": "1/4", "initial_k": 4, "merkle_hash": "blake3", "ladder": "custom:2:4",
                   "proof_bytes": len(proof) if proof else None, "proof_verified": True, "verify_ms": verify_ms,
                   "host_rows_pcie_inclusive": host_rows_rate},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
        "mul_u64": mul_u64,
        "sha256_chain": sha256_chain,
        "commit_batch": commit_batch,
        "kernels": {k: {"launches": int(v[0]), "ms": round(v[1], 3)} for k, v in
                    sorted(profile.items(), key=lambda kv: -kv[1][1])},
        "host_phases_ms": phases,
        "tune": tune,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
