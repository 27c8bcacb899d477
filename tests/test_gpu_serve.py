"""Batched and served commit+open jobs: every job's root, claim and proof
bytes are the same run alone and in a mix of concurrent jobs of different
sizes and code rates; the CPU verifier accepts them; admission keeps the
device-memory footprint of the jobs in flight within the budget."""
import numpy as np
import pytest

import bitz_protocol as P
from gpu_util import log

pytestmark = pytest.mark.gpu

# (t, s, log_inv_rate): n = t + s in 20..23, rates 1/2, 1/4, 1/8
SHAPES = [(12, 8, 1), (12, 8, 2), (12, 8, 3), (13, 8, 2), (13, 9, 1), (12, 9, 3), (14, 9, 2), (13, 10, 3)]


def _jobs(seed):
    from bitz_amd.serve import Job
    rng = np.random.default_rng(seed)
    return [Job(rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64), t, s, 4, r)
            for t, s, r in SHAPES]


def _run_alone(jobs):
    from bitz_amd.serve import Server
    with Server(workers=1) as srv:
        for j in jobs:
            srv.submit(j)
            srv.wait_all()
            j.check()
    return [(j.root, j.claimed, j.proof) for j in jobs]


def test_jobs_alone_and_in_the_mix_are_identical():
    from bitz_amd.serve import Server, run_batch
    log("serve: jobs alone")
    alone = _run_alone(_jobs(11))
    log("serve: the same jobs on a 4-worker server, submitted in a shuffled order")
    mix = _jobs(11)
    order = [5, 0, 7, 2, 6, 1, 3, 4]
    with Server(workers=4) as srv:
        for i in order:
            srv.submit(mix[i])
        srv.wait_all()
    for a, j in zip(alone, mix):
        j.check()
        assert (j.root, j.claimed, j.proof) == a
    # intervals overlap: the jobs really ran concurrently
    spans = sorted((j.s_.start_ms, j.s_.end_ms) for j in mix)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    log("serve: bitz_batch_run with 3 workers")
    batch = _jobs(11)
    run_batch(batch, workers=3)
    for a, j in zip(alone, batch):
        assert (j.root, j.claimed, j.proof) == a
    for (t, s, r), j in list(zip(SHAPES, mix))[:3]:
        P.verify_standalone(j.root, t, s, r, 4, j.claimed, j.proof)


def test_admission_respects_the_budget():
    from bitz_amd.serve import Server
    jobs = [j for j in _jobs(12) if (j.s_.t, j.s_.s, j.s_.log_inv_rate) == (12, 8, 2)]
    jobs += _jobs(13)[1:2] + _jobs(14)[1:2]  # three jobs of one shape and footprint
    fp = jobs[0].footprint
    log(f"serve: budget of 1.5 job footprints ({fp} B each)")
    with Server(workers=3, hbm_budget=fp * 3 // 2) as srv:
        for j in jobs:
            srv.submit(j)
        srv.wait_all()
    spans = sorted((j.s_.start_ms, j.s_.end_ms) for j in jobs)
    for j in jobs:
        j.check()
    # one at a time: no two jobs in flight together
    assert all(spans[i + 1][0] >= spans[i][1] for i in range(len(spans) - 1))


def test_bad_job_reports_its_error_and_the_server_continues():
    from bitz_amd.serve import Job, Server
    from bitz_amd._lib import BitzError
    good = _jobs(15)[:2]
    rng = np.random.default_rng(3)
    # rate 1/256 commits, but no custom Johnson ladder has it (inverse-rate
    # exponents 1..4): the opening rejects the job
    bad = Job(rng.integers(0, 2**64, size=(1 << 8) << 6, dtype=np.uint64), 12, 8, 6, 8)
    with Server(workers=2) as srv:
        srv.submit(good[0])
        srv.submit(bad)
        srv.submit(good[1])
        srv.wait_all()
    good[0].check()
    good[1].check()
    with pytest.raises(BitzError):
        bad.check()


def test_multiplication_jobs_in_a_verified_mix():
    """u64-multiplication jobs (the witness grid packed on the device by the
    worker) beside row jobs on one server, each proof verified by the worker:
    a multiplication job's root, claim and proof equal those of the grid
    packed through bitz_mul_bit_rows_dev and committed directly, and do not
    change in the mix"""
    import torch

    from bitz_amd.device import Commitment
    from bitz_amd.mul import MulWitness
    from bitz_amd.serve import Job, MulJob, Server
    from gpu_util import ctx
    rng = np.random.default_rng(21)
    ops = [(rng.integers(0, 2**64, size=n, dtype=np.uint64), rng.integers(0, 2**64, size=n, dtype=np.uint64), r)
           for n, r in [(1 << 12, 2), (3000, 1), (1 << 13, 3)]]
    log("serve: multiplication jobs alone (direct packing + commit)")
    direct = []
    for x, y, r in ops:
        w = MulWitness(ctx(), 64, torch.from_numpy(x.view(np.int64)).cuda(), torch.from_numpy(y.view(np.int64)).cuda())
        com = Commitment(ctx(), w.bitz_bit_rows(), w.layout.t, w.layout.s, 4, r)
        claimed = com.evaluation()
        direct.append((com.root, claimed, com.prove(claimed)))
        com.close()
    log("serve: multiplication and row jobs on a 4-worker server, verified by the workers")
    mjobs = [MulJob(x, y, 4, r, verify=True) for x, y, r in ops]
    rjobs = [Job(rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64), t, s, 4, r, verify=True)
             for t, s, r in SHAPES[:4]]
    with Server(workers=4) as srv:
        for j in [mjobs[0], rjobs[0], mjobs[1], rjobs[1], rjobs[2], mjobs[2], rjobs[3]]:
            srv.submit(j)
        srv.wait_all()
    for j, d in zip(mjobs, direct):
        j.check()
        assert j.verified
        assert (j.root, j.claimed, j.proof) == d
    for j in rjobs:
        j.check()
        assert j.verified
    x, y, r = ops[1]
    t, s = mjobs[1].t, mjobs[1].s
    P.verify_standalone(mjobs[1].root, t, s, r, 4, mjobs[1].claimed, mjobs[1].proof)


def test_multiplication_proof_jobs_in_a_verified_mix():
    """full u64 multiplication proofs (Spartan + BitZ) as server jobs beside
    row jobs, each verified by its worker: root, transcript state and proof
    bytes equal bitz_prove_mul_u64 run alone, and one is accepted by the CPU
    verifier"""
    import spartan_u64 as S
    import torch

    from bitz_amd.mul import prove_u64
    from bitz_amd.serve import Job, MulProofJob, Server
    from gpu_util import ctx
    rng = np.random.default_rng(31)
    ops = [(rng.integers(0, 2**64, size=n, dtype=np.uint64), rng.integers(0, 2**64, size=n, dtype=np.uint64))
           for n in (1 << 12, 5000, 1 << 14)]
    log("serve: u64 multiplication proofs alone")
    direct = []
    for x, y in ops:
        root, state, proof = prove_u64(ctx(), torch.from_numpy(x.view(np.int64)).cuda(),
                                       torch.from_numpy(y.view(np.int64)).cuda())
        direct.append((root, int.from_bytes(state[:16], "little"), proof))
    log("serve: u64 multiplication proofs and row jobs on a 4-worker server")
    mjobs = [MulProofJob(x, y, verify=True) for x, y in ops]
    rjobs = [Job(rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64), t, s, 4, r, verify=True)
             for t, s, r in SHAPES[:3]]
    with Server(workers=4) as srv:
        for j in [mjobs[0], rjobs[0], mjobs[1], rjobs[1], mjobs[2], rjobs[2]]:
            srv.submit(j)
        srv.wait_all()
    for j, d in zip(mjobs, direct):
        j.check()
        assert j.verified
        assert (j.root, j.claimed, j.proof) == d
    for j in rjobs:
        j.check()
        assert j.verified
    S.verify_u64(mjobs[1].root, 5000, 0, mjobs[1].proof)


def test_poisson_arrivals_continuous_batching():
    """continuous batching under Poisson arrivals: 24 jobs (row commit+open
    jobs and full u64 multiplication proofs, alternating) submitted at
    exponential inter-arrival times while earlier jobs run; every job is
    verified by its worker, admitted jobs overlap in time, and every result
    equals the same job run alone"""
    import time

    from bitz_amd.serve import Job, MulProofJob, Server
    rng = np.random.default_rng(41)
    specs = []
    for i in range(24):
        if i % 2 == 0:
            t, s, r = SHAPES[int(rng.integers(0, len(SHAPES)))]
            specs.append(("rows", rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64), t, s, r))
        else:
            n = int(rng.choice([1 << 12, 5000, 1 << 13]))
            specs.append(("mul", rng.integers(0, 2**64, size=n, dtype=np.uint64),
                          rng.integers(0, 2**64, size=n, dtype=np.uint64)))

    def make(sp, verify):
        if sp[0] == "rows":
            return Job(sp[1], sp[2], sp[3], 4, sp[4], verify=verify)
        return MulProofJob(sp[1], sp[2], verify=verify)

    log("serve: 24 jobs alone")
    alone = [make(sp, False) for sp in specs]
    with Server(workers=1) as srv:
        for j in alone:
            srv.submit(j)
            srv.wait_all()
            j.check()
    log("serve: 24 Poisson arrivals (mean gap 3 ms) on a 4-worker server")
    gaps = rng.exponential(0.003, size=len(specs))
    live = [make(sp, True) for sp in specs]
    with Server(workers=4) as srv:
        for j, gap in zip(live, gaps):
            time.sleep(float(gap))
            srv.submit(j)
        srv.wait_all()
    for a, j in zip(alone, live):
        j.check()
        assert j.verified
        assert (j.root, j.claimed, j.proof) == (a.root, a.claimed, a.proof)
        assert j.s_.start_ms >= j.s_.submit_ms and j.s_.end_ms >= j.s_.start_ms
    spans = sorted((j.s_.start_ms, j.s_.end_ms) for j in live)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))


def test_config5_mix_of_multiplication_and_chain_proofs():
    """the serving mix of config 5 under Poisson arrivals: u64 multiplication
    proofs and SHA-256 chain proofs alternating, submitted at exponential
    inter-arrival times on a 4-worker server; each is verified by its worker
    and equals the same proof made alone (root, transcript state, proof
    bytes, and the chain's digest)"""
    import time

    from bitz_amd.chain import prove_chain
    from bitz_amd.serve import ChainJob, MulProofJob, Server
    from gpu_util import ctx
    rng = np.random.default_rng(55)
    specs = []
    for i in range(12):
        if i % 2 == 0:
            n = int(rng.choice([1 << 12, 3000, 1 << 13]))
            specs.append(("mul", rng.integers(0, 2**64, size=n, dtype=np.uint64),
                          rng.integers(0, 2**64, size=n, dtype=np.uint64)))
        else:
            k = int(rng.integers(7, 11))
            specs.append(("chain", rng.integers(0, 2**32, size=(1 << k, 16), dtype=np.uint64).astype(np.uint32)))
    log("serve: chain proofs alone")
    alone = {}
    for i, sp in enumerate(specs):
        if sp[0] == "chain":
            root, digest, state, proof = prove_chain(ctx(), sp[1])
            alone[i] = (root, digest, int.from_bytes(state[:16], "little"), proof)
    log("serve: 12 Poisson arrivals (u64 and chain proofs) on a 4-worker server")
    live = [MulProofJob(sp[1], sp[2], verify=True) if sp[0] == "mul" else ChainJob(sp[1], verify=True)
            for sp in specs]
    with Server(workers=4) as srv:
        for j, gap in zip(live, rng.exponential(0.003, size=len(specs))):
            time.sleep(float(gap))
            srv.submit(j)
        srv.wait_all()
    for i, j in enumerate(live):
        j.check()
        assert j.verified
        if i in alone:
            assert (j.root, j.digest, j.claimed, j.proof) == alone[i]
