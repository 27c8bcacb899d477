"""The batched-commitment workload: 64 independent polynomials whose sizes
are 2^14..2^22 u64 coefficients (log2 uniform, n = 20..28) at code rates
mixed over 1/2, 1/4 and 1/8, each committed and opened at its
transcript-derived point. Every job's root, claim and proof bytes are the
same run one at a time and run as one batch on eight workers, and the CPU
restatement's verifier accepts a sample of the batched proofs (the smallest,
a middle and the largest job)."""
import numpy as np
import pytest

import bitz_protocol as P
from gpu_util import log

pytestmark = pytest.mark.gpu


def shape_for(log_coeffs: int):
    """t = max(7, ceil(3n/5) - 1) rows (src/bitz/params.rs:30-35), n = log2
    coefficients + 6 committed bits"""
    n = log_coeffs + 6
    t = min(max(7, -(-3 * n // 5) - 1), n - 1)
    return t, n - t


def _jobs(seed, count=64):
    from bitz_amd.serve import Job
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        lc = int(rng.integers(14, 23))
        r = int(rng.choice([1, 2, 3]))
        t, s = shape_for(lc)
        specs.append((t, s, r, rng.integers(0, 2**64, size=(1 << s) << (t - 6), dtype=np.uint64)))
    return specs, [Job(rows, t, s, 4, r) for t, s, r, rows in specs]


def test_64_mixed_jobs_batched_equal_sequential():
    from bitz_amd.serve import Server, run_batch
    specs, alone = _jobs(2025)
    log("batch: 64 jobs one at a time")
    with Server(workers=1) as srv:
        for j in alone:
            srv.submit(j)
            srv.wait_all()
            j.check()
    log("batch: the same 64 jobs as one batch on 8 workers")
    _, batch = _jobs(2025)
    run_batch(batch, workers=8)
    for a, b in zip(alone, batch):
        b.check()
        assert (b.root, b.claimed, b.proof) == (a.root, a.claimed, a.proof)
    sizes = [t + s for t, s, _, _ in specs]
    order = np.argsort(sizes, kind="stable")
    for i in (order[0], order[len(order) // 2], order[-1]):
        t, s, r, _ = specs[i]
        log(f"batch: CPU verifier on job {i} (n={t + s}, rate 1/{1 << r})")
        P.verify_standalone(batch[i].root, t, s, r, 4, batch[i].claimed, batch[i].proof)


def test_commit_batch_one_graph_launch_equals_single_commits():
    """64 commitments of mixed sizes (2^14..2^22 coefficients) and rates
    (1/2, 1/4, 1/8) in one graph launch (bitz_commit_batch): every root,
    codeword and tree equals the single commit of the same rows; the same
    batch refilled with new rows replays the kept graph, and every replay is
    checked against single commits, with other device work between replays"""
    import torch

    from bitz_amd.device import CommitBatch
    from gpu_util import ctx
    rng = np.random.default_rng(64)
    specs = []
    for _ in range(64):
        lc = int(rng.integers(14, 23))
        t, s = shape_for(lc)
        specs.append((t, s, 4, int(rng.choice([1, 2, 3]))))
    batch = CommitBatch(ctx(), specs)
    for rep in range(3):
        log(f"commit batch: 64 mixed commitments in one graph launch, replay {rep}")
        for p in batch.packed:
            p.copy_(torch.randint(-(2**62), 2**62, p.shape, dtype=torch.int64, device="cuda"))
        roots = batch.commit()
        for i, (t, s, k, r) in enumerate(specs):
            root, cw, tree = ctx().commit(batch.packed[i], t + s - 7, k, r)
            assert roots[i] == root
            assert torch.equal(cw, batch.codeword[i]) and torch.equal(tree, batch.tree[i])
        torch.cuda.synchronize()
