"""The traceback pool is planned within what the device can still give.

A context's default budget is the HBM free at halign_init less a reserve, so
a context created before other allocations (another context's pool, a
framework's tensors) would plan past the memory that is left. The batch
runner therefore plans each call within min(budget, its pool + free HBM - 1
GiB): under that pressure the launches shrink (fewer workgroups, more rounds
of tickets) and the results stay those of an unconstrained run."""
import sys

import numpy as np
import pytest

from seqgen import mutate_codes, rand_codes
from test_dp_gpu import _tasks

pytestmark = pytest.mark.gpu


def _jobs(n, tlen, seed):
    """n pairs of a tlen-base target and a 1%-mutated copy, band 4,665 (n_col_ 293:
    the one-wave ring of five slots, in either form)."""
    rng = np.random.default_rng(seed)
    base = rand_codes(rng, tlen + 64)
    cases = []
    for i in range(n):
        s = int(rng.integers(0, 64))
        t = base[s:s + tlen].copy()
        q = mutate_codes(rng, t, 0.008, 0.002)
        cases.append((q, t, 4665, 0x06))
    return cases


def test_pool_shrinks_under_memory_pressure():
    import torch
    import halign
    cases = _jobs(2048, 8000, 99)
    tasks, qc, tc = _tasks(cases)
    ref = halign.Context(0)
    try:
        print("2,048 pairs of 8 kb, unconstrained", file=sys.stderr, flush=True)
        r0, c0, _ = ref.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
    finally:
        ref.close()
    torch.cuda.empty_cache()
    ctx = halign.Context(0)   # budget taken now, with nearly all HBM free
    hog = None
    try:
        free, _ = torch.cuda.mem_get_info(0)
        left = 24 << 30
        hog = torch.empty(int(free - left), dtype=torch.uint8, device="cuda:0")
        print(f"holding {hog.numel() / 2**30:.0f} GiB elsewhere, {left >> 30} GiB left", file=sys.stderr, flush=True)
        ctx.set_bucket_timing(True)
        r1, c1, _ = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
        rows = ctx.last_buckets()
        ctx.set_bucket_timing(False)
    finally:
        ctx.close()
        del hog
        torch.cuda.empty_cache()
    assert rows and all(b["jobs"] == 2048 for b in rows if b["family"] == 0), rows
    assert np.array_equal(r1["score"], r0["score"]) and np.array_equal(r1["n_cigar"], r0["n_cigar"])
    assert all(np.array_equal(a, b) for a, b in zip(c1, c0))
    assert (r1["status"] == 0).all()
