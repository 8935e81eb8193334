"""GPU parity of the query-fixed DP kernel (dp_qfix.hip) against the oracle.

The kernel takes CIGAR-mode jobs whose band is the whole matrix and whose
query is the shorter side (the anchor-less whole pairs and the long end
segments of the per-query path). Every shape is forced on cases it holds:
queries from one base to its full capacity, targets from the query's length
to tens of kilobases, N bases, both gap directions, waves whose jobs end at
different anti-diagonals, and a launch under a small traceback budget (each
workgroup's slots reused ticket after ticket). score, mte, mte_q, the cell count and the CIGAR must be
bit-identical to the oracle (pinned to the reference by test_oracle.py).
"""
import sys

import numpy as np
import pytest

from oracle_lib import FLAG_LEFT, FLAG_RIGHT, Oracle
from seqgen import mutate_codes, rand_codes
from test_dp_gpu import _compare, _tasks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oracle():
    return Oracle()


def _qfix_cases(rng, cap_bases, n, max_t=6000):
    """Full-band cases with qlen <= min(tlen, cap_bases)."""
    cases = []
    for it in range(n):
        kind = it % 5
        ql = int(rng.integers(1, cap_bases + 1))
        if it < 4:   # edges: one base, full capacity, square, one more target base
            ql = (1, cap_bases, cap_bases, max(1, cap_bases - 3))[it]
        if kind == 0 or it < 4:
            tl = ql + (0, 1, int(rng.integers(0, 40)), int(rng.integers(0, 3)))[it % 4]
        else:
            tl = int(rng.integers(ql, max(ql + 1, max_t)))
        t = rand_codes(rng, tl, p_n=0.02 if kind == 3 else 0.0)
        if kind in (1, 2) and tl > ql:
            # a mutated piece of the target (an end segment before the first anchor)
            s = int(rng.integers(0, tl - ql + 1))
            q = mutate_codes(rng, t[s:s + ql], 0.05, 0.02)[:ql]
            if len(q) == 0:
                q = rand_codes(rng, 1)
        else:
            q = rand_codes(rng, ql, p_n=0.03 if kind == 3 else 0.0)
        if len(q) > len(t):
            q = q[:len(t)]
        flag = FLAG_RIGHT if rng.random() < 0.6 else FLAG_LEFT
        w = -1 if rng.random() < 0.7 else max(len(q), len(t)) - 1 + int(rng.integers(0, 3))
        cases.append((q, t, w, flag))
    return cases


def _shape_ids():
    import halign
    return [(g, s) for g, s in halign.qfix_shapes()]


@pytest.mark.parametrize("shape", _shape_ids(), ids=lambda p: f"q{p[0]}x{p[1]}")
def test_qfix_shape_vs_oracle(shape, oracle):
    import halign
    g, s = shape
    cap = 16 * g * s
    rng = np.random.default_rng(1000 + 17 * g + s)
    n = 96 if g <= 4 else 48
    cases = _qfix_cases(rng, cap, n, max_t=6000 if cap <= 256 else 3000)
    ctx = halign.Context(0)
    try:
        ctx.set_kernel_policy(halign.qfix_policy(g, s))
        ctx.set_bucket_timing(True)
        tasks, qc, tc = _tasks(cases)
        print(f"qfix {g}x{s}: {n} cases, query <= {cap}", file=sys.stderr, flush=True)
        res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
        rows = ctx.last_buckets()
        ctx.set_bucket_timing(False)
    finally:
        ctx.close()
    want = halign.shape_names().index(f"q{g}x{s}")
    assert rows and all(b["shape"] == want for b in rows if b["family"] == 0), rows
    assert _compare(oracle, cases, res, cigs, False, f"q{g}x{s}") == 0


def test_qfix_auto_whole_pairs_and_end_segments(oracle):
    """The automatic policy on the per-query path's shapes: queries of 100-550
    bases and of 3-16 bases against 20 kb targets, unbanded, ragged lengths in
    every wave."""
    import halign
    rng = np.random.default_rng(4242)
    cases = []
    for i in range(40):
        tl = int(rng.integers(15000, 20001))
        t = rand_codes(rng, tl)
        ql = int(rng.integers(100, 551)) if i % 2 else int(rng.integers(3, 17))
        if i % 4 < 2:
            s = int(rng.integers(0, tl - ql))
            q = mutate_codes(rng, t[s:s + ql], 0.03, 0.01)[:ql]
        else:
            q = rand_codes(rng, ql)
        cases.append((q, t, -1, FLAG_RIGHT if i % 3 else FLAG_LEFT))
    ctx = halign.Context(0)
    try:
        ctx.set_bucket_timing(True)
        tasks, qc, tc = _tasks(cases)
        print("qfix auto: 40 pairs against 15-20 kb targets", file=sys.stderr, flush=True)
        res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
        rows = ctx.last_buckets()
        ctx.set_bucket_timing(False)
    finally:
        ctx.close()
    names = halign.shape_names()
    assert all(names[b["shape"]].startswith("q") for b in rows if b["family"] == 0), rows
    assert _compare(oracle, cases, res, cigs, False, "qfix-auto") == 0


def test_qfix_slots_reused(oracle):
    """One-lane jobs under a small traceback budget: the planner keeps a few
    workgroups, and each one's traceback slots serve several tickets in a row."""
    import halign
    rng = np.random.default_rng(77)
    base = _qfix_cases(rng, 16, 64, max_t=400)
    ctx = halign.Context(0)
    try:
        ctx.set_kernel_policy(halign.qfix_policy(1, 1))
        # 2 MB: about 9 workgroups of 64 slots of up to 3.3 KB for 40 tickets
        assert ctx.lib.halign_set_traceback_budget(ctx.h, 2 << 20) == 0
        cases = base * 40
        tasks, qc, tc = _tasks(cases)
        print(f"qfix 1x1: {len(cases)} jobs", file=sys.stderr, flush=True)
        res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
    finally:
        ctx.close()
    picks = list(range(0, len(cases), 37)) + [len(cases) - 1]
    sub = [cases[i] for i in picks]
    assert _compare(oracle, sub, res[picks], [cigs[i] for i in picks], False, "qfix-reuse") == 0


def test_qfix_combined_launch_vs_oracle(oracle):
    """Queries of 1-128 bases against 8-12 kb targets, both gap directions:
    several query-fixed buckets bound by their anti-diagonal chains, which the
    batch runner merges into one combined launch (bucket statistics family 3)."""
    import halign
    rng = np.random.default_rng(31337)
    cases = []
    for i in range(96):
        tl = int(rng.integers(8000, 12001))
        t = rand_codes(rng, tl)
        ql = int(rng.integers(1, 17)) if i % 3 == 0 else int(rng.integers(17, 65)) if i % 3 == 1 else \
            int(rng.integers(65, 129))
        if i % 2:
            s0 = int(rng.integers(0, tl - ql))
            q = mutate_codes(rng, t[s0:s0 + ql], 0.04, 0.01)[:ql]
            if len(q) == 0:
                q = rand_codes(rng, 1)
        else:
            q = rand_codes(rng, ql)
        cases.append((q, t, -1, FLAG_RIGHT if (i // 2) % 2 else FLAG_LEFT))
    ctx = halign.Context(0)
    try:
        ctx.set_bucket_timing(True)
        tasks, qc, tc = _tasks(cases)
        print("qfix combined: 96 short queries against 8-12 kb targets", file=sys.stderr, flush=True)
        res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
        rows = ctx.last_buckets()
        ctx.set_bucket_timing(False)
        # the same batch again: the plan's graph is captured and replayed
        res2, cigs2, _ = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
        res3, cigs3, _ = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
    finally:
        ctx.close()
    comb = [b for b in rows if b["family"] == 3]
    assert len(comb) == 1 and comb[0]["jobs"] >= 2, rows
    assert _compare(oracle, cases, res, cigs, False, "qfix-combined") == 0
    for r2, c2 in ((res2, cigs2), (res3, cigs3)):
        assert np.array_equal(r2["score"], res["score"]) and all(np.array_equal(a, b) for a, b in zip(c2, cigs))
