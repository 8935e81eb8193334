"""The DP planner's measured cost table (halign_tune, csrc/dp_tune_table.h).

CPU: the table compiled into the library covers every ring shape the kernel
has, with positive, finite costs. GPU: a fresh measurement fills the same
table, and the shape the planner gives a batch of equal jobs is the one its
greedy rule picks from the measured figures (restated here), for the bands
of the headline pairwise workload (n_col_ 293), a whole pair of a short query
against a 20 kb center, and short banded segments."""
import math
import os
import re
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "..", "halign-4_amd", "csrc", "dp_tune_table.h")


def table_rows():
    rows = []
    for line in open(TABLE):
        m = re.match(r"\s*\{(-?\d+), (\d+), (\d+), (\d+), (\d+), (\d+), \{([^}]*)\}, \{([^}]*)\}, ([0-9.]+)\},", line)
        if m:
            n = int(m.group(6))
            rows.append({"waves": int(m.group(1)), "lanes": int(m.group(2)), "ring": int(m.group(3)),
                         "uniform": int(m.group(4)), "jobs_in_flight": int(m.group(5)),
                         "ncol": [int(x) for x in m.group(7).split(",")][:n],
                         "ns_diag": [float(x) for x in m.group(8).split(",")][:n], "ns_job": float(m.group(9))})
    return rows


def check_rows(rows, shapes):
    seen = {(r["waves"], r["lanes"], r["ring"], r["uniform"]) for r in rows}
    # ring_shapes lists each sub-wave shape twice (per-group form, then uniform
    # form: 1) and a ring shape with a two-per-SIMD form twice (then form 2)
    want = set()
    for w, l, ring in shapes:
        u = (2 if w else 1) if (w, l, ring, 0) in want else 0
        want.add((w, l, ring, u))
    assert want <= seen, f"shapes without a measurement: {sorted(want - seen)}"
    for r in rows:
        assert r["jobs_in_flight"] > 0 and 1 <= len(r["ncol"]) <= 4
        # a ring shape serves n_col_ <= ring - 2, a query-fixed one (waves -1)
        # queries of up to ring blocks
        assert r["ncol"] == sorted(r["ncol"]) and r["ncol"][-1] <= r["ring"] - (0 if r["waves"] < 0 else 2)
        assert all(0 < t < 1e6 and math.isfinite(t) for t in r["ns_diag"]) and r["ns_job"] >= 0


def test_builtin_table_covers_every_shape():
    from halign import _lib
    rows = table_rows()
    assert rows, "no measured entries in dp_tune_table.h"
    check_rows(rows, _lib.ring_shapes())


def td_at(r, nc):
    td = r["ns_diag"][0]
    for k in range(1, len(r["ncol"])):
        if nc <= r["ncol"][k - 1]:
            break
        f = min(1.0, (nc - r["ncol"][k - 1]) / (r["ncol"][k] - r["ncol"][k - 1]))
        td = r["ns_diag"][k - 1] + f * (r["ns_diag"][k] - r["ns_diag"][k - 1])
    return td


def round_slot(b):
    g = 256
    while g * 64 <= b:
        g <<= 1
    return (b + g - 1) // g * g


def greedy_pick(table, shapes, n, qlen, tlen, w, budget):
    """The planner's choice for one bucket of n equal jobs: candidates in
    rising chip time (tp); it starts on the first and moves to the shape with
    the lowest max(crit, tp) when that is lower (crit: one job's time), the
    first such shape in that order. Uniform-ticket forms are not candidates;
    a two-per-SIMD form is one only for at least its jobs in flight whose
    traceback slots (n_col_ x 8 bytes per anti-diagonal, rounded up to 1/32 of
    their power of two) fit the budget."""
    we = max(qlen, tlen) if w < 0 else w
    nc = (min(qlen, tlen, we + 1) + 15) // 16 + 1
    nd = qlen + tlen - 1
    by_key = {(r["waves"], r["lanes"], r["ring"], r["uniform"]): r for r in table}
    cand = []
    seen = set()
    for i, (wv, l, ring) in enumerate(shapes):
        form = (2 if wv else 1) if (wv, l, ring) in seen else 0
        key = (wv, l, ring, form)
        seen.add((wv, l, ring))
        if form == 1 or nc > ring - 2 or key not in by_key or nc < by_key[key]["ncol"][0]:
            continue
        r = by_key[key]
        if form == 2 and (n < r["jobs_in_flight"] or r["jobs_in_flight"] * round_slot(nd * nc * 8) > budget):
            continue
        td = td_at(r, nc)
        crit = nd * td + r["ns_job"]
        tp = (n * nd * td + n * r["ns_job"]) / r["jobs_in_flight"]
        cand.append((tp, i, max(crit, tp)))
    cand.sort()
    best = min(c[2] for c in cand)
    if best >= cand[0][2] * (1 - 1e-9):
        return cand[0][1]
    return next(c[1] for c in cand if c[2] <= best * (1 + 1e-12))


@pytest.mark.gpu
def test_tune_measures_and_plans():
    from halign import _lib
    ctx = _lib.Context(0)
    try:
        shapes = _lib.ring_shapes()
        builtin = ctx.tune_table()
        check_rows(builtin, shapes)
        print("tuning every ring shape", file=sys.stderr, flush=True)
        ctx.tune()
        fresh = ctx.tune_table()
        check_rows(fresh, shapes)
        names = _lib.shape_names()
        # planning allocates nothing: budgets that hold the config-2 launch's
        # 2,048 slots of 140 MB for the two-per-SIMD form, and that do not
        for budget in (300e9, 150e9):
            assert ctx.lib.halign_set_traceback_budget(ctx.h, int(budget)) == 0
            for n, q, t, w in [(4096, 29500, 29903, 4665), (4096, 2000, 2100, 300), (8192, 300, 320, 230),
                               (2000, 1500, 1500, 100)]:
                got = ctx.plan_shape(n, q, t, w)
                want = greedy_pick(fresh, shapes, n, q, t, w, budget)
                print(f"plan {n} x ({q}, {t}, w {w}), budget {budget / 1e9:.0f} GB: shape {got} {names[got]}",
                      file=sys.stderr, flush=True)
                assert got == want, f"({q}, {t}, {w}): planner shape {got}, measured-best rule {want}"
            if budget < 200e9:   # 287 GB of slots do not fit: the one-per-SIMD form
                assert not names[ctx.plan_shape(4096, 29500, 29903, 4665)].endswith("d")
        # an unbanded pair whose query is the short side goes to a query-fixed
        # shape that holds the query: the measured-best one by the same rule
        qf = _lib.qfix_shapes()
        assert len([r for r in fresh if r["waves"] < 0]) == len(qf), "query-fixed shapes without a measurement"
        for n, q, t, w in [(4096, 400, 20000, -1), (6000, 12, 20000, -1), (100, 1024, 3000, 2999)]:
            got = ctx.plan_shape(n, q, t, w)
            print(f"plan {n} x ({q}, {t}, w {w}): shape {got} {names[got]}", file=sys.stderr, flush=True)
            assert names[got].startswith("q")
            g, s = qf[got - len(shapes)]
            assert (q + 15) // 16 <= g * s
    finally:
        ctx.close()


@pytest.mark.gpu
def test_bucket_timing_same_cigars():
    """Per-launch DP timing (a diagnostic mode that skips graph replay and
    brackets each launch with events) leaves every CIGAR as it is, and
    records launches whose jobs and cells add up."""
    import numpy as np
    import halign
    from seqgen import codes_to_ascii, mutate_codes, rand_codes
    rng = np.random.default_rng(21)
    center = rand_codes(rng, 8000)
    qs = []
    for i in range(300):
        if i % 10 == 0:
            qs.append(codes_to_ascii(rand_codes(rng, int(rng.integers(50, 400)))))   # anchor-less: whole pair
        else:
            a = int(rng.integers(0, 4000))
            qs.append(codes_to_ascii(mutate_codes(rng, center[a:a + int(rng.integers(500, 4000))], 0.02, 0.004)))
    ctx = halign.Context(0)
    try:
        c = halign.Center(ctx, codes_to_ascii(center))
        base, _, _ = c.align(qs)
        ctx.set_bucket_timing(True)
        timed, _, _ = c.align(qs)
        rows = ctx.last_buckets()
        ctx.set_bucket_timing(False)
        assert ctx.last_buckets() == []
    finally:
        ctx.close()
    assert all(np.array_equal(a, b) for a, b in zip(base, timed))
    print(f"bucket timing: {len(rows)} launches", file=sys.stderr, flush=True)
    assert rows and all(r["jobs"] > 0 and r["cells"] > 0 and r["kernel_ms"] > 0 for r in rows)
    assert any(r["family"] == 0 for r in rows) and any(r["family"] == 1 for r in rows)
