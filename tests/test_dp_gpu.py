"""GPU parity of the DP kernel (ksw_extz2_sse semantics) against the oracle.

Every ksw_extz_t field and the CIGAR must be bit-identical to the oracle
(oracle/halign_oracle.c), which is itself pinned to the compiled reference
and the golden fixtures by tests/test_oracle.py.
"""
import sys

import numpy as np
import pytest

from oracle_lib import FLAG_LEFT, FLAG_RIGHT, Oracle, cigar_str
from seqgen import codes_to_ascii, mutate_codes, rand_codes

pytestmark = pytest.mark.gpu

FIELDS = ("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score")


# auto, multi-wave only, and forced ring shapes (waves, slots per lane) so that
# small cases also run through the multi-wave LDS exchange of the ring kernel;
# waves 0: sub-wave shapes (slots, lanes per job), forced where the band fits them;
# dense1x5: the two-per-SIMD form of the one-wave ring of five slots
POLICIES = ["auto", "multiwave", "wide", "dense1x5", (2, 1), (3, 2), (5, 1), (1, 5), (0, 1, 16), (0, 3, 16), (0, 1, 4),
            (0, 2, 4)]


@pytest.fixture(scope="module", params=POLICIES, ids=lambda p: p if isinstance(p, str) else (f"ring{p[0]}x{p[1]}" if p[0] else f"sub{p[2]}x{p[1]}"))
def ctx(request):
    import halign
    c = halign.Context(0)
    p = request.param
    named = {"auto": 0, "multiwave": 1, "wide": halign.wide_policy(), "dense1x5": halign.ring_policy(1, 5, dense=True)}
    pol = named[p] if isinstance(p, str) else halign.ring_policy(*p)
    c.set_kernel_policy(pol)
    print(f"kernel policy {p} -> {pol}", file=sys.stderr, flush=True)
    yield c
    c.close()


@pytest.fixture(scope="module")
def oracle():
    return Oracle()


def _tasks(cases):
    import halign
    qc, tc = [], []
    tasks = np.zeros(len(cases), halign.DP_TASK_DTYPE)
    qo = to = 0
    for i, (q, t, w, flag) in enumerate(cases):
        tasks[i] = (len(q), len(t), w, flag, qo, to)
        qc.append(q)
        tc.append(t)
        qo += len(q)
        to += len(t)
    return tasks, np.concatenate(qc) if qc else np.zeros(0, np.uint8), \
        np.concatenate(tc) if tc else np.zeros(0, np.uint8)


def _compare(oracle, cases, res, cigs, full=True, label=""):
    bad = 0
    for i, (q, t, w, flag) in enumerate(cases):
        ez, cig, cells = oracle.ksw(q, t, w, flag)
        got = res[i]
        names = FIELDS if full else ("zdropped", "score", "mte", "mte_q")
        exp = dict(zip(("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q",
                        "score", "n_cigar"), ez))
        ok = all(int(got[f]) == int(exp[f]) for f in names) and np.array_equal(cigs[i], cig)
        ok = ok and int(got["cells"]) == cells
        if not ok:
            bad += 1
            if bad <= 3:
                print(f"[{label}] case {i} qlen={len(q)} tlen={len(t)} w={w} flag={flag}",
                      {f: (int(got[f]), int(exp[f])) for f in names}, "cells", int(got["cells"]),
                      cells, "\n gpu:", cigar_str(cigs[i])[:120], "\n ora:", cigar_str(cig)[:120],
                      file=sys.stderr, flush=True)
    return bad


def _random_cases(rng, n):
    cases = []
    for it in range(n):
        kind = it % 6
        if kind == 0:
            t = rand_codes(rng, int(rng.integers(1, 80)))
            q = rand_codes(rng, int(rng.integers(1, 80)))
        elif kind == 1:
            t = rand_codes(rng, int(rng.integers(1, 300)), p_n=0.05)
            q = mutate_codes(rng, t, 0.1, 0.05)
        elif kind == 2:
            t = rand_codes(rng, int(rng.integers(50, 900)))
            q = mutate_codes(rng, t, 0.05, 0.02)
        elif kind == 3:
            t = rand_codes(rng, int(rng.integers(1, 200)), alphabet=2)
            q = rand_codes(rng, int(rng.integers(1, 200)), alphabet=2)
        elif kind == 4:
            t = rand_codes(rng, int(rng.integers(100, 1500)))
            q = mutate_codes(rng, t, 0.2, 0.05)
        else:
            t = rand_codes(rng, int(rng.integers(1, 40)))
            q = rand_codes(rng, int(rng.integers(1, 40)))
        if len(q) == 0:
            q = rand_codes(rng, 1)
        flag = FLAG_RIGHT if rng.random() < 0.6 else FLAG_LEFT
        r = rng.random()
        if r < 0.3:
            w = -1
        elif r < 0.6:
            w = int(rng.integers(0, 48))
        else:
            from oracle_lib import Oracle as _O  # auto band as globalAlignKSW2 computes it
            w = _O().auto_band(len(t), len(q))
        cases.append((q, t, w, flag))
    return cases


def test_dp_random_batched_full_stats(ctx, oracle):
    rng = np.random.default_rng(1234)
    cases = _random_cases(rng, 1200)
    tasks, qc, tc = _tasks(cases)
    print("dp batched full-stats: 1200 random cases", file=sys.stderr, flush=True)
    res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=True, plain=False)
    assert _compare(oracle, cases, res, cigs, True, "batched-stats") == 0


def test_dp_random_batched_cigar_mode(ctx, oracle):
    rng = np.random.default_rng(99)
    cases = _random_cases(rng, 1200)
    tasks, qc, tc = _tasks(cases)
    print("dp batched cigar-mode: 1200 random cases", file=sys.stderr, flush=True)
    res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
    # in CIGAR mode max/mqe are only exact where the band overflows
    assert _compare(oracle, cases, res, cigs, False, "batched-cigar") == 0


def test_dp_plain_form(ctx, oracle):
    rng = np.random.default_rng(7)
    cases = _random_cases(rng, 60)
    tasks, qc, tc = _tasks(cases)
    print("dp plain: 60 random cases", file=sys.stderr, flush=True)
    res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=True, plain=True)
    assert _compare(oracle, cases, res, cigs, True, "plain") == 0


def test_dp_long_banded_pairs(ctx, oracle):
    """Whole-pair sized DPs with auto band: several ring shapes."""
    rng = np.random.default_rng(5)
    cases = []
    for n in (3000, 8000, 16569):
        t = rand_codes(rng, n)
        q = mutate_codes(rng, t, 0.01, 0.002)
        cases.append((q, t, oracle.auto_band(len(t), len(q)), FLAG_RIGHT))
    t = rand_codes(rng, 5000)
    q = mutate_codes(rng, t[:2600], 0.01, 0.002)
    cases.append((q, t, oracle.auto_band(len(t), len(q)), FLAG_RIGHT))   # band overflow
    tasks, qc, tc = _tasks(cases)
    print("dp long pairs", [(len(c[0]), len(c[1]), c[2]) for c in cases], file=sys.stderr,
          flush=True)
    res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=True, plain=False)
    assert _compare(oracle, cases, res, cigs, True, "long") == 0


def test_dp_asymmetric_unbanded(ctx, oracle):
    """Tall-thin and short-wide pairs with w = -1 (auto_band's ratio > 0.5):
    the band slides along the whole longer side (the empty-chain whole pair of
    a short query)."""
    rng = np.random.default_rng(31)
    cases = []
    for it in range(60):
        lt = int(rng.integers(300, 5000))
        ls = int(rng.integers(1, 200))
        t = rand_codes(rng, lt)
        a = int(rng.integers(0, lt - 1))
        q = mutate_codes(rng, t[a:a + ls], 0.05, 0.02) if it % 3 else rand_codes(rng, ls)
        if len(q) == 0:
            q = rand_codes(rng, 1)
        flag = FLAG_RIGHT if it % 2 else FLAG_LEFT
        cases.append((q, t, -1, flag) if it % 4 < 2 else (t, q, -1, flag))
    tasks, qc, tc = _tasks(cases)
    print("dp asymmetric unbanded: 60 cases", file=sys.stderr, flush=True)
    for full in (False, True):
        res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=full, plain=False)
        assert _compare(oracle, cases, res, cigs, full, "asym") == 0


def test_global_align_batch(ctx, oracle):
    rng = np.random.default_rng(11)
    refs, qrys, flags = [], [], []
    for i in range(300):
        t = rand_codes(rng, int(rng.integers(0, 400)), p_n=0.01)
        q = mutate_codes(rng, t, 0.05, 0.02) if i % 4 else rand_codes(rng, int(rng.integers(0, 50)))
        refs.append(codes_to_ascii(t))
        qrys.append(codes_to_ascii(q).lower() if i % 7 == 0 else codes_to_ascii(q))
        flags.append(FLAG_RIGHT if i % 3 else FLAG_LEFT)
    print("global_align_batch: 300 pairs", file=sys.stderr, flush=True)
    cigs, scores = ctx.global_align_batch(refs, qrys, flags)
    bad = 0
    for i in range(300):
        exp, _ = oracle.global_align_ksw2(refs[i], qrys[i], flags[i])
        if not np.array_equal(exp, cigs[i]):
            bad += 1
    assert bad == 0


def test_pairset_whole_pairs(ctx, oracle):
    """Resident center + queries (config-2 sized pairs), device-side query prep."""
    import halign
    from halign import synth
    center = synth.center(2)
    qb, off = synth.queries(2, 0, 3)
    print("pairset: 3 whole pairs of ~29.5 kb", file=sys.stderr, flush=True)
    ps = halign.PairSet(ctx, center, qb, off)
    ms, cells = ps.run(0, 3)
    for i in range(3):
        q = qb[off[i]:off[i + 1]]
        exp, ecells = oracle.global_align_ksw2(center, q)
        sc, cig = ps.result(i)
        assert np.array_equal(cig, exp), i
    ps.close()


def _golden(name):
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_dp_golden_fixtures(ctx):
    """The reference's own outputs (tests/golden/ksw_cases.npz, written from
    the compiled reference): every ksw_extz_t field in full-stats mode, score,
    mte and CIGAR in CIGAR mode."""
    import halign
    g = _golden("ksw_cases.npz")
    n = len(g["w"])
    tasks = np.zeros(n, halign.DP_TASK_DTYPE)
    tasks["qlen"] = np.diff(g["q_off"])
    tasks["tlen"] = np.diff(g["t_off"])
    tasks["w"] = g["w"]
    tasks["flag"] = g["flag"]
    tasks["query_off"] = g["q_off"][:-1]
    tasks["target_off"] = g["t_off"][:-1]
    print(f"golden ksw cases: {n}", file=sys.stderr, flush=True)
    for full in (True, False):
        res, cigs, ms = ctx.dp_batch(tasks, g["q"], g["t"], full_stats=full, plain=False)
        names = FIELDS if full else ("zdropped", "score", "mte", "mte_q")
        bad = 0
        for i in range(n):
            exp = dict(zip(g["ez_fields"], g["ez"][i]))
            ok = all(int(res[i][f]) == int(exp[f]) for f in names)
            ok = ok and np.array_equal(cigs[i], g["cigar"][g["cigar_off"][i]:g["cigar_off"][i + 1]])
            bad += not ok
        assert bad == 0, (full, bad)


def test_pairset_golden_whole_pairs(ctx):
    """20 whole config-2 pairs against the reference's CIGARs
    (tests/golden/ksw_whole_pairs.npz)."""
    import halign
    from halign import synth
    g = _golden("ksw_whole_pairs.npz")
    n = int(g["count"])
    center = synth.center(int(g["config"]))
    qb, off = synth.queries(int(g["config"]), int(g["first"]), n)
    assert np.array_equal(off, g["query_off"])
    print(f"golden whole pairs: {n}", file=sys.stderr, flush=True)
    ps = halign.PairSet(ctx, center, qb, off)
    ps.run(0, n)
    for i in range(n):
        sc, cig = ps.result(i)
        assert np.array_equal(cig, g["cigar"][g["cigar_off"][i]:g["cigar_off"][i + 1]]), i
    ps.close()


UNIFORM_SHAPES = [(16, 1, 16), (16, 2, 16), (16, 3, 16), (4, 1, 4), (4, 2, 4)]   # (lanes, slots, max side)


@pytest.mark.parametrize("lanes,slots,_", UNIFORM_SHAPES, ids=lambda v: str(v))
def test_dp_uniform_tickets_forced(oracle, lanes, slots, _):
    """The uniform-ticket form of each sub-wave shape, forced: jobs in groups
    of equal (qlen, tlen, band) whose sizes are not multiples of the jobs per
    wave (the library pads each group with repeats), different sequences in
    every job, unbanded and auto-band, both gap directions, N bases; every
    result equals the oracle's."""
    import halign
    ctx = halign.Context(0)
    try:
        ctx.set_kernel_policy(halign.ring_policy(0, slots, lanes, uniform=True))
        rng = np.random.default_rng(100 * lanes + slots)
        side = 16 * (lanes * slots - 3)   # the shorter side's limit for n_col <= ring - 2
        cases = []
        for key in range(6):
            ls = int(rng.integers(1, max(2, side)))
            lt = int(rng.integers(ls, 6 * ls + 50))
            w = -1 if key % 2 else oracle.auto_band(lt, ls)
            flag = FLAG_RIGHT if key % 3 else FLAG_LEFT
            for j in range(int(rng.integers(1, 8))):
                t = rand_codes(rng, lt, p_n=0.01 if key == 4 else 0.0)
                a = int(rng.integers(0, lt - ls + 1))
                q = mutate_codes(rng, t[a:a + ls], 0.05, 0.0)[:ls]
                if len(q) < ls:
                    q = np.concatenate([q, rand_codes(rng, ls - len(q))])
                cases.append((q, t, w, flag) if key < 3 else (t, q, w, flag))
        tasks, qc, tc = _tasks(cases)
        print(f"uniform tickets {lanes}x{slots}: {len(cases)} jobs", file=sys.stderr, flush=True)
        res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
        assert _compare(oracle, cases, res, cigs, False, f"uniform{lanes}x{slots}") == 0
    finally:
        ctx.close()


def test_dp_uniform_tickets_automatic(oracle):
    """Automatic policy: whole pairs of short queries against one long center
    (unbanded, many queries per length, as the anchor-less queries of the
    uneven workload) go to the uniform-ticket form; results equal the oracle's."""
    import halign
    ctx = halign.Context(0)
    try:
        rng = np.random.default_rng(77)
        t = rand_codes(rng, 3000)
        cases = []
        for ls in (180, 300, 420):
            for j in range(16):   # 8 per gap direction: whole tickets
                a = int(rng.integers(0, len(t) - ls))
                q = mutate_codes(rng, t[a:a + ls], 0.02, 0.004)
                q = np.concatenate([q, rand_codes(rng, max(0, ls - len(q)))])[:ls]
                cases.append((q, t, -1, FLAG_RIGHT if j % 2 else FLAG_LEFT))
        tasks, qc, tc = _tasks(cases)
        res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
        assert _compare(oracle, cases, res, cigs, False, "uniform-auto") == 0
    finally:
        ctx.close()


def test_dp_wide_bands(oracle):
    """Bands wider than the widest ring (n_col_ > 1,535 blocks) run on the
    wide-band kernel, which ksw_extz2_sse allows for any w and length
    (src/align/ksw2_extz2_sse.c:72-76): an unbanded ~24,650 x 49,400 pair
    (n_col_ 1,542) both ways round and with N bases; and, forced onto that
    kernel, banded pairs of several passes per anti-diagonal (n_col_ 564 and
    1,100) including a band overflow. Every field and CIGAR equals the oracle's."""
    import halign
    rng = np.random.default_rng(2600)
    big_t = rand_codes(rng, 49400)
    big_q = mutate_codes(rng, big_t[20000:44650], 0.02, 0.004)   # ~24,650: n_col_ 1,542
    big_qn = big_q.copy()
    big_qn[rng.integers(0, len(big_qn), 40)] = 4
    auto_cases = [(big_q, big_t, -1, FLAG_RIGHT), (big_t, big_qn, -1, FLAG_LEFT)]
    t1 = rand_codes(rng, 9500)
    q1 = mutate_codes(rng, t1, 0.03, 0.01)
    t2 = rand_codes(rng, 18000)
    q2 = mutate_codes(rng, t2, 0.02, 0.004)
    t3 = rand_codes(rng, 12000)
    q3 = mutate_codes(rng, t3[:6100], 0.02, 0.004)
    forced_cases = [(q1, t1, 8999, FLAG_RIGHT), (q2, t2, 17590, FLAG_LEFT), (q3, t3, 2000, FLAG_RIGHT)]
    class Once:   # the oracle's answer per case, computed once for both modes
        def __init__(self):
            self.memo = {}

        def ksw(self, q, t, w, flag):
            k = (id(q), id(t), w, flag)
            if k not in self.memo:
                self.memo[k] = oracle.ksw(q, t, w, flag)
            return self.memo[k]
    once = Once()
    for label, cases, pol in (("wide auto", auto_cases, 0), ("wide forced", forced_cases, halign.wide_policy())):
        ctx = halign.Context(0)
        try:
            ctx.set_kernel_policy(pol)
            tasks, qc, tc = _tasks(cases)
            print(label, [(len(c[0]), len(c[1]), c[2]) for c in cases], file=sys.stderr, flush=True)
            for full in (True, False):
                res, cigs, ms = ctx.dp_batch(tasks, qc, tc, full_stats=full, plain=False)
                assert _compare(once, cases, res, cigs, full, label) == 0
        finally:
            ctx.close()
