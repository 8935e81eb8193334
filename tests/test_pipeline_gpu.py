"""GPU parity of the per-query path against a resident center
(RefAligner::globalAlign, src/align/ref_aligner.cpp:91-125): anchors
(collect_anchors), chains (chainAnchors) and the final CIGAR (globalAlignMM2),
bit-exact against the reference-written fixtures and against the oracle on
seeded batches."""
import os
import sys

import numpy as np
import pytest

from oracle_lib import ANCHOR_DTYPE, Oracle, cigar_str
from seqgen import codes_to_ascii, mutate_codes, rand_codes

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def seg(d, o, i):
    return d[o[i]:o[i + 1]]


@pytest.fixture(scope="module")
def ctx():
    import halign
    c = halign.Context(0)
    yield c
    c.close()


def test_anchors_chains_mm2_golden(ctx):
    """Each fixture case is its own center: anchors, the chain and the
    globalAlignMM2 CIGAR of the reference (anchors_chain.npz)."""
    import halign
    g = load("anchors_chain.npz")
    n = len(g["ref_off"]) - 1
    rec = ANCHOR_DTYPE.itemsize
    print(f"anchors/chains/mm2 golden: {n} cases", file=sys.stderr, flush=True)
    bad = {"anchors": [], "chain": [], "mm2": []}
    for i in range(n):
        ref = seg(g["ref"], g["ref_off"], i).tobytes()
        qry = seg(g["qry"], g["qry_off"], i).tobytes()
        c = halign.Center(ctx, ref)
        try:
            exp_a = g["anchors"][g["anchors_off"][i] * rec:g["anchors_off"][i + 1] * rec].view(ANCHOR_DTYPE)
            got_a = c.anchors([qry])[0]
            if not np.array_equal(got_a.view(np.uint8), exp_a.view(np.uint8)):
                bad["anchors"].append(i)
            exp_c = g["chain"][g["chain_off"][i] * rec:g["chain_off"][i + 1] * rec].view(ANCHOR_DTYPE)
            got_c = c.chains([qry])[0]
            if not np.array_equal(got_c.view(np.uint8), exp_c.view(np.uint8)):
                bad["chain"].append(i)
            cig, st, _ = c.align([qry])
            exp = seg(g["mm2"], g["mm2_off"], i)
            if not np.array_equal(cig[0], exp):
                bad["mm2"].append(i)
                if len(bad["mm2"]) <= 3:
                    direct, _ = ctx.global_align_batch([ref], [qry], [0x06])
                    print(f"mm2 case {i}: ref {len(ref)} query {len(qry)} chain {len(exp_c)} stats {st[0]}\n"
                          f"  gpu    {cigar_str(cig[0])[:160]}\n  expect {cigar_str(exp)[:160]}\n"
                          f"  direct globalAlignKSW2 {cigar_str(direct[0])[:160]}", file=sys.stderr, flush=True)
        finally:
            c.close()
    assert not any(bad.values()), {k: (len(v), v[:8]) for k, v in bad.items()}


def test_ref_global_golden(ctx):
    """RefAligner::globalAlign fixtures (ref_global.npz): 150 short pairs, each
    its own center, and 30 config-1 queries against the 16,569 bp center in
    one batch."""
    import halign
    g = load("ref_global.npz")
    n = len(g["ref_off"]) - 1
    print(f"ref_global golden: {n} cases", file=sys.stderr, flush=True)
    groups = {}
    for i in range(n):
        groups.setdefault(seg(g["ref"], g["ref_off"], i).tobytes(), []).append(i)
    bad = []
    for ref, idx in groups.items():
        c = halign.Center(ctx, ref)
        try:
            cigs, st, _ = c.align([seg(g["qry"], g["qry_off"], i).tobytes() for i in idx])
            bad += [i for j, i in enumerate(idx) if not np.array_equal(cigs[j], seg(g["cigar"], g["cigar_off"], i))]
        finally:
            c.close()
    assert not bad, (len(bad), bad[:10])


def test_align_batch_vs_oracle(ctx):
    """One 12 kb center, 300 queries of 0-15 kb: exact windows, mutated
    substrings, extensions with random ends, long indels (large segments and
    band overflows), queries too short for anchors, N runs."""
    import halign
    O = Oracle()
    rng = np.random.default_rng(77)
    center = rand_codes(rng, 12000)
    cstr = codes_to_ascii(center)
    qs = []
    for i in range(300):
        kind = i % 6
        if kind == 0:
            a = int(rng.integers(0, 11000))
            q = mutate_codes(rng, center[a:a + int(rng.integers(1, 4000))], 0.01, 0.002)
        elif kind == 1:
            q = mutate_codes(rng, center, 0.05, 0.01)
        elif kind == 2:
            q = np.concatenate([rand_codes(rng, int(rng.integers(0, 800))), mutate_codes(rng, center, 0.01, 0.002),
                                rand_codes(rng, int(rng.integers(0, 800)))])
        elif kind == 3:
            q = center.copy()
            a = int(rng.integers(100, 11000))
            q = np.concatenate([q[:a], rand_codes(rng, int(rng.integers(50, 500))), q[a + int(rng.integers(0, 500)):]])
            q = mutate_codes(rng, q, 0.01, 0.002)
        elif kind == 4:
            q = rand_codes(rng, int(rng.integers(0, 500)))
        else:
            q = mutate_codes(rng, center, 0.02, 0.005)
            q[rng.random(len(q)) < 0.001] = 4
        qs.append(codes_to_ascii(q))
    print("align batch vs oracle: 300 queries, 12 kb center", file=sys.stderr, flush=True)
    c = halign.Center(ctx, cstr)
    cigs, st, ms = c.align(qs)
    bad = []
    for i, q in enumerate(qs):
        exp, _, _, ub = O.ref_global_align(cstr, q)
        if ub & 3:
            continue
        if not np.array_equal(cigs[i], exp):
            bad.append(i)
    assert not bad, f"{len(bad)} differ, first {bad[:5]}"
    c.close()


@pytest.mark.parametrize("clen", [8000, 40000, 120000])
def test_anchors_center_sizes_vs_oracle(ctx, clen):
    """collect_anchors on the device against the oracle for centers whose
    distinct-hash table fits the kernel's LDS with the hashes (8 kb), only
    its per-query counters (40 kb) and neither (120 kb: counters in global
    scratch); queries with a repeated block exercise the query-occurrence
    filter (q_occ_frac) and a low-complexity center stretch the sampled
    high-occurrence hashes."""
    import halign
    O = Oracle()
    rng = np.random.default_rng(clen)
    center = np.concatenate([rand_codes(rng, clen - 2000), np.tile(rand_codes(rng, 20), 100)])
    cstr = codes_to_ascii(center)
    rx, ry, _ = O.minimizers(cstr)
    qs = []
    for i in range(40):
        a = int(rng.integers(0, len(center) - 2500))
        q = mutate_codes(rng, center[a:a + int(rng.integers(200, 2500))], 0.01, 0.002)
        if i % 3 == 0:
            q = np.concatenate([q, np.tile(q[:60], 6)])
        qs.append(codes_to_ascii(q))
    qs.append(codes_to_ascii(center[-2000:]))
    print(f"anchors vs oracle: center {clen} bp, {len(qs)} queries", file=sys.stderr, flush=True)
    c = halign.Center(ctx, cstr)
    try:
        got = c.anchors(qs)
    finally:
        c.close()
    bad = []
    for i, q in enumerate(qs):
        qx, qy, _ = O.minimizers(q)
        exp = O.collect_anchors(rx, ry, qx, qy)
        if not np.array_equal(got[i].view(np.uint8), exp.view(np.uint8)):
            bad.append(i)
    assert not bad, f"{len(bad)} differ, first {bad[:5]}"


def test_chains_vs_oracle(ctx):
    """chainAnchors on the device against the oracle's, anchor for anchor, on a
    center with tandem repeats (many equal-hash anchors: long predecessor
    windows over several 64-candidate chunks, improving and skipped
    candidates interleaved, max_skip breaks) and plain random stretches;
    queries cut from both, mutated, with swapped blocks."""
    import halign
    O = Oracle()
    rng = np.random.default_rng(123)
    unit = rand_codes(rng, 180)
    rep = np.concatenate([mutate_codes(rng, unit, 0.02, 0.0) for _ in range(25)])
    center = np.concatenate([rand_codes(rng, 3000), rep, rand_codes(rng, 3000), np.tile(unit[:37], 40)])
    cstr = codes_to_ascii(center)
    rx, ry, _ = O.minimizers(cstr)
    qs = []
    for i in range(160):
        a = int(rng.integers(0, len(center) - 300))
        q = mutate_codes(rng, center[a:a + int(rng.integers(300, 6000))], 0.02 if i % 2 else 0.005, 0.004)
        if i % 5 == 0 and len(q) > 600:   # two blocks swapped
            h = len(q) // 2
            q = np.concatenate([q[h:], q[:h]])
        qs.append(codes_to_ascii(q))
    print(f"chains vs oracle: {len(qs)} queries, {len(center)} bp center with repeats", file=sys.stderr, flush=True)
    c = halign.Center(ctx, cstr)
    try:
        got = c.chains(qs)
    finally:
        c.close()
    bad, long_chains = [], 0
    for i, q in enumerate(qs):
        qx, qy, _ = O.minimizers(q)
        _, exp = O.chain(O.collect_anchors(rx, ry, qx, qy))
        long_chains += len(exp) > 64
        if not np.array_equal(got[i].view(np.uint8), exp.view(np.uint8)):
            bad.append(i)
    assert not bad, f"{len(bad)} differ, first {bad[:5]}"
    assert long_chains > 20


def test_large_batch_staged_upload(ctx):
    """A batch of ~19 MB of queries goes up through the pinned staging buffers
    in 8 MiB chunks (a query straddles each chunk boundary); a query's CIGAR
    does not depend on its batch, so every one must equal the CIGAR of the
    same query aligned in small batches (direct uploads), and a sample of them
    the oracle's."""
    import halign
    O = Oracle()
    rng = np.random.default_rng(2024)
    center = rand_codes(rng, 12000)
    cstr = codes_to_ascii(center)
    qs = []
    for i in range(1700):
        a = int(rng.integers(0, 1500))
        qs.append(codes_to_ascii(mutate_codes(rng, center[a:a + int(rng.integers(9000, 12000))], 0.01, 0.002)))
    total = sum(len(q) for q in qs)
    print(f"staged upload: {len(qs)} queries, {total / 1e6:.1f} MB in one batch", file=sys.stderr, flush=True)
    assert total > 2 * (8 << 20)
    c = halign.Center(ctx, cstr)
    try:
        big, _, _ = c.align(qs)
        small = []
        for k in range(0, len(qs), 150):
            cg, _, _ = c.align(qs[k:k + 150])
            small += cg
    finally:
        c.close()
    bad = [i for i in range(len(qs)) if not np.array_equal(big[i], small[i])]
    assert not bad, f"{len(bad)} differ between the batch sizes, first {bad[:5]}"
    for i in range(0, len(qs), 170):
        exp, _, _, ub = O.ref_global_align(cstr, qs[i])
        if not ub & 3:
            assert np.array_equal(big[i], exp), i


def test_one_gap_segments_vs_oracle(ctx):
    """Queries whose end segments have a one-gap optimum (CIGAR written
    without the DP) and near misses that must run the DP: exact windows of the
    center (a long reference part against the query bases before the first
    anchor), windows whose first or last bases equal the center's first or
    last bases (ties, broken by the segment's gap direction), the same with
    only the edge base set to the center's (near misses), windows reaching
    past the center's ends (insertions against a short reference part), and
    N at the edges. The oracle runs every segment through its DP."""
    import halign
    O = Oracle()
    rng = np.random.default_rng(91)
    center = rand_codes(rng, 8000)
    cstr = codes_to_ascii(center)
    # window starts sharing the center's first 3 bases, ends sharing its last 3
    starts = [p for p in range(1, 7000) if np.array_equal(center[p:p + 3], center[:3])]
    ends = [e for e in range(3000, 7999) if np.array_equal(center[e - 3:e], center[-3:])]
    qs = []
    for i in range(300):
        kind = i % 10
        a = int(rng.integers(0, 7000))
        L = int(rng.integers(200, 3000))
        q = center[a:a + L].copy()
        if kind == 1:
            q[0] = center[0]              # tie at the front
        elif kind == 2:
            q[-1] = center[-1]            # tie at the back
        elif kind == 3:                   # runs past the center's start
            q = np.concatenate([rand_codes(rng, int(rng.integers(1, 40))), center[:L]])
        elif kind == 4:                   # runs past its end
            q = np.concatenate([center[-L:], rand_codes(rng, int(rng.integers(1, 40)))])
        elif kind == 5:
            q[0] = 4                      # N at the front
        elif kind == 6:
            q = mutate_codes(rng, q, 0.01, 0.002)
        elif kind == 7:                   # past the start, with the shared first base
            q = np.concatenate([center[:1], rand_codes(rng, int(rng.integers(1, 40))), center[:L]])
        elif kind == 8:                   # tie: starts like the center
            a = int(rng.choice(starts))
            q = center[a:a + L].copy()
        elif kind == 9:                   # tie: ends like the center
            e = int(rng.choice(ends))
            q = center[e - L:e].copy()
        qs.append(codes_to_ascii(q))
    print(f"one-gap segments vs oracle: {len(qs)} queries, 8 kb center, {len(starts)} tie starts, "
          f"{len(ends)} tie ends", file=sys.stderr, flush=True)
    c = halign.Center(ctx, cstr)
    cigs, st, ms = c.align(qs)
    bad = []
    for i, q in enumerate(qs):
        exp, _, _, ub = O.ref_global_align(cstr, q)
        if ub & 3:
            continue
        if not np.array_equal(cigs[i], exp):
            bad.append(i)
    assert not bad, f"{len(bad)} differ, first {bad[:5]}"
    c.close()


def test_rows_and_insertion_flags(ctx):
    """delQueryToRefByCigar rows and hasInsertion from the device, against the
    oracle's projection of the same CIGARs (and the CIGARs against the oracle's
    RefAligner::globalAlign), on queries with and without insertions, N bases,
    lower case and queries without anchors (whole-pair CIGAR)."""
    import halign
    O = Oracle()
    rng = np.random.default_rng(21)
    center = rand_codes(rng, 5000)
    cstr = codes_to_ascii(center)
    qs = []
    for i in range(120):
        a = int(rng.integers(0, 4000))
        q = mutate_codes(rng, center[a:a + int(rng.integers(1, 1000))], 0.02, 0.01 if i % 3 else 0.0)
        if i % 10 == 0:
            q = rand_codes(rng, int(rng.integers(0, 200)))
        s = codes_to_ascii(q)
        if i % 7 == 0 and s:
            s = s[:len(s) // 2].lower() + s[len(s) // 2:]
        qs.append(s)
    c = halign.Center(ctx, cstr)
    try:
        cigs, rows, hins = c.align_rows(qs)
    finally:
        c.close()
    print(f"rows: {len(qs)} queries, {int(hins.sum())} with insertions", file=sys.stderr, flush=True)
    bad = []
    for i, q in enumerate(qs):
        exp_cig, _, _, ub = O.ref_global_align(cstr, q)
        if ub & 3:
            continue
        exp_row = O.del_query_to_ref(q, cigs[i])
        if not np.array_equal(cigs[i], exp_cig) or rows[i] != exp_row or hins[i] != any((int(x) & 0xf) == 1 for x in cigs[i]):
            bad.append(i)
    assert not bad, f"{len(bad)} differ, first {bad[:5]}"
    assert 0 < int(hins.sum()) < len(qs)


def test_whole_pair_wider_than_every_ring_vs_oracle(ctx):
    """An anchor-less query of ~24,650 bases against a 49,400-base center:
    globalAlignMM2 falls back to the whole pair (src/align/align.cpp:204-206)
    with auto_band -1 (length ratio above 0.5), so its band is n_col_ 1,542
    blocks, wider than every ring kernel; it runs on the wide-band kernel
    inside the batch, beside ordinary queries, and equals the oracle's CIGAR."""
    import halign
    rng = np.random.default_rng(49400)
    center = codes_to_ascii(rand_codes(rng, 49400))
    lone = codes_to_ascii(rand_codes(rng, 24650))
    cc = np.frombuffer(center, np.uint8)
    lut = np.zeros(256, np.uint8)
    lut[np.frombuffer(b"ACGT", np.uint8)] = np.arange(4, dtype=np.uint8)
    near = [codes_to_ascii(mutate_codes(rng, lut[cc[a:a + 3000]], 0.01, 0.002)) for a in (100, 20000, 46000)]
    qs = [near[0], lone, near[1], near[2]]
    O = Oracle()
    print("wide whole pair: 49,400 center, 24,650 anchor-less query", file=sys.stderr, flush=True)
    exp = [O.ref_global_align(center, q)[0] for q in qs]
    c = halign.Center(ctx, center)
    try:
        got, st, _ = c.align(qs)
    finally:
        c.close()
    for i in range(len(qs)):
        assert np.array_equal(got[i], exp[i]), (i, cigar_str(got[i])[:200], cigar_str(exp[i])[:200])
