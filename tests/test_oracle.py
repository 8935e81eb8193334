"""The oracle (oracle/halign_oracle.c) pinned to the reference, on the CPU.

Three kinds of pins, per function:
  * golden fixtures written by tests/golden/make_golden.py from the reference
    compiled from its own sources (oracle/_ref): bit-exact;
  * the reference's own known-answer tests (test/test_align.cpp,
    test/test_anchor.cpp, test/test_consensus.cpp), restated here;
  * what the mathematics fixes independently of both: a textbook Gotoh global
    DP (full band) for the ksw2 score and for the score of the returned CIGAR,
    and the lengths every global CIGAR must consume.
When oracle/_ref is present (the container that has the reference tree), the
oracle is also run against it live on fresh seeded inputs.
"""
import os

import numpy as np
import pytest

from oracle_lib import (ANCHOR_DTYPE, FLAG_LEFT, FLAG_RIGHT, Oracle, RefLib, cigar_str, pack_hits,
                        ref_available)
from seqgen import codes_to_ascii, mutate_codes, rand_codes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}   # decompress each array once


def seg(data, off, i):
    return data[off[i]:off[i + 1]]


@pytest.fixture(scope="module")
def O():
    return Oracle()


# ---------------------------------------------------------------- ksw2 / DP --

def test_ksw_golden(O):
    g = load("ksw_cases.npz")
    n = len(g["w"])
    bad = []
    for i in range(n):
        q, t = seg(g["q"], g["q_off"], i), seg(g["t"], g["t_off"], i)
        ez, cig, _ = O.ksw(q, t, int(g["w"][i]), int(g["flag"][i]))
        if tuple(int(v) for v in g["ez"][i]) != tuple(ez) or \
                not np.array_equal(cig, seg(g["cigar"], g["cigar_off"], i)):
            bad.append(i)
    assert n >= 3000 and not bad, f"{len(bad)} of {n} differ, first {bad[:5]}"
    # the set covers band overflow, both gap directions and N
    assert (g["ez"][:, 1] == 1).sum() > 50
    assert set(np.unique(g["flag"])) == {FLAG_LEFT, FLAG_RIGHT}
    assert (g["q"] == 4).any() and (g["t"] == 4).any()


def test_ksw_golden_discriminates(O):
    """A flipped gap direction or a narrower band must show in the fixtures:
    they would catch an oracle (or kernel) with that mistake."""
    g = load("ksw_cases.npz")
    flipped = narrowed = 0
    for i in range(600):
        q, t = seg(g["q"], g["q_off"], i), seg(g["t"], g["t_off"], i)
        w, flag = int(g["w"][i]), int(g["flag"][i])
        exp = seg(g["cigar"], g["cigar_off"], i)
        _, c1, _ = O.ksw(q, t, w, FLAG_LEFT if flag == FLAG_RIGHT else FLAG_RIGHT)
        flipped += not np.array_equal(c1, exp)
        if w > 2:
            e2, c2, _ = O.ksw(q, t, w - 2, flag)
            narrowed += (not np.array_equal(c2, exp)) or tuple(e2) != tuple(int(v) for v in g["ez"][i])
    assert flipped > 50 and narrowed > 10


def test_ksw_whole_pairs_golden(O):
    """Whole 29.9 kb pairs (auto band ~4,665) regenerated from csrc/synth.c."""
    from halign import synth
    g = load("ksw_whole_pairs.npz")
    center = synth.center(int(g["config"]))
    assert len(center) == int(g["center_len"])
    qb, off = synth.queries(int(g["config"]), int(g["first"]), 3)
    assert np.array_equal(off, g["query_off"][:4])
    for i in range(3):
        cig, cells = O.global_align_ksw2(center, qb[off[i]:off[i + 1]], FLAG_RIGHT)
        assert np.array_equal(cig, seg(g["cigar"], g["cigar_off"], i)), i
        assert cells > 2.5e8


def test_global_align_golden(O):
    g = load("global_align.npz")
    n = len(g["flag"])
    for i in range(n):
        ref = seg(g["ref"], g["ref_off"], i).tobytes()
        qry = seg(g["qry"], g["qry_off"], i).tobytes()
        cig, _ = O.global_align_ksw2(ref, qry, int(g["flag"][i]))
        assert np.array_equal(cig, seg(g["cigar"], g["cigar_off"], i)), (i, len(ref), len(qry))


def test_ksw_known_answers(O):
    """test/test_align.cpp:234-245 (12M), :260-271 (insertion shows as I),
    :295-302 (empty sides do not fail)."""
    cig, _ = O.global_align_ksw2(b"ACGTACGTACGT", b"ACGTACGTACGT")
    assert cigar_str(cig) == "12M"
    cig, _ = O.global_align_ksw2(b"ACGTACGTACGT", b"ACGTAACGTACGT")
    assert "I" in cigar_str(cig)
    cig, _ = O.global_align_ksw2(b"", b"ACGT")
    assert cigar_str(cig) == "4I"
    cig, _ = O.global_align_ksw2(b"ACGT", b"")
    assert cigar_str(cig) == "4D"


MATCH, MISMATCH, GAPO, GAPE = 5, -4, 6, 2


def gotoh(q, t):
    """Textbook Gotoh global alignment score (dna5: +5 / -4 / N 0; gap of
    length l costs 6 + 2l), independent of ksw2's difference recurrence."""
    n, m = len(t), len(q)
    NEG = -10 ** 9
    H = np.full((n + 1, m + 1), NEG, np.int64)
    E = np.full((n + 1, m + 1), NEG, np.int64)   # gap in t direction (consumes q)
    F = np.full((n + 1, m + 1), NEG, np.int64)   # consumes t
    H[0, 0] = 0
    for j in range(1, m + 1):
        E[0, j] = H[0, j] = -(GAPO + GAPE * j)
    for i in range(1, n + 1):
        F[i, 0] = H[i, 0] = -(GAPO + GAPE * i)
        for j in range(1, m + 1):
            E[i, j] = max(E[i, j - 1] - GAPE, H[i, j - 1] - GAPO - GAPE)
            F[i, j] = max(F[i - 1, j] - GAPE, H[i - 1, j] - GAPO - GAPE)
            a, b = int(t[i - 1]), int(q[j - 1])
            s = 0 if (a == 4 or b == 4) else (MATCH if a == b else MISMATCH)
            H[i, j] = max(H[i - 1, j - 1] + s, E[i, j], F[i, j])
    return int(H[n, m])


def cigar_score(cig, q, t):
    sc = i = j = 0
    for c in cig:
        ln, op = int(c) >> 4, int(c) & 15
        if op == 0:
            for _ in range(ln):
                a, b = int(t[i]), int(q[j])
                sc += 0 if (a == 4 or b == 4) else (MATCH if a == b else MISMATCH)
                i += 1
                j += 1
        elif op == 1:
            sc -= GAPO + GAPE * ln
            j += ln
        elif op == 2:
            sc -= GAPO + GAPE * ln
            i += ln
    return sc, i, j


def test_ksw_full_band_is_gotoh_optimum(O):
    """Full band (w = -1): ez.score is the Gotoh optimum and the CIGAR is an
    optimal global path (its own score equals it, it consumes both sides)."""
    rng = np.random.default_rng(77)
    for it in range(120):
        t = rand_codes(rng, int(rng.integers(1, 60)), p_n=0.05 if it % 3 == 0 else 0.0)
        q = mutate_codes(rng, t, 0.1, 0.1) if it % 2 else rand_codes(rng, int(rng.integers(1, 60)))
        if len(q) == 0:
            q = rand_codes(rng, 1)
        best = gotoh(q, t)
        for flag in (FLAG_RIGHT, FLAG_LEFT):
            ez, cig, _ = O.ksw(q, t, -1, flag)
            assert ez[8] == best, (it, flag)
            sc, ti, qi = cigar_score(cig, q, t)
            assert (sc, ti, qi) == (best, len(t), len(q)), (it, flag, cigar_str(cig))


def one_gap_closed_form(R, Q, right):
    """The one-gap optimum of globalAlignKSW2 under a full band (auto_band < 0),
    from the scores alone: a shorter side equal to the longer side's tail or
    head aligns as one gap run of |n - m| and s = min(n, m) matching ACGT
    pairs, and every other path scores less. The optimal paths are
    M^a gap M^(s-a) for a up to the common prefix (tail case) or from s minus
    the common suffix (head case); ksw2 puts the gap leftmost without
    KSW_EZ_RIGHT and rightmost with it. None when neither case holds."""
    n, m = len(R), len(Q)
    acgt = b"ACGT"
    if n == m:
        return None
    g, s = abs(n - m), min(n, m)
    gop = g << 4 | (2 if m < n else 1)
    long_, short = (R, Q) if m < n else (Q, R)
    tail = all(short[k] in acgt and short[k] == long_[g + k] for k in range(s))
    head = all(short[k] in acgt and short[k] == long_[k] for k in range(s))
    lcp = lcs = 0
    while lcp < s and R[lcp] in acgt and R[lcp] == Q[lcp]:
        lcp += 1
    while lcs < s and R[n - 1 - lcs] in acgt and R[n - 1 - lcs] == Q[m - 1 - lcs]:
        lcs += 1
    if tail:
        a = lcp if right else 0
    elif head:
        a = s if right else s - lcs
    else:
        return None
    return [x for x in (a << 4, gop, (s - a) << 4) if x >> 4]


def test_one_gap_closed_form(O):
    """Segments whose optimum is one gap run get their CIGAR from the closed
    form above, without the DP: the form against the oracle's DP on both
    flags, including the ties (shared bases at the far end, 1- and 2-letter
    alphabets) where only the flag picks the path, insertions (roles
    swapped) and near misses (first base changed) that must not qualify."""
    rng = np.random.default_rng(5)
    A = b"ACGT"
    tot = 0
    for it in range(6000):
        kind = it % 6
        m, g = int(rng.integers(1, 30)), int(rng.integers(1, 300))
        alph = {1: 2, 4: 2, 5: 1}.get(kind, 4)
        n = m + g
        R = bytes(A[int(x)] for x in rng.integers(0, alph, n))
        Q = R[n - m:] if it % 2 == 0 else R[:m]
        if kind == 3:
            Q = bytes([A[int(rng.integers(0, 4))]]) + Q[1:]
        if it % 3 == 0:
            R, Q = Q, R
        if O.auto_band(len(R), len(Q)) >= 0:
            continue
        for flag in (FLAG_LEFT, FLAG_RIGHT):
            exp = one_gap_closed_form(R, Q, flag == FLAG_RIGHT)
            if exp is None:
                continue
            tot += 1
            cig, _ = O.global_align_ksw2(R, Q, flag)
            assert list(cig) == exp, (it, flag, R[:12], Q[:12], cigar_str(cig), cigar_str(exp))
    assert tot > 8000


def test_band_overflow_never_consumes(O):
    """A segment whose auto band leaves the matrix before the last
    anti-diagonal (ksw2 sets zdropped and backtracks from its maximum) gets a
    CIGAR that does not consume both sides, so globalAlignMM2 always replaces
    it by I(query) D(ref) and the device skips its DP. Checked on the oracle's
    globalAlignKSW2 over related and unrelated pairs, both flags, every shape
    found to overflow."""
    rng = np.random.default_rng(404)
    A = b"ACGT"
    seen = 0
    for it in range(400):
        n = int(rng.integers(60, 2500))
        m = int(rng.integers(max(1, n // 2), n))
        if it % 2:
            n, m = m, n          # longer query (insertions)
        w = O.auto_band(n, m)    # ref length n, query length m
        if w < 0:
            continue
        r = m + n - 2            # last anti-diagonal; ksw2's qlen = m, tlen = n
        if not (w == 0 or r - m + 1 > (r + w) >> 1 or ((r - w + 1) >> 1) > n - 1):
            continue
        R = bytes(A[int(x)] for x in rng.integers(0, 4, n))
        if it % 3 == 0:          # the query a piece of the reference
            a = int(rng.integers(0, max(1, n - m + 1)))
            Q = (R[a:a + m] + bytes(A[int(x)] for x in rng.integers(0, 4, m)))[:m]
        else:
            Q = bytes(A[int(x)] for x in rng.integers(0, 4, m))
        for flag in (FLAG_RIGHT, FLAG_LEFT):
            cig, _ = O.global_align_ksw2(R, Q, flag)
            _, ti, qi = cigar_score(cig, np.frombuffer(Q, np.uint8) * 0, np.frombuffer(R, np.uint8) * 0)
            assert (ti, qi) != (n, m), (it, n, m, w, flag)
            seen += 1
    assert seen > 100


def test_banded_cigars_consume_both_sides(O):
    """test/test_align.cpp:1418-1626: a global CIGAR consumes the whole query
    and reference, also under a narrow band (no overflow)."""
    g = load("ksw_cases.npz")
    for i in range(len(g["w"])):
        if int(g["ez"][i][1]):
            continue
        cig = seg(g["cigar"], g["cigar_off"], i)
        q, t = seg(g["q"], g["q_off"], i), seg(g["t"], g["t_off"], i)
        _, ti, qi = cigar_score(cig, q, t)
        assert (ti, qi) == (len(t), len(q)), i


# ----------------------------------------------------------------- seeding --

def test_minimizers_golden(O):
    g = load("minimizers.npz")
    n = len(g["k"])
    for i in range(n):
        s = seg(g["seq"], g["seq_off"], i).tobytes()
        x, y, ub = O.minimizers(s, int(g["k"][i]), int(g["w"][i]), bool(g["non_canonical"][i]))
        assert ub & 3 == 0
        assert np.array_equal(x, seg(g["x"], g["hit_off"], i)), i
        assert np.array_equal(y, seg(g["y"], g["hit_off"], i)), i
    assert n > 1000


def test_anchors_chain_mm2_golden(O):
    g = load("anchors_chain.npz")
    n = len(g["ref_off"]) - 1
    nonempty = 0
    for i in range(n):
        rx, ry = seg(g["rx"], g["rx_off"], i), seg(g["ry"], g["rx_off"], i)
        qx, qy = seg(g["qx"], g["qx_off"], i), seg(g["qy"], g["qx_off"], i)
        exp_a = seg(g["anchors"], g["anchors_off"] * ANCHOR_DTYPE.itemsize, i).view(ANCHOR_DTYPE)
        a = O.collect_anchors(rx, ry, qx, qy)
        assert np.array_equal(a, exp_a), i
        srt, chain = O.chain(a)
        assert np.array_equal(srt, exp_a[seg(g["sorted"], g["sorted_off"], i)]), i
        exp_c = seg(g["chain"], g["chain_off"] * ANCHOR_DTYPE.itemsize, i).view(ANCHOR_DTYPE)
        assert np.array_equal(chain, exp_c), i
        nonempty += len(chain) > 0
        ref = seg(g["ref"], g["ref_off"], i).tobytes()
        qry = seg(g["qry"], g["qry_off"], i).tobytes()
        cig, _, _ = O.mm2(ref, qry, a)
        assert np.array_equal(cig, seg(g["mm2"], g["mm2_off"], i)), i
    assert nonempty > n // 2


def test_ref_global_align_golden(O):
    g = load("ref_global.npz")
    n = len(g["ref_off"]) - 1
    for i in range(n):
        ref = seg(g["ref"], g["ref_off"], i).tobytes()
        qry = seg(g["qry"], g["qry_off"], i).tobytes()
        cig, _, _, ub = O.ref_global_align(ref, qry)
        assert np.array_equal(cig, seg(g["cigar"], g["cigar_off"], i)), (i, len(ref), len(qry))


NOFILTER = dict(q_occ_frac=0.0, f_top_frac=0.0)


def test_collect_anchors_known_answers(O):
    """test/test_anchor.cpp:60-247."""
    ca = O.collect_anchors_ex
    e = pack_hits([])
    assert len(ca(*e, *e)) == 0
    assert len(ca(*e, *pack_hits([(0x111111, 100)]))) == 0
    assert len(ca(*pack_hits([(0x111111, 100)]), *e)) == 0
    a = ca(*pack_hits([(0x123456, 100, 0, False, 20)]), *pack_hits([(0x123456, 50, 0, False, 20)]),
           **NOFILTER)
    assert len(a) == 1 and a[0]["hash"] == 0x123456 and a[0]["pos_ref"] == 100
    assert a[0]["pos_qry"] == 50 and a[0]["span"] == 20 and a[0]["is_rev"] == 0
    a = ca(*pack_hits([(0xABCDEF, 200, 0, False, 15)]), *pack_hits([(0xABCDEF, 80, 0, True, 15)]),
           **NOFILTER)
    assert len(a) == 1 and a[0]["is_rev"] == 1
    assert len(ca(*pack_hits([(0x111111, 100), (0x222222, 200)]),
                  *pack_hits([(0x333333, 50), (0x444444, 80)]))) == 0
    a = ca(*pack_hits([(0x555555, 100), (0x555555, 200), (0x555555, 300)]),
           *pack_hits([(0x555555, 50)]), **NOFILTER)
    assert len(a) == 3 and set(a["pos_qry"]) == {50} and sorted(a["pos_ref"]) == [100, 200, 300]
    q = [(0x888888, i) for i in range(50)] + [(0x999900 + i, i) for i in range(50)]
    assert len(ca(*pack_hits([(0x888888, 100)]), *pack_hits(q), q_occ_frac=0.10,
                  f_top_frac=0.0)) == 0
    a = ca(*pack_hits([(0xAAAAAA, 100, 0, False, 30)]), *pack_hits([(0xAAAAAA, 50, 0, False, 20)]),
           **NOFILTER)
    assert len(a) == 1 and a[0]["span"] == 20
    a = ca(*pack_hits([(0xBBBBBB, 100, 2, False, 15)]), *pack_hits([(0xBBBBBB, 50, 3, False, 15)]),
           **NOFILTER)
    assert len(a) == 1 and a[0]["rid_ref"] == 2 and a[0]["rid_qry"] == 3


def anchors(rows):
    a = np.zeros(len(rows), ANCHOR_DTYPE)
    for i, (h, pr, pq, rr, rq, rev, span) in enumerate(rows):
        a[i] = (h, rr, pr, rq, pq, span, int(rev))
    return a


def test_chain_known_answers(O):
    """test/test_anchor.cpp:260-482."""
    ch = O.chain_ex
    assert len(ch(anchors([]))[1]) == 0
    _, c = ch(anchors([(0x111111, 100, 50, 0, 0, False, 20)]), min_cnt=1, min_score=10)
    assert len(c) == 1 and c[0]["span"] == 20
    two = [(0x111111, 100, 50, 0, 0, False, 20), (0x222222, 150, 100, 0, 0, False, 20)]
    _, c = ch(anchors(two), min_cnt=2, min_score=30)
    assert len(c) == 2 and c[0]["pos_ref"] < c[1]["pos_ref"]
    _, c = ch(anchors(two + [(0x555555, 200, 150, 0, 0, False, 20),
                             (0x333333, 200, 150, 1, 0, False, 15),
                             (0x444444, 250, 200, 1, 0, False, 15)]), min_cnt=2, min_score=20)
    assert len(c) >= 2 and c[0]["rid_ref"] == 0
    _, c = ch(anchors(two + [(0x666666, 200, 150, 0, 0, False, 20),
                             (0x333333, 300, 250, 0, 0, True, 15),
                             (0x444444, 350, 300, 0, 0, True, 15)]), min_cnt=2, min_score=20)
    assert len(c) >= 2 and c[0]["is_rev"] == 0
    _, c = ch(anchors([(0x111111, 100, 50, 0, 0, False, 20), (0x222222, 6000, 5500, 0, 0, False, 20)]),
              min_cnt=1, min_score=10, max_dist_x=5000)
    assert len(c) == 1
    _, c = ch(anchors([(0x111111, 100, 50, 0, 0, False, 20), (0x222222, 250, 150, 0, 0, False, 20)]),
              min_cnt=1, min_score=10, bw=30)
    assert len(c) == 1
    _, c = ch(anchors(two + [(0x333333, 200, 150, 0, 0, False, 20),
                             (0x444444, 10000, 5000, 0, 0, False, 15),
                             (0x555555, 10100, 5100, 0, 0, False, 15)]), min_cnt=2, min_score=20)
    assert len(c) == 3
    _, c = ch(anchors([(0x111111, 100, 50, 0, 0, False, 15), (0x222222, 150, 100, 0, 0, False, 15)]),
              min_cnt=3, min_score=10)
    assert len(c) == 0
    _, c = ch(anchors([(0x111111, 100, 50, 0, 0, False, 5), (0x222222, 150, 100, 0, 0, False, 5)]),
              min_cnt=2, min_score=50)
    assert len(c) == 0
    _, c = ch(anchors([(0x111111, 100, 50, 0, 0, False, 20), (0x222222, 200, 150, 0, 0, False, 25)]),
              min_cnt=2, min_score=30)
    assert [(int(x["pos_ref"]), int(x["pos_qry"])) for x in c] == [(100, 50), (200, 150)]


def test_chain_log2_table(O):
    """The chain penalty's log2 values come from a table of the C library's
    log2f (the reference calls std::log2(float), src/seed/anchor.cpp:227-230)."""
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.log2f.argtypes = [ctypes.c_float]
    libm.log2f.restype = ctypes.c_float
    assert O.lut[0] == 0.0
    exp = np.array([libm.log2f(float(d)) for d in range(1, 502)], np.float32)
    assert np.array_equal(O.lut[1:], exp)


# --------------------------------------------------------------- consensus --

def test_consensus_golden(O):
    g = load("consensus.npz")
    for ci, (row_set, lim) in enumerate(g["case"]):
        nrows, aln = (int(v) for v in g["shape"][row_set])
        rows = seg(g["rows"], g["rows_off"], row_set).reshape(nrows, aln)
        counts, cons, n = O.consensus(rows, seq_limit=int(lim))
        assert np.array_equal(counts.reshape(-1), seg(g["counts"], g["counts_off"], ci)), ci
        assert cons == seg(g["cons"], g["cons_off"], ci).tobytes(), ci
        assert n == int(g["num_seqs"][ci]), ci


def rows_of(*seqs):
    return np.array([np.frombuffer(s, np.uint8) for s in seqs])


def test_consensus_known_answers(O):
    """test/test_consensus.cpp:200-349."""
    assert O.consensus(rows_of(b"ACGT-", b"AC-T-", b"ACGT-"))[1] == b"ACGTA"
    assert O.consensus(rows_of(b"A", b"C"))[1] == b"A"
    assert O.consensus(rows_of(b"---", b"---", b"---"))[1] == b"AAA"
    assert O.consensus(rows_of(b"ACGTACGT"))[1] == b"ACGTACGT"
    assert O.consensus(rows_of(b"UAAAA", b"TAAAA", b"UAAAA"))[1][:1] == b"U"
    five = rows_of(b"A", b"A", b"C", b"C", b"C")
    assert O.consensus(five)[1] == b"C"
    assert O.consensus(five, seq_limit=2)[1] == b"A"


def test_consensus_is_a_histogram(O):
    """Closed form on one batch: counts are the 7-way column histogram and the
    consensus is its A>C>G>T>U argmax (all-zero -> A)."""
    rng = np.random.default_rng(3)
    alpha = np.frombuffer(b"ACGTUN-.acgtuRY", np.uint8)
    rows = alpha[rng.integers(0, len(alpha), (300, 50))]
    counts, cons, n = O.consensus(rows)
    cls = {**{c: i for i, c in enumerate(b"ACGTU")}, **{c: i for i, c in enumerate(b"acgtu")},
           ord("-"): 6, ord("."): 6}
    exp = np.zeros((50, 7), np.uint32)
    for r in rows:
        for j, c in enumerate(r):
            exp[j, cls.get(int(c), 5)] += 1
    assert n == 300 and np.array_equal(counts, exp)
    best = b"".join(b"ACGTU"[int(np.argmax(exp[j, :5]))].to_bytes(1, "little")
                    if exp[j, :5].max() > 0 else b"A" for j in range(50))
    assert cons == best


# ------------------------------------------------ live check against _ref --

@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (no reference tree here)")
def test_oracle_matches_compiled_reference_live(O):
    R = RefLib()
    rng = np.random.default_rng(20261203)
    for it in range(300):
        t = rand_codes(rng, int(rng.integers(1, 400)), p_n=0.02)
        q = mutate_codes(rng, t, 0.08, 0.04)
        if len(q) == 0:
            q = rand_codes(rng, 1)
        w = [-1, int(rng.integers(0, 30)), O.auto_band(len(t), len(q))][it % 3]
        flag = FLAG_RIGHT if it % 2 else FLAG_LEFT
        e1, c1, _ = O.ksw(q, t, w, flag)
        e2, c2 = R.ksw(q, t, w, flag)
        assert e1 == e2 and np.array_equal(c1, c2), it
    for it in range(40):
        rc = rand_codes(rng, int(rng.integers(50, 3000)))
        ref = codes_to_ascii(rc)
        qry = codes_to_ascii(mutate_codes(rng, rc, 0.03, 0.01))   # codes, not ASCII: no T read as N
        assert b"T" in qry and b"N" not in qry
        assert np.array_equal(O.ref_global_align(ref, qry)[0], R.ref_global_align(ref, qry)), it


# ------------------------------------------- workload samples (configs 3-5) --

@pytest.mark.parametrize("cfg", [3, 4, 5])
def test_oracle_matches_reference_program_on_workloads(O, cfg):
    """The oracle's RefAligner::globalAlign against the CIGARs the reference
    program wrote for the workload samples (config<c>.npz): the first 40
    queries and 40 spread over the sample (forced cases included), and the
    reference's SAM group equals hasInsertion of its CIGAR."""
    from halign import synth
    from workloads import workload_queries
    with np.load(os.path.join(GOLD, "config%d.npz" % cfg), allow_pickle=False) as z:
        f = {k: z[k] for k in z.files}
    n = len(f["idx"])
    pick = np.unique(np.concatenate([np.arange(40), np.linspace(40, n - 1, 40).astype(int)]))
    q, off = workload_queries(cfg, f["idx"][pick])
    center = synth.center(cfg)
    co = f["cigar_off"]
    for j, k in enumerate(pick):
        exp = f["cigar"][co[k]:co[k + 1]]
        assert bool(f["has_insertion"][k]) == bool(np.any((exp & 0xF) == 1)), k
        got = O.ref_global_align(center, q[off[j]:off[j + 1]])[0]
        assert np.array_equal(got, exp), (cfg, int(f["idx"][k]))


def test_oracle_cigar_to_string_matches_reference_sam():
    """or_cigar_to_string (cigarToString, src/align/cigar.cpp:82-97) against
    the CIGAR fields of the reference program's own SAM files (sam_cases.npz):
    every field parsed with the BAM op numbering and written back must give
    the reference's text."""
    import ctypes as C
    import os
    import numpy as np
    from oracle_lib import Oracle, parse_cigar_string
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sam_cases.npz")
    L = Oracle().lib
    L.or_cigar_to_string.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int]
    n = 0
    with np.load(gold, allow_pickle=False) as z:
        for key in ("sam", "sam_ins"):
            blob, off = z[key].tobytes(), z[key + "_off"]
            for k in range(len(off) - 1):
                for line in blob[off[k]:off[k + 1]].split(b"\n"):
                    if not line or line.startswith(b"@"):
                        continue
                    field = line.split(b"\t")[5]
                    if field == b"*":
                        continue
                    cig = parse_cigar_string(field)
                    buf = C.create_string_buffer(len(field) + 8)
                    got = L.or_cigar_to_string(cig.ctypes.data_as(C.c_void_p), len(cig), buf, len(buf))
                    assert got == len(field) and buf.value == field
                    n += 1
    assert n > 100
