"""SAM output and the second alignment to the consensus against the
reference program's own per-thread SAM files (sam_cases.npz, written by
halign4 -t 1 --save-workdir): RefAligner::alignQueryToRef with one thread
(src/align/ref_aligner.cpp:286-394) writes every query to thread0.sam or
thread0_insertion.sam by cigar::hasInsertion of its result (:212-283), as
writeSamRecord formats it (:128-136). Cases: one center, three references,
four references with missing columns and their aligned file, each with and
without --keep-length (without it and with several references, the result is
a second alignment to the consensus, seeded with the consensus's own
minimizers, :268-282)."""
import os
import sys
import tempfile

import numpy as np
import pytest

from oracle_lib import parse_cigar_string

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    import halign
    c = halign.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def cases():
    with np.load(os.path.join(GOLD, "sam_cases.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def seg(g, k, i):
    return g[k][g[k + "_off"][i]:g[k + "_off"][i + 1]].tobytes()


def write_case(g, i, d):
    cp, ap, ip = (os.path.join(d, "%s%d.fa" % (x, i)) for x in ("c", "a", "q"))
    open(cp, "wb").write(seg(g, "center", i))
    open(ip, "wb").write(seg(g, "input", i))
    aligned = seg(g, "aligned", i)
    if aligned:
        open(ap, "wb").write(aligned)
    return cp, (ap if aligned else None), ip


def sam_lines(text):
    out = {}
    for line in text.split(b"\n")[1:]:
        if line:
            f = line.split(b"\t")
            out[f[0]] = (f[2], parse_cigar_string(f[5]))
    return out


@pytest.mark.parametrize("case", range(6))
def test_sam_files_match_reference_program(ctx, cases, case):
    """halign_align_sam writes the reference program's two SAM files byte
    for byte, in chunks of 16 queries."""
    import halign
    keep, has_aligned, nref, _ = (int(x) for x in cases["par"][case])
    d = tempfile.mkdtemp()
    cp, ap, ip = write_case(cases, case, d)
    sp, si = os.path.join(d, "t0.sam"), os.path.join(d, "t0_ins.sam")
    print(f"sam case {case}: {nref} reference(s), keep_length {keep}", file=sys.stderr, flush=True)
    st = halign.align_sam(ctx, cp, ip, sp, si, keep_length=keep, batch=16, aligned_path=ap)
    assert open(sp, "rb").read() == seg(cases, "sam", case)
    assert open(si, "rb").read() == seg(cases, "sam_ins", case)
    assert st["records"] == 42


@pytest.mark.parametrize("kind", [1, 2])
def test_refset_first_and_second_alignment(ctx, cases, kind):
    """halign_refset_align without keep_length returns both alignments: the
    first (to the chosen reference) equals the keep-length run's SAM CIGAR,
    the second (to the consensus) the other run's, and the SAM group follows
    the second; with missing columns some queries change group."""
    import halign

    def read_records(path):
        return [(i, q) for i, _, q in halign.read_fastx(path)]
    d = tempfile.mkdtemp()
    cp, ap, ip = write_case(cases, 2 * kind + 1, d)
    refs = read_records(cp)
    aligned = read_records(ap) if ap else refs
    cons = halign.consensus_file(ctx, ap or cp, os.path.join(d, "cons.fa"), os.path.join(d, "cons.json"))
    queries = read_records(ip)
    rs = halign.RefSet(ctx, [s for _, s in refs], cons)
    try:
        best, first, result, hins, fref = rs.align([s for _, s in queries], keep_length=False)
    finally:
        rs.close()
    keep_sam = {**sam_lines(seg(cases, "sam", 2 * kind)), **sam_lines(seg(cases, "sam_ins", 2 * kind))}
    plain = sam_lines(seg(cases, "sam", 2 * kind + 1))
    ins = sam_lines(seg(cases, "sam_ins", 2 * kind + 1))
    names = [refs[r][0] for r in range(len(refs))]
    flips = 0
    for i, (name, _) in enumerate(queries):
        rname, c1 = keep_sam[name]
        assert names[best[i]] == rname and np.array_equal(first[i], c1), name
        _, c2 = plain[name] if name in plain else ins[name]
        assert np.array_equal(result[i], c2) and bool(hins[i]) == (name in ins) and fref[i] == -1, name
        flips += bool(np.any((first[i] & 0xF) == 1)) != bool(hins[i])
    if kind == 2:
        assert flips >= 3
