"""The keep-length FASTA run on the CPU side: FASTA/FASTQ reading and the
merge of rows into the final FASTA.

Pins:
  * fastx_cases.npz: seq_io::KseqReader's records (or its error) for fixed
    edge cases, 400 random texts and two texts larger than the readers'
    buffers, written by the reference (tests/golden/make_golden.py). Checked
    against the oracle's restatement and against the library's reader (host
    code of libhalign_hip.so, no GPU), plain and gzip.
  * msa_keep_length.npz: halign4 -c center --keep-length run by the
    reference's own readers, aligners and writers. The oracle recomputes it
    from the files: its parser, the consensus of the center, its per-query
    path with the minimizers of the center as written, and its merge; the
    bytes must agree.
"""
import gzip
import os

import numpy as np
import pytest

from oracle_lib import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fx():
    z = np.load(os.path.join(GOLDEN, "fastx_cases.npz"))
    out = []
    r = 0
    for i in range(len(z["status"])):
        text = z["text"][z["text_off"][i]:z["text_off"][i + 1]].tobytes()
        if z["status"][i] != 0:
            out.append((text, int(z["status"][i])))
            continue
        recs = []
        for _ in range(int(z["nrec"][i])):
            recs.append(tuple(z[k][z[k + "_off"][r]:z[k + "_off"][r + 1]].tobytes() for k in ("id", "desc", "seq")))
            r += 1
        out.append((text, recs))
    return out


@pytest.fixture(scope="module")
def fx_cases():
    return _fx()


def test_fastx_fixture_shape(fx_cases):
    assert len(fx_cases) > 400
    errs = sum(isinstance(e, int) for _, e in fx_cases)
    big = [t for t, _ in fx_cases if len(t) > 2_000_000]
    assert errs >= 2 and len(big) == 2   # malformed FASTQ cases, and texts beyond one read buffer


def test_fastx_oracle_vs_reference(fx_cases):
    O = Oracle()
    for i, (text, expect) in enumerate(fx_cases):
        got = O.fastx_parse(text)
        if isinstance(expect, int):
            assert got == -2, i
        else:
            assert got == expect, i


@pytest.mark.parametrize("gz", [False, True])
def test_fastx_library_reader_vs_reference(fx_cases, tmp_path, gz):
    import halign
    for i, (text, expect) in enumerate(fx_cases):
        if gz and len(text) < 1000 and i % 4:
            continue   # every big text, and a quarter of the small ones, compressed
        path = str(tmp_path / ("t.fa.gz" if gz else "t.fa"))
        with (gzip.open(path, "wb") if gz else open(path, "wb")) as f:
            f.write(text)
        try:
            got = list(halign.read_fastx(path, chunk=1 + i % 5))
        except halign.HalignError:
            got = -2
        if isinstance(expect, int):
            assert got == -2, i
        else:
            assert got == expect, i


def test_fastx_missing_file(tmp_path):
    import halign
    with pytest.raises(halign.HalignError):
        list(halign.read_fastx(str(tmp_path / "absent.fa")))


def _msa_cases():
    z = np.load(os.path.join(GOLDEN, "msa_keep_length.npz"))
    return [tuple(z[k][z[k + "_off"][i]:z[k + "_off"][i + 1]].tobytes() for k in ("center", "input", "output"))
            for i in range(len(z["center_off"]) - 1)]


def oracle_keep_length(O, center_text, input_text):
    """The run restated with the oracle: records, consensus center, per-query
    CIGARs (minimizers of the center as written), merge."""
    (cid, cdesc, cseq), = O.fastx_parse(center_text)
    rows = np.frombuffer(cseq, np.uint8)[None, :]
    _, cons, _ = O.consensus(rows)
    rx, ry, _ = O.minimizers(cseq)
    hits = (rx, ry) if len(rx) else None
    recs = O.fastx_parse(input_text)
    cigs = [O.ref_global_align(cons, q, ref_hits=hits)[0] for _, _, q in recs]
    return O.keep_length_fasta(cid, cdesc, cseq, [q for _, _, q in recs], [n for n, _, _ in recs], cigs), recs, cigs


@pytest.mark.parametrize("case", range(5))
def test_keep_length_oracle_vs_reference(case):
    O = Oracle()
    center, inp, expect = _msa_cases()[case]
    got, recs, cigs = oracle_keep_length(O, center, inp)
    assert got == expect
    # the case covers both SAM groups and the dropped '@' name
    has_ins = [any((int(c) & 0xF) == 1 for c in cg) for cg in cigs]
    assert any(has_ins) and not all(has_ins)
    assert any(n.startswith(b"@") for n, _, _ in recs)
    assert b"\n>@" not in expect


def test_keep_length_planted_mistakes_fail():
    """Swapping the two groups, keeping the '@' query, or not cleaning the
    center each change the bytes of a reference-written case."""
    O = Oracle()
    center, inp, expect = _msa_cases()[1]
    got, recs, cigs = oracle_keep_length(O, center, inp)
    assert got == expect
    (cid, cdesc, cseq), = O.fastx_parse(center)
    raw_center = b">" + cid + b" " + cdesc + b"\n" + cseq + b"\n"
    assert not expect.startswith(raw_center)   # the center record is cleaned (lower case, n)
    names = [n for n, _, _ in recs]
    queries = [q for _, _, q in recs]
    renamed = [b"x" + n if n.startswith(b"@") else n for n in names]
    assert O.keep_length_fasta(cid, cdesc, cseq, queries, renamed, cigs) != expect
    # the reference's record order: queries with an insertion first, each group
    # in input order, the '@' query absent
    out_names = [n for n, _, _ in O.fastx_parse(expect)][1:]
    has_ins = [any((int(c) & 0xF) == 1 for c in cg) for cg in cigs]
    first = [n for n, h in zip(names, has_ins) if h and not n.startswith(b"@")]
    rest = [n for n, h in zip(names, has_ins) if not h and not n.startswith(b"@")]
    assert out_names == first + rest
    assert out_names != [n for n in names if not n.startswith(b"@")]


def test_keep_length_center_index_is_the_written_center():
    """RefAligner aligns to the consensus of the center file but seeds with
    the minimizers of the record as written; in case 4 (scattered N, queries
    following the consensus) seeding from the consensus changes the output."""
    O = Oracle()
    center, inp, expect = _msa_cases()[4]
    (cid, cdesc, cseq), = O.fastx_parse(center)
    _, cons, _ = O.consensus(np.frombuffer(cseq, np.uint8)[None, :])
    recs = O.fastx_parse(inp)
    cigs = [O.ref_global_align(cons, q)[0] for _, _, q in recs]
    wrong = O.keep_length_fasta(cid, cdesc, cseq, [q for _, _, q in recs], [n for n, _, _ in recs], cigs)
    assert wrong != expect
    assert oracle_keep_length(O, center, inp)[0] == expect
