"""Keep-length FASTA on the GPU: halign_center_align_fasta (rows and records
made on the device) and halign_msa_keep_length (file to file, chunks
overlapped with reading and writing), against the reference-written runs of
tests/golden/msa_keep_length.npz (halign4 -c center --keep-length, one
thread, copy MSA command), then at a larger size against the oracle's merge
of the GPU's own CIGARs and against a single-chunk run."""
import os

import numpy as np
import pytest

from oracle_lib import Oracle
from test_msa_host import _msa_cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import halign
    c = halign.Context(0)
    yield c
    c.close()


def _write(tmp_path, name, data):
    p = str(tmp_path / name)
    with open(p, "wb") as f:
        f.write(data)
    return p


@pytest.mark.parametrize("case", range(5))
@pytest.mark.parametrize("batch", [0, 7])
def test_file_run_vs_reference(ctx, tmp_path, case, batch):
    import halign
    center, inp, expect = _msa_cases()[case]
    print(f"msa case {case} batch {batch}", flush=True)
    cp = _write(tmp_path, "c.fa", center)
    ip = _write(tmp_path, "q.fq" if case == 3 else "q.fa", inp)
    op = str(tmp_path / "out.fa")
    st = halign.msa_keep_length(ctx, cp, ip, op, batch=batch)
    got = open(op, "rb").read()
    assert got == expect
    assert st["bytes"] == len(expect)
    assert st["records"] == expect.count(b"\n>") + 1
    assert not os.path.exists(op + ".rest.tmp")


@pytest.mark.parametrize("case", [1, 4])
def test_align_fasta_records_vs_reference(ctx, case):
    """Center.align_fasta's two record groups are the reference's FASTA after
    its center record (the center aligned to is the consensus of the file,
    its index built from the record as written)."""
    import halign
    O = Oracle()
    center, inp, expect = _msa_cases()[case]
    (cid, cdesc, cseq), = O.fastx_parse(center)
    _, cons, _, _ = ctx.consensus(np.frombuffer(cseq, np.uint8)[None, :])
    recs = O.fastx_parse(inp)
    c = halign.Center(ctx, cons, index_seq=cseq)
    out, ins, nins, _ = c.align_fasta([q for _, _, q in recs], [n for n, _, _ in recs])
    body = expect[expect.index(b"\n>") + 1:]
    assert out == body
    # the split lies on a record boundary, after nins records
    assert ins == 0 or out[ins - 1:ins + 1] == b"\n>" or ins == len(out)
    assert out[:ins].count(b">") == nins
    c.close()


def test_chunked_run_matches_merge_of_gpu_cigars(ctx, tmp_path):
    """600 config-1 queries (16,569 bp center): the chunked file run equals
    (a) the oracle's merge of the CIGARs halign_center_align returns for the
    same queries and (b) one halign_center_align_fasta call over all of them."""
    import halign
    from halign import synth
    O = Oracle()
    center = synth.center(1)
    qb, off = synth.queries(1, 0, 600)
    qs = [bytes(qb[off[i]:off[i + 1]]) for i in range(600)]
    rng = np.random.default_rng(5)
    for i in range(0, 600, 3):   # substitutions only: rows without an insertion
        b = bytearray(center)
        for j in rng.integers(0, len(b), 30):
            b[j] = b"ACGT"[(b"ACGT".index(b[j]) + 1) % 4]
        qs[i] = bytes(b)
    names = [b"read_%d" % i for i in range(600)]
    text = b"".join(b">" + n + b" x\n" + q + b"\n" for n, q in zip(names, qs))
    cp = _write(tmp_path, "c.fa", b">cen\n" + center + b"\n")
    ip = _write(tmp_path, "q.fa", text)
    op = str(tmp_path / "o.fa")
    st = halign.msa_keep_length(ctx, cp, ip, op, batch=97)
    got = open(op, "rb").read()
    c = halign.Center(ctx, center)
    cigs, _, _ = c.align(qs)
    want = O.keep_length_fasta(b"cen", b"", center, qs, names, cigs)
    assert got == want
    one, ins, nins, _ = c.align_fasta(qs, names)
    assert got == b">cen\n" + center + b"\n" + one
    assert st["insertion_records"] == nins
    assert nins == sum(any((int(x) & 0xF) == 1 for x in cg) for cg in cigs)
    assert 0 < nins < 600
    c.close()


def test_errors(ctx, tmp_path):
    import halign
    one = _write(tmp_path, "c1.fa", b">a\n" + b"ACGT" * 100 + b"\n")
    q = _write(tmp_path, "q.fa", b">q\nACGTACGT\n")
    bad = _write(tmp_path, "bad.fq", b"@q\nACGT\n+\nII\n")
    out = str(tmp_path / "o.fa")
    with pytest.raises(halign.HalignError):
        halign.msa_keep_length(ctx, one, str(tmp_path / "absent.fa"), out)
    with pytest.raises(halign.HalignError, match="quality"):
        halign.msa_keep_length(ctx, one, bad, out)
    assert not os.path.exists(out + ".rest.tmp")
    # the context still works after the errors
    st = halign.msa_keep_length(ctx, one, q, out)
    assert st["records"] == 2
