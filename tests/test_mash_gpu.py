"""Several references on the GPU: halign_mash_sketch and halign_mash_best
against the reference-written mash_cases.npz, and the multi-reference
keep-length file run (halign_msa_keep_length_refs) against msa_refs.npz."""
import os

import numpy as np
import pytest

from test_mash_host import mash_fixture, msa_refs_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import halign
    c = halign.Context(0)
    yield c
    c.close()


def test_sketches_vs_reference(ctx):
    cases, _, _, _, _ = mash_fixture()
    groups = {}
    for i, (seq, p, s) in enumerate(cases):
        groups.setdefault(p, []).append(i)
    for (k, size, nc, seed), idx in groups.items():
        print(f"sketch k={k} size={size} nc={nc} seed={seed}: {len(idx)} sequences", flush=True)
        got = ctx.mash_sketch([cases[i][0] for i in idx], k, size, bool(nc), seed)
        for i, g in zip(idx, got):
            assert np.array_equal(g, cases[i][2]), i


def test_reference_choice_vs_reference(ctx):
    _, refs, qs, best, jac = mash_fixture()
    rsk = ctx.mash_sketch(refs)
    qsk = ctx.mash_sketch(qs)
    b, j = ctx.mash_best(qsk, rsk)
    assert np.array_equal(b, best)
    assert np.array_equal(j, jac)


def test_large_batch_of_sketches_matches_per_sequence(ctx):
    """Enough k-mers that the sketch pass splits (32 Mi keys per pass): the
    sketches equal those of the sequences sketched one call each."""
    rng = np.random.default_rng(11)
    seqs = [np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, int(n))].tobytes()
            for n in rng.integers(200_000, 400_000, 120)]
    many = ctx.mash_sketch(seqs)
    for i in range(0, 120, 17):
        assert np.array_equal(many[i], ctx.mash_sketch([seqs[i]])[0])
    assert all(len(s) == 2000 for s in many)


@pytest.mark.parametrize("case", range(3))
@pytest.mark.parametrize("batch", [0, 5])
def test_file_run_refs_vs_reference(ctx, tmp_path, case, batch):
    import halign
    center, aligned, inp, expect = msa_refs_fixture()[case]
    print(f"refs case {case} batch {batch}", flush=True)
    cp = str(tmp_path / "c.fa")
    open(cp, "wb").write(center)
    ap = None
    if aligned is not None:
        ap = str(tmp_path / "a.fa")
        open(ap, "wb").write(aligned)
    ip = str(tmp_path / ("q.fq" if case == 2 else "q.fa"))
    open(ip, "wb").write(inp)
    op = str(tmp_path / "o.fa")
    st = halign.msa_keep_length(ctx, cp, ip, op, batch=batch, aligned_path=ap)
    assert open(op, "rb").read() == expect
    assert st["bytes"] == len(expect)
    assert not os.path.exists(op + ".rest.tmp")


def test_refs_errors(ctx, tmp_path):
    import halign
    q = str(tmp_path / "q.fa")
    open(q, "wb").write(b">q\nACGTACGTAC\n")
    uneq = str(tmp_path / "u.fa")
    open(uneq, "wb").write(b">a\n" + b"ACGT" * 30 + b"\n>b\n" + b"ACGT" * 31 + b"\n")
    with pytest.raises(halign.HalignError, match="differ in length"):
        halign.msa_keep_length(ctx, uneq, q, str(tmp_path / "o.fa"))
    two = str(tmp_path / "t.fa")
    open(two, "wb").write(b">a\n" + b"ACGT" * 30 + b"\n>b\n" + b"ACGA" * 30 + b"\n")
    bad = str(tmp_path / "bad.fa")
    open(bad, "wb").write(b">a\n" + b"ACGT" * 30 + b"\n>c\n" + b"ACGA" * 30 + b"\n")
    with pytest.raises(halign.HalignError, match="not found"):
        halign.msa_keep_length(ctx, two, q, str(tmp_path / "o.fa"), aligned_path=bad)
    st = halign.msa_keep_length(ctx, two, q, str(tmp_path / "o.fa"))
    assert st["records"] == 3
