"""The auto-center path on the GPU: halign_consensus_file's FASTA and JSON
against generateConsensusSequence's files (incl. the dropped record after
each batch and seq_limit), and halign4 --keep-length without -c
(halign_msa_auto_center) against the reference's final FASTA
(tests/golden/auto_center.npz)."""
import pytest

from test_auto_center_host import auto_center_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import halign
    c = halign.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("case", range(6))
def test_consensus_file_vs_reference(ctx, tmp_path, case):
    import halign
    text, lim, batch, fa, js = auto_center_fixture()[1][case]
    print(f"consensus file case {case}: seq_limit {lim}, batch {batch}", flush=True)
    ip = str(tmp_path / "a.fa")
    open(ip, "wb").write(text)
    cons = halign.consensus_file(ctx, ip, str(tmp_path / "c.fa"), str(tmp_path / "c.json"), lim, batch)
    assert open(tmp_path / "c.fa", "rb").read() == fa
    assert open(tmp_path / "c.json", "rb").read() == js
    assert b">consensus\n" + b"".join(cons[i:i + 80] + b"\n" for i in range(0, len(cons), 80)) == fa


def test_consensus_file_errors(ctx, tmp_path):
    import halign
    bad = str(tmp_path / "b.fa")
    open(bad, "wb").write(b">a\nACGT\n>b\nACG\n")
    with pytest.raises(halign.HalignError, match="mismatch"):
        halign.consensus_file(ctx, bad, str(tmp_path / "c.fa"), str(tmp_path / "c.json"))
    empty = str(tmp_path / "e.fa")
    open(empty, "wb").write(b"")
    with pytest.raises(halign.HalignError):
        halign.consensus_file(ctx, empty, str(tmp_path / "c.fa"), str(tmp_path / "c.json"))


@pytest.mark.parametrize("case", range(2))
def test_auto_center_run_vs_reference(ctx, tmp_path, case):
    import halign
    text, aligned, k, expect = auto_center_fixture()[2][case]
    print(f"auto-center case {case}: cons_n {k}, aligned file {aligned is not None}", flush=True)
    ip = str(tmp_path / "in.fa")
    open(ip, "wb").write(text)
    ap = None
    if aligned is not None:
        ap = str(tmp_path / "aligned.fa")
        open(ap, "wb").write(aligned)
    op = str(tmp_path / "out.fa")
    halign.msa_auto_center(ctx, ip, op, cons_n=k, aligned_topk_path=ap)
    assert open(op, "rb").read() == expect
    assert sorted(p.name for p in tmp_path.iterdir()) == sorted(["in.fa", "out.fa"] + (["aligned.fa"] if ap else []))
