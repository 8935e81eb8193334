"""The auto-center path's host part on the CPU: halign_topk_longest (host
code of libhalign_hip.so) against preprocessInputFasta's Top-K files written
by the reference (tests/golden/auto_center.npz)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def auto_center_fixture():
    z = np.load(os.path.join(GOLDEN, "auto_center.npz"))

    def seg(k, i):
        return z[k][z[k + "_off"][i]:z[k + "_off"][i + 1]].tobytes()

    topk = [(seg("topk_in", i), int(z["topk_k"][i]), seg("topk_out", i)) for i in range(len(z["topk_k"]))]
    cons = [(seg("cons_in", i), int(z["cons_par"][i][0]), int(z["cons_par"][i][1]), seg("cons_fa", i),
             seg("cons_json", i)) for i in range(len(z["cons_par"]))]
    runs = [(seg("run_in", i), seg("run_aligned", i) if z["run_par"][i][1] else None, int(z["run_par"][i][0]),
             seg("run_out", i)) for i in range(len(z["run_par"]))]
    return topk, cons, runs


@pytest.mark.parametrize("case", range(5))
def test_topk_vs_reference(tmp_path, case):
    import halign
    text, k, expect = auto_center_fixture()[0][case]
    ip = str(tmp_path / "in.fa")
    open(ip, "wb").write(text)
    op = str(tmp_path / "topk.fa")
    n = halign.topk_longest(ip, k, op)
    assert open(op, "rb").read() == expect
    assert n == text.count(b">")


def test_topk_fixture_covers_ties_and_cleaning():
    """The reference-written Top-K files: longest first, input order among equal
    lengths, cleaned bases; and some case cuts through a group of equal
    lengths, so the tie order decides which records are in."""
    from oracle_lib import Oracle
    O = Oracle()
    cut_tie = False
    for text, k, out in auto_center_fixture()[0]:
        recs = O.fastx_parse(text)
        got = O.fastx_parse(out)
        assert len(got) == min(k, len(recs))
        order = {r[0]: i for i, r in enumerate(recs)}
        keys = [(-len(sq), order[n]) for n, _, sq in got]
        assert keys == sorted(keys)
        assert all(set(sq) <= set(b"ACGTUN-") for _, _, sq in got)
        if got and len(got) < len(recs):
            last = len(got[-1][2])
            cut_tie |= sum(len(sq) == last for _, _, sq in recs) > sum(len(sq) == last for _, _, sq in got)
    assert cut_tie
