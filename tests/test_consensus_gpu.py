"""GPU parity of the column-count / consensus kernel against the oracle
(consensus::generateConsensusSequence's counting, src/consensus/consensus.cpp:
611-703) and the reference-written fixtures, bit-exact."""
import os
import sys

import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

ALPHA = np.frombuffer(b"ACGTUN-.acgtuRYX*", np.uint8)


@pytest.fixture(scope="module")
def ctx():
    import halign
    c = halign.Context(0)
    yield c
    c.close()


def test_consensus_golden(ctx):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "consensus.npz")
    with np.load(path, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    print("consensus golden cases:", len(g["case"]), file=sys.stderr, flush=True)
    for ci, (row_set, lim) in enumerate(g["case"]):
        nrows, aln = (int(v) for v in g["shape"][row_set])
        rows = g["rows"][g["rows_off"][row_set]:g["rows_off"][row_set + 1]].reshape(nrows, aln)
        counts, cons, n, _ = ctx.consensus(rows, seq_limit=int(lim))
        assert np.array_equal(counts.reshape(-1), g["counts"][g["counts_off"][ci]:g["counts_off"][ci + 1]]), ci
        assert cons == g["cons"][g["cons_off"][ci]:g["cons_off"][ci + 1]].tobytes(), ci
        assert n == int(g["num_seqs"][ci]), ci


@pytest.mark.parametrize("nrows,aln,batch,lim", [(1, 1, 4096, 0), (3, 5, 4096, 0), (5000, 37, 4096, 0),
                                                  (9000, 13, 4096, 4500), (700, 101, 7, 0),
                                                  (700, 101, 7, 333), (64, 4099, 4096, 2)])
def test_consensus_vs_oracle(ctx, nrows, aln, batch, lim):
    rng = np.random.default_rng(nrows * 31 + aln)
    rows = ALPHA[rng.integers(0, len(ALPHA), (nrows, aln))]
    rows[rng.random(rows.shape) < 0.5] = ord("G")
    print(f"consensus vs oracle: {nrows} x {aln}, batch {batch}, limit {lim}", file=sys.stderr, flush=True)
    c1, s1, n1, _ = ctx.consensus(rows, batch=batch, seq_limit=lim)
    c2, s2, n2 = Oracle().consensus(rows, batch=batch, seq_limit=lim)
    assert np.array_equal(c1, c2) and s1 == s2 and n1 == n2


def test_consensus_large_histogram(ctx):
    """120 MB of rows: counts equal a numpy histogram of the counted rows and
    sum to rows x columns."""
    nrows, aln = 4100, 29903
    rng = np.random.default_rng(17)
    rows = ALPHA[rng.integers(0, len(ALPHA), (nrows, aln))]
    print(f"consensus large: {nrows} x {aln}", file=sys.stderr, flush=True)
    counts, cons, n, ms = ctx.consensus(rows)
    kept = np.ones(nrows, bool)
    kept[4096::4097] = False          # the record read after a full batch is dropped
    lut = np.full(256, 5, np.int64)
    for i, ch in enumerate(b"ACGTU"):
        lut[ch] = i
        lut[ch + 32] = i
    lut[ord("-")] = lut[ord(".")] = 6
    cls = lut[rows[kept]]
    exp = np.zeros((aln, 7), np.int64)
    for k in range(7):
        exp[:, k] = (cls == k).sum(axis=0)
    assert n == kept.sum() and np.array_equal(counts, exp)
    assert int(counts.sum()) == int(kept.sum()) * aln
