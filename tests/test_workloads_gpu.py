"""GPU parity on samples of the uneven (3), divergent (4) and serving (5)
workloads against the reference program's own output (tests/golden/
config<c>.npz, written by oracle/_ref/halign4 -c center --keep-length -t 1
with the copy command): every query's CIGAR (RefAligner::globalAlign,
src/align/ref_aligner.cpp:91-125), its SAM group (cigar::hasInsertion,
:253-260), the final aligned FASTA of the file run (:676-740), and the
serving engine's CIGARs for requests cut from the same sample. The samples
hold the first 64 queries of each workload, random ones, and the cases each
workload stresses: anchor-less short queries and the longest tail (3), long
indels (4), queries of fewer than 20 bases, longer than the center or near
the auto-band switch (5). A last test plants diverged stretches whose
segments overflow their band (forced I(q)D(r), src/align/align.cpp:243-261)
and checks them against the oracle."""
import hashlib
import os
import sys
import tempfile

import numpy as np
import pytest

from oracle_lib import Oracle, cigar_str
from workloads import expected_cigar, load_config, write_fasta_inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import halign
    c = halign.Context(0)
    yield c
    c.close()


_CACHE = {}


def fixture(cfg):
    if cfg not in _CACHE:
        _CACHE[cfg] = load_config(cfg)
    return _CACHE[cfg]


def _mismatches(f, cig, coff):
    bad = []
    for k in range(len(f["idx"])):
        if not np.array_equal(cig[coff[k]:coff[k + 1]], expected_cigar(f, k)):
            bad.append(k)
    return bad


@pytest.mark.parametrize("cfg", [3, 4, 5])
def test_workload_cigars_one_batch(ctx, cfg):
    """The whole sample as one batch: every CIGAR equals the reference's."""
    import halign
    f = fixture(cfg)
    n = len(f["idx"])
    print(f"workload {cfg}: {n} queries, one batch", file=sys.stderr, flush=True)
    from halign import synth
    c = halign.Center(ctx, synth.center(cfg))
    try:
        cig, coff, st, _ = c.align_flat(f["queries"], f["qoff"])
    finally:
        c.close()
    bad = _mismatches(f, cig, coff)
    assert not bad, (f"workload {cfg}: {len(bad)} of {n} CIGARs differ; first: query {int(f['idx'][bad[0]])} "
                     f"got {cigar_str(cig[coff[bad[0]]:coff[bad[0] + 1]])[:200]} "
                     f"expected {cigar_str(expected_cigar(f, bad[0]))[:200]}")


@pytest.mark.parametrize("cfg", [3, 4, 5])
def test_workload_rows_and_routing_uneven_batches(ctx, cfg):
    """The sample in batches of uneven sizes: the CIGARs do not depend on
    the batch, the keep-length rows have the center's length, and the SAM
    group (hasInsertion) is the one the reference wrote each query to."""
    import halign
    from halign import synth
    f = fixture(cfg)
    n = len(f["idx"])
    rng = np.random.default_rng(cfg)
    c = halign.Center(ctx, synth.center(cfg))
    q, off = f["queries"], f["qoff"]
    try:
        a, bad_cig, bad_ins, bad_row = 0, [], [], []
        while a < n:
            b = min(n, a + int(rng.integers(1, 700)))
            print(f"workload {cfg}: rows batch {a}..{b}", file=sys.stderr, flush=True)
            qs = [q[off[k]:off[k + 1]] for k in range(a, b)]
            cigs, rows, hins = c.align_rows(qs)
            for j, k in enumerate(range(a, b)):
                if not np.array_equal(cigs[j], expected_cigar(f, k)):
                    bad_cig.append(k)
                if bool(hins[j]) != bool(f["has_insertion"][k]):
                    bad_ins.append(k)
                if len(rows[j]) != c.rlen:
                    bad_row.append(k)
            a = b
    finally:
        c.close()
    assert not bad_cig, f"workload {cfg}: CIGARs of queries {[int(f['idx'][k]) for k in bad_cig[:10]]} differ"
    assert not bad_ins, f"workload {cfg}: SAM group of queries {[int(f['idx'][k]) for k in bad_ins[:10]]} differs"
    assert not bad_row, f"workload {cfg}: rows of queries {[int(f['idx'][k]) for k in bad_row[:10]]} not center length"


@pytest.mark.parametrize("cfg", [3, 4, 5])
def test_workload_final_fasta(ctx, cfg):
    """The file run (halign4 -c center --keep-length) over the sample writes
    the reference program's final FASTA, byte for byte (SHA-256)."""
    import halign
    f = fixture(cfg)
    d = tempfile.mkdtemp()
    cpath, ipath = write_fasta_inputs(f, cfg, d)
    out = os.path.join(d, "out.fa")
    print(f"workload {cfg}: keep-length file run", file=sys.stderr, flush=True)
    halign.msa_keep_length(ctx, cpath, ipath, out, batch=500)
    data = open(out, "rb").read()
    assert len(data) == int(f["fasta_bytes"][0])
    assert hashlib.sha256(data).digest() == f["fasta_sha256"].tobytes()


def test_serving_workload5_requests():
    """Serving workload: the sample as requests of 1..300 queries submitted
    at once to a two-worker server; every CIGAR equals the reference's."""
    import halign
    from halign import synth
    f = fixture(5)
    n = len(f["idx"])
    q, off = f["queries"], f["qoff"]
    rng = np.random.default_rng(55)
    srv = halign.Server(synth.center(5), workers=2, max_batch_queries=512)
    try:
        reqs, a = [], 0
        while a < n:
            b = min(n, a + int(rng.integers(1, 301)))
            o = off[a:b + 1] - off[a]
            reqs.append((a, b, srv.submit_flat(q[off[a]:off[b]], o)))
            a = b
        print(f"serving workload 5: {len(reqs)} requests", file=sys.stderr, flush=True)
        bad = []
        for a, b, rid in reqs:
            cig, coff, _, _ = srv.result(rid)
            for j, k in enumerate(range(a, b)):
                if not np.array_equal(cig[coff[j]:coff[j + 1]], expected_cigar(f, k)):
                    bad.append(k)
    finally:
        srv.close()
    assert not bad, f"served CIGARs of queries {[int(f['idx'][k]) for k in bad[:10]]} differ"


def test_serving_small_request_not_queued_behind_large():
    """Chunked round-robin admission on the serving workload: one worker,
    batches of 256 queries, chunks of 64. A request of 1,400 queries goes in
    first, one of 20 right after; the small one rides the next batch the
    worker forms, so it is done while the large one still has batches to
    go. Both match the reference's CIGARs, and the server's footprint
    estimate and measured device use are reported."""
    import halign
    from halign import synth
    f = fixture(5)
    q, off = f["queries"], f["qoff"]
    srv = halign.Server(synth.center(5), workers=1, max_batch_queries=256, chunk_queries=64)
    try:
        big = srv.submit_flat(q[off[0]:off[1400]], off[:1401] - off[0])
        small = srv.submit_flat(q[off[1400]:off[1420]], off[1400:1421] - off[1400])
        cs, ofs, _, ts = srv.result(small)
        big_pending = not srv.wait(big, 0)
        cb, ofb, _, tb = srv.result(big)
        m = srv.stats()
    finally:
        srv.close()
    print(f"serving chunks: small done at {ts[1]:.1f} ms, large at {tb[1]:.1f} ms, {m['batches']} batches, "
          f"largest batch estimate {m['max_batch_estimate'] / 2**20:.1f} MiB, working buffers measured "
          f"{m['working_measured'] / 2**20:.1f} MiB, traceback reserved {m['traceback_reserved'] / 2**30:.1f} GiB",
          file=sys.stderr, flush=True)
    assert big_pending, "the small request waited for the whole large one"
    assert m["batches"] >= 1400 // 256
    assert m["traceback_reserved"] > 0 and m["device_used"] >= m["traceback_reserved"]
    assert ts[2] <= 256 and tb[2] <= 256
    bad = [k for k in range(1400) if not np.array_equal(cb[ofb[k]:ofb[k + 1]], expected_cigar(f, k))]
    bad += [1400 + j for j in range(20) if not np.array_equal(cs[ofs[j]:ofs[j + 1]], expected_cigar(f, 1400 + j))]
    assert not bad, f"served CIGARs of queries {[int(f['idx'][k]) for k in bad[:10]]} differ"


def test_band_overflow_segments_vs_oracle(ctx):
    """Queries of the divergent workload's center with 1-3 stretches of
    600-2,000 bases replaced by random ones of 50-80% of that length: the
    segment holding a stretch leaves its auto band (|dq - dr| > 200 + 0.1 *
    (dr + dq / 2), ratio <= 0.5), so the reference writes I(q) D(r) for it
    (src/align/align.cpp:243-261). The batch must hold many such segments
    (the oracle's forced I+D count), and every CIGAR equals the oracle's."""
    import halign
    from halign import synth
    center = np.frombuffer(synth.center(4), np.uint8)
    rng = np.random.default_rng(404)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    qs = []
    for i in range(120):
        s = center.copy()
        for _ in range(int(rng.integers(1, 4))):
            L = int(rng.integers(600, 2001))
            a = int(rng.integers(0, len(s) - L))
            m = int(L * rng.uniform(0.5, 0.8))
            s = np.concatenate([s[:a], acgt[rng.integers(0, 4, m)], s[a + L:]])
        sub = rng.random(len(s)) < 0.01
        s[sub] = acgt[rng.integers(0, 4, int(sub.sum()))]
        qs.append(s.tobytes())
    O = Oracle()
    exp, forced = [], 0
    for s in qs:
        cig, _, st, _ = O.ref_global_align(center.tobytes(), s)
        exp.append(cig)
        forced += int(st[1])
    print(f"band overflow: {len(qs)} queries, {forced} forced I+D segments", file=sys.stderr, flush=True)
    assert forced >= len(qs) // 2
    c = halign.Center(ctx, center.tobytes())
    try:
        got, _, _ = c.align(qs)
    finally:
        c.close()
    bad = [i for i in range(len(qs)) if not np.array_equal(got[i], exp[i])]
    assert not bad, f"{len(bad)} CIGARs differ, first {bad[0]}"
