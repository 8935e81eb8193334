"""Samples of the seeded benchmark workloads (csrc/synth.c through
halign.synth) and the fixtures the reference program wrote for them
(tests/golden/config<c>.npz: query indices, CIGARs, SAM group, final FASTA
digest)."""
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def workload_queries(cfg, idx):
    """Queries idx (in that order) of workload cfg: (bytes, int64 offsets)."""
    from halign import synth
    parts = []
    for i in idx:
        q, _ = synth.queries(cfg, int(i), 1)
        parts.append(q)
    off = np.zeros(len(parts) + 1, np.int64)
    off[1:] = np.cumsum([len(p) for p in parts])
    return b"".join(parts), off


def load_config(cfg):
    """The fixture of workload cfg and its queries regenerated from the seeds."""
    with np.load(os.path.join(GOLD, "config%d.npz" % cfg), allow_pickle=False) as z:
        f = {k: z[k] for k in z.files}
    q, off = workload_queries(cfg, f["idx"])
    f["queries"], f["qoff"] = q, off
    return f


def expected_cigar(f, k):
    return f["cigar"][f["cigar_off"][k]:f["cigar_off"][k + 1]]


def write_fasta_inputs(f, cfg, d):
    """The center and query files the fixture's reference run read."""
    from halign import synth
    cpath, ipath = os.path.join(d, "center.fa"), os.path.join(d, "q.fa")
    with open(cpath, "wb") as fh:
        fh.write(b">center\n" + synth.center(cfg) + b"\n")
    q, off = f["queries"], f["qoff"]
    with open(ipath, "wb") as fh:
        for k, i in enumerate(f["idx"]):
            fh.write(b">q%d\n" % int(i) + q[off[k]:off[k + 1]] + b"\n")
    return cpath, ipath
