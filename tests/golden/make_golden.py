#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ from the reference itself.

Every expected output here comes from oracle/_ref/libhalign_ref.so, the
reference's own sources (src/align/ksw2_extz2_sse.c, src/align/align.cpp,
src/seed/minimizer.cpp, src/seed/anchor.cpp, src/align/ref_aligner.cpp,
src/consensus/consensus.cpp, ...) compiled with the production flags by
`make -C oracle ref` (oracle/Makefile) behind the thin C harness
oracle/ref_harness.cpp. Inputs are seeded (numpy PCG64) and stored in the
fixtures, so the fixtures do not depend on this script's random stream.

Run from the repository root:  python tests/golden/make_golden.py
Outputs (np.savez_compressed, no pickles):
  ksw_cases.npz        ksw_extz2_sse on 0..4 codes: every ksw_extz_t field + CIGAR
  ksw_whole_pairs.npz  globalAlignKSW2 on whole 29.9 kb synthetic pairs: CIGAR
  global_align.npz     globalAlignKSW2 on ASCII pairs (IUPAC, lowercase, U, empty)
  minimizers.npz       extractMinimizer over (k, w, non_canonical)
  anchors_chain.npz    collect_anchors -> chainAnchors (in-place sort order + best chain)
  mm2_cases.npz        globalAlignMM2 on (ref, query, reference anchors)
  ref_global.npz       RefAligner::globalAlign (seed -> anchors -> chain -> MM2)
  consensus.npz        generateConsensusSequence: counts, consensus, row count
                       (case = (row set, seq_limit))
  fastx_cases.npz      seq_io::KseqReader records of FASTA/FASTQ texts (plain and
                       gzip), or its error
  msa_keep_length.npz  halign4 -c center --keep-length (one thread, copy MSA
                       command): the final FASTA bytes of center + query files
  mash_cases.npz       mash::sketchFromSequence over (k, sketch size, canonical,
                       seed), and the Jaccard choice of a reference
  msa_refs.npz         halign4 -c refs --keep-length with several references
                       (copy command, or a given aligned file): final FASTA bytes
  auto_center.npz      preprocessInputFasta's Top-K file, generateConsensusSequence's
                       FASTA and JSON files, and halign4 --keep-length without -c
  sam_cases.npz        halign4 -t 1 --save-workdir: the thread-0 SAM files (plain and
                       insertion) of one and several references, with and without
                       --keep-length (the second alignment to the consensus)
  config3.npz, config4.npz, config5.npz
                       samples of the uneven (3), divergent (4) and serving (5)
                       workloads through the reference program itself
                       (oracle/_ref/halign4 -c center --keep-length, one thread,
                       copy command): each query's SAM CIGAR and group, and the
                       SHA-256 of the final FASTA; the queries are stored as
                       indices into the seeded workload (csrc/synth.c, pinned
                       by tests/test_synth.py)

Run a subset:  python tests/golden/make_golden.py config3 config4
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "halign-4_amd"))

from oracle_lib import ANCHOR_DTYPE, FLAG_LEFT, FLAG_RIGHT, Halign4, Oracle, RefLib  # noqa: E402
from seqgen import codes_to_ascii, mutate_codes, rand_codes  # noqa: E402
from workloads import workload_queries  # noqa: E402

ACGT = np.frombuffer(b"ACGT", np.uint8)


def concat(arrs, dtype):
    off = np.zeros(len(arrs) + 1, np.int64)
    for i, a in enumerate(arrs):
        off[i + 1] = off[i] + len(a)
    data = np.concatenate([np.asarray(a, dtype) for a in arrs]) if arrs else np.zeros(0, dtype)
    return data.astype(dtype), off


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"  {name}: {os.path.getsize(path) / 1024:.0f} KiB", flush=True)


def ksw_cases(R, O, rng):
    """G1: ~3,000 DP cases over the shapes the kernel meets."""
    cases = []
    for it in range(3000):
        kind = it % 8
        if kind == 0:      # tiny random (0..64)
            t = rand_codes(rng, int(rng.integers(1, 65)))
            q = rand_codes(rng, int(rng.integers(1, 65)))
        elif kind == 1:    # mutated with N
            t = rand_codes(rng, int(rng.integers(1, 300)), p_n=0.05)
            q = mutate_codes(rng, t, 0.1, 0.05)
        elif kind == 2:    # 64..512 full band
            t = rand_codes(rng, int(rng.integers(64, 513)))
            q = mutate_codes(rng, t, 0.05, 0.02)
        elif kind == 3:    # binary alphabet: tie-heavy
            t = rand_codes(rng, int(rng.integers(1, 200)), alphabet=2)
            q = rand_codes(rng, int(rng.integers(1, 200)), alphabet=2)
        elif kind == 4:    # homopolymer runs
            n = int(rng.integers(1, 300))
            t = np.repeat(rand_codes(rng, n // 6 + 1), 6)[:n]
            q = mutate_codes(rng, t, 0.05, 0.05)
        elif kind == 5:    # banded 500..3000
            t = rand_codes(rng, int(rng.integers(500, 3001)))
            q = mutate_codes(rng, t, 0.02, 0.01)
        elif kind == 6:    # |tlen - qlen| near the band: overflow (zdropped) and just inside
            t = rand_codes(rng, int(rng.integers(300, 2000)))
            w = int(rng.integers(8, 120))
            cut = max(1, len(t) - 2 * w + int(rng.integers(-20, 21)))
            q = mutate_codes(rng, t[:cut], 0.02, 0.01)
        else:              # unrelated pairs
            t = rand_codes(rng, int(rng.integers(1, 120)))
            q = rand_codes(rng, int(rng.integers(1, 120)))
        if len(q) == 0:
            q = rand_codes(rng, 1)
        flag = FLAG_RIGHT if rng.random() < 0.6 else FLAG_LEFT
        r = rng.random()
        if kind == 6:
            pass
        elif r < 0.3:
            w = -1
        elif r < 0.6:
            w = int(rng.integers(0, 48))
        else:
            w = O.auto_band(len(t), len(q))
        cases.append((q, t, w, flag))
    ez, cigs = [], []
    for q, t, w, flag in cases:
        e, c = R.ksw(q, t, w, flag)
        ez.append(e)
        cigs.append(c)
    qd, qo = concat([c[0] for c in cases], np.uint8)
    td, to = concat([c[1] for c in cases], np.uint8)
    cd, co = concat(cigs, np.uint32)
    save("ksw_cases.npz", q=qd, q_off=qo, t=td, t_off=to,
         w=np.array([c[2] for c in cases], np.int32), flag=np.array([c[3] for c in cases], np.int32),
         ez=np.array(ez, np.int64), ez_fields=np.array(
             ["max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score",
              "n_cigar"]), cigar=cd, cigar_off=co)


def whole_pairs(R):
    """G1b: whole config-2 pairs (center 29,903 bp, queries 29.0-30.0 kb from
    halign-4_amd/csrc/synth.c, regenerated by the tests) through globalAlignKSW2."""
    from halign import synth
    center = synth.center(2)
    n = 20
    qb, off = synth.queries(2, 0, n)
    cigs = []
    for i in range(n):
        cigs.append(R.global_align_ksw2(center, qb[off[i]:off[i + 1]], FLAG_RIGHT))
        print(f"    whole pair {i}: {len(cigs[-1])} ops", flush=True)
    cd, co = concat(cigs, np.uint32)
    save("ksw_whole_pairs.npz", config=np.int32(2), first=np.int64(0), count=np.int64(n),
         center_len=np.int64(len(center)), query_off=off, cigar=cd, cigar_off=co)


def rand_ascii(rng, n, p_n=0.0, extra=b""):
    s = ACGT[rng.integers(0, 4, n)].copy()
    if p_n and n:
        pool = np.frombuffer(b"NRYKMSWBDHVN" + extra, np.uint8)
        m = rng.random(n) < p_n
        s[m] = pool[rng.integers(0, len(pool), int(m.sum()))]
    return s.tobytes()


def global_align(R, rng):
    """G1c: globalAlignKSW2 (src/align/align.cpp:22-65) on ASCII pairs."""
    refs, qrys, flags, cigs = [], [], [], []
    for i in range(400):
        rl = int(rng.integers(0, 500)) if i % 10 else 0
        ref = rand_ascii(rng, rl, 0.02, b"U")
        codes = np.frombuffer(ref, np.uint8)
        lut = np.full(256, 4, np.uint8)
        for c, v in zip(b"ACGT", range(4)):
            lut[c] = v
        q = mutate_codes(rng, lut[codes], 0.05, 0.02) if i % 4 else rand_codes(rng, int(rng.integers(0, 60)))
        qry = codes_to_ascii(q)
        if i % 7 == 0:
            qry = qry.lower()
        if i % 11 == 0:
            qry = qry.replace(b"T", b"U")
        if i % 13 == 0:
            qry = b""
        flag = FLAG_RIGHT if i % 3 else FLAG_LEFT
        refs.append(np.frombuffer(ref, np.uint8))
        qrys.append(np.frombuffer(qry, np.uint8))
        flags.append(flag)
        cigs.append(R.global_align_ksw2(ref, qry, flag))
    rd, ro = concat(refs, np.uint8)
    qd, qo = concat(qrys, np.uint8)
    cd, co = concat(cigs, np.uint32)
    save("global_align.npz", ref=rd, ref_off=ro, qry=qd, qry_off=qo,
         flag=np.array(flags, np.int32), cigar=cd, cigar_off=co)


def minimizers(R, O, rng):
    """G2: extractMinimizer (src/seed/minimizer.cpp:103-177)."""
    seqs, ks, ws, ncs, xs, ys = [], [], [], [], [], []
    skipped = 0
    for it in range(1500):
        n = int(rng.integers(0, 600))
        k = int(rng.choice([3, 5, 15, 21, 31, 32]))
        w = int(rng.choice([1, 2, 3, 5, 10, 10, 10, 255, 256]))
        s = rand_ascii(rng, n, 0.01 if it % 3 == 0 else 0.0)
        if it % 5 == 0:
            s = bytes(ACGT[rng.integers(0, 2, n)])
        if it % 7 == 0:
            s = s.lower()
        if it % 17 == 0:
            s = s.replace(b"T", b"U")
        nc = bool(it % 2)
        # the reference indexes past its ring buffer for some small windows
        # (undefined behaviour there); such inputs are not fixtures
        _, _, ub = O.minimizers(s, k, w, nc)
        if ub & 3:
            skipped += 1
            continue
        x, y = R.minimizers(s, k, w, nc)
        seqs.append(np.frombuffer(s, np.uint8))
        ks.append(k)
        ws.append(w)
        ncs.append(int(nc))
        xs.append(x)
        ys.append(y)
    sd, so = concat(seqs, np.uint8)
    xd, xo = concat(xs, np.uint64)
    yd, _ = concat(ys, np.uint64)
    print(f"    minimizer inputs skipped (reference UB): {skipped}")
    save("minimizers.npz", seq=sd, seq_off=so, k=np.array(ks, np.int32), w=np.array(ws, np.int32),
         non_canonical=np.array(ncs, np.int32), x=xd, y=yd, hit_off=xo)


def make_pair(rng, it, max_len=6000):
    rl = int(rng.integers(20, max_len))
    ref = rand_ascii(rng, rl, 0.001 if it % 4 == 0 else 0)
    if it % 10 == 0:
        ref = (ref[:max(1, rl // 3)] * 4)[:rl]   # tandem repeats: frequent minimizers
    st = int(rng.integers(0, rl // 2 + 1))
    en = int(rng.integers(st + 1, rl + 1))
    lut = np.full(256, 4, np.uint8)
    for c, v in zip(b"ACGT", range(4)):
        lut[c] = v
    src = ref[st:en] if it % 3 else ref
    q = mutate_codes(rng, lut[np.frombuffer(src, np.uint8)], [0.01, 0.05, 0.1][it % 3],
                     [0.002, 0.01, 0.05][it % 3])
    return ref, codes_to_ascii(q)


def anchors_chain_mm2(R, rng):
    """G3 + G4: collect_anchors -> chainAnchors -> globalAlignMM2 on pairs."""
    refs, qrys = [], []
    rxs, rys, qxs, qys, anc, srt, chn, mm2 = [], [], [], [], [], [], [], []
    for it in range(160):
        ref, qry = make_pair(rng, it, 4000)
        rx, ry = R.minimizers(ref)
        qx, qy = R.minimizers(qry)
        a = R.collect_anchors(rx, ry, qx, qy)
        s, c = R.chain(a)
        # chainAnchors sorts its input in place: keep that order as a permutation of a
        rec = ANCHOR_DTYPE.itemsize
        key = {a.tobytes()[i * rec:(i + 1) * rec]: i for i in range(len(a))}
        s = np.array([key[s.tobytes()[i * rec:(i + 1) * rec]] for i in range(len(s))], np.uint32)
        m = R.mm2(ref, qry, a)
        refs.append(np.frombuffer(ref, np.uint8))
        qrys.append(np.frombuffer(qry, np.uint8))
        rxs.append(rx)
        rys.append(ry)
        qxs.append(qx)
        qys.append(qy)
        anc.append(a)
        srt.append(s)
        chn.append(c)
        mm2.append(m)
    out = {}
    for name, arrs, dt in (("ref", refs, np.uint8), ("qry", qrys, np.uint8), ("rx", rxs, np.uint64),
                           ("ry", rys, np.uint64), ("qx", qxs, np.uint64), ("qy", qys, np.uint64),
                           ("mm2", mm2, np.uint32), ("sorted", srt, np.uint32)):
        d, o = concat(arrs, dt)
        out[name] = d
        out[name + "_off"] = o
    for name, arrs in (("anchors", anc), ("chain", chn)):
        o = np.zeros(len(arrs) + 1, np.int64)
        for i, a in enumerate(arrs):
            o[i + 1] = o[i] + len(a)
        out[name] = np.concatenate(arrs).view(np.uint8) if o[-1] else np.zeros(0, np.uint8)
        out[name + "_off"] = o
    save("anchors_chain.npz", **out)


def ref_global(R, rng):
    """G4b: RefAligner::globalAlign (src/align/ref_aligner.cpp:91-125), the
    whole per-query path, on short pairs and on config-1-sized queries."""
    from halign import synth
    refs, qrys, cigs = [], [], []
    for it in range(150):
        ref, qry = make_pair(rng, it + 1000)
        refs.append(np.frombuffer(ref, np.uint8))
        qrys.append(np.frombuffer(qry, np.uint8))
        cigs.append(R.ref_global_align(ref, qry))
    center = synth.center(1)
    qb, off = synth.queries(1, 0, 30)
    for i in range(30):
        q = qb[off[i]:off[i + 1]]
        refs.append(np.frombuffer(center, np.uint8))
        qrys.append(np.frombuffer(q, np.uint8))
        cigs.append(R.ref_global_align(center, q))
    rd, ro = concat(refs, np.uint8)
    qd, qo = concat(qrys, np.uint8)
    cd, co = concat(cigs, np.uint32)
    save("ref_global.npz", ref=rd, ref_off=ro, qry=qd, qry_off=qo, cigar=cd, cigar_off=co)


def consensus(R, rng):
    """G6: generateConsensusSequence (src/consensus/consensus.cpp:611-703),
    including more than one 4,096-row batch."""
    d = tempfile.mkdtemp()
    alpha = np.frombuffer(b"ACGTUN-.acgtuRYX", np.uint8)
    rows_all, shapes, limits, counts, conss, nseq = [], [], [], [], [], []
    for it, nrows in enumerate([1, 2, 5, 4095, 4096, 4097, 8193, 9000]):
        aln = int(rng.integers(1, 24))
        rows = alpha[rng.integers(0, len(alpha), (nrows, aln))]
        if nrows > 100:   # skewed columns, so the argmax is not a coin toss
            rows[rng.random(rows.shape) < 0.6] = ord("A") + it % 3
        fa = os.path.join(d, "a.fa")
        with open(fa, "wb") as f:
            for i in range(nrows):
                f.write(b">r%d\n" % i + rows[i].tobytes() + b"\n")
        rows_all.append(rows.reshape(-1))
        shapes.append((nrows, aln))
        for lim in (0, 3, 4097, 5000):
            cons = R.consensus_file(fa, os.path.join(d, "c.fa"), os.path.join(d, "c.json"),
                                    seq_limit=lim)
            js = json.load(open(os.path.join(d, "c.json")))["consensus"]
            cnt = np.array([[c["A"], c["C"], c["G"], c["T"], c["U"], c["N"], c["-"]]
                            for c in js["counts"]], np.uint32)
            limits.append((it, lim))
            counts.append(cnt.reshape(-1))
            conss.append(np.frombuffer(cons, np.uint8))
            nseq.append(js["num_seqs"])
    rd, ro = concat(rows_all, np.uint8)
    cd, co = concat(counts, np.uint32)
    sd, so = concat(conss, np.uint8)
    save("consensus.npz", rows=rd, rows_off=ro, shape=np.array(shapes, np.int64),
         case=np.array(limits, np.int64), counts=cd, counts_off=co, cons=sd, cons_off=so,
         num_seqs=np.array(nseq, np.int64))


FASTX_FIXED = [
    b">a desc here\nACGT\nacgt\n>b\nNNNN\n",
    b">a\r\nAC\r\nGT\r\n>b x\r\nTT\r\n",
    b"junk\n>a\n\n\nAC\n\n>b\n",
    b">a\n\r\n>b\n\r\nA\n",
    b"@q1 c\nACGT\n+\nIIII\n@q2\nAC\nGT\n+q2\nII\nII\n",
    b"@q1\nACGT\n+\nIII\n",
    b"@q1\nACGT\n+",
    b">a\tb c\nA C G\n@x\n",
    b">\nACGT\n> d\nAA\n>",
    b">a \r\nAAA\r",
    b">a\nAAA\n\r",
    b"\r",
    b"",
    b">only",
    b">a\n+\n\n>b\nAC\n",
    b">a b\r\nA\rC\n",
    b"@r\n\n+\n\n@s\nA\n+\n!\n",
    b"@r\n+\n@s\nA\n+\n!\n",
]


def fastx_text(rng, size):
    """Random FASTA/FASTQ-like text over header, separator and base bytes."""
    alpha = np.frombuffer(b">@+\n\n\r \tACGTacgtNn-", np.uint8)
    p = np.array([2, 1, 1, 8, 4, 1, 1, 1] + [6] * 8 + [1, 1, 1], np.float64)
    return alpha[rng.choice(len(alpha), size, p=p / p.sum())].tobytes()


def fastx_cases(R, rng):
    """seq_io::KseqReader (src/utils/seq_io.cpp:67-155 over include/kseq.h) on
    FASTA/FASTQ texts: fixed edge cases, random texts, and two texts larger
    than the readers' buffers; each read plain and gzip-compressed."""
    import gzip
    texts = list(FASTX_FIXED)
    texts += [fastx_text(rng, int(rng.integers(1, 300))) for _ in range(400)]
    recs = []
    unit = ACGT[rng.integers(0, 4, 997)].tobytes() * 41   # repetitive, so the fixture compresses
    for i in range(2):   # records of 0.5-40 kb, 3 MB in all
        parts = []
        tot = 0
        k = 0
        while tot < 3_000_000:
            n = int(rng.integers(500, 40_000))
            a = int(rng.integers(0, 997))
            seq = unit[a:a + n]
            if i == 0:
                lines = [seq[j:j + 60] for j in range(0, n, 60)]
                parts.append(b">r%d d%d\r\n" % (k, k) + b"\r\n".join(lines) + b"\r\n")
            else:
                parts.append(b"@r%d\n%s\n+\n%s\n" % (k, seq, b"I" * n))
            tot += n
            k += 1
        texts.append(b"".join(parts))
    d = tempfile.mkdtemp()
    text_data, status, ids, descs, seqs, nrec = [], [], [], [], [], []
    for i, t in enumerate(texts):
        res = []
        for gz in (False, True):
            path = os.path.join(d, "t%d.fa" % i + (".gz" if gz else ""))
            with (gzip.open(path, "wb") if gz else open(path, "wb")) as f:
                f.write(t)
            res.append(R.read_fastx(path))
        assert res[0] == res[1], i
        r = res[0]
        text_data.append(np.frombuffer(t, np.uint8))
        if isinstance(r, int):
            status.append(r)
            nrec.append(0)
            continue
        status.append(0)
        nrec.append(len(r))
        for a, b, c in r:
            ids.append(np.frombuffer(a, np.uint8))
            descs.append(np.frombuffer(b, np.uint8))
            seqs.append(np.frombuffer(c, np.uint8))
    td, to = concat(text_data, np.uint8)
    idd, ido = concat(ids, np.uint8)
    dd, do = concat(descs, np.uint8)
    sd, so = concat(seqs, np.uint8)
    save("fastx_cases.npz", text=td, text_off=to, status=np.array(status, np.int32),
         nrec=np.array(nrec, np.int64), id=idd, id_off=ido, desc=dd, desc_off=do, seq=sd, seq_off=so)


def msa_keep_length(H, rng):
    """halign4 -c center --keep-length (src/halign4.cpp:96-147) with one
    thread and the copy MSA command, run as the reference program itself
    (oracle/_ref/halign4): the final FASTA for a few center/query files. Queries cover both SAM groups (with and
    without an insertion), a query too short for minimizers, an empty one,
    descriptions, a name starting with '@' (its SAM line is skipped), FASTQ
    input, and a center with lower case and N (aligned as its consensus,
    indexed as written)."""
    d = tempfile.mkdtemp()
    centers, inputs, outputs = [], [], []
    for case in range(5):
        clen = [3000, 5000, 16569, 2000, 4000][case]
        cc = rand_codes(rng, clen)
        center = codes_to_ascii(cc)
        if case == 4:   # scattered N: the queries follow the consensus (A there), so
            # anchors differ between the index of the center as written and of its consensus
            pos = rng.choice(clen, 80, replace=False)
            cc[pos] = 0
            b = bytearray(codes_to_ascii(cc))
            for j in pos:
                b[j] = ord("N")
            center = bytes(b)
        if case == 1:   # lower case and N in the center
            b = bytearray(center)
            for j in rng.integers(0, clen, 40):
                b[j] = ord("n") if j % 3 == 0 else (b[j] | 0x20)
            b[100:160] = b"N" * 60
            center = bytes(b)
        cfa = b">center%d some description\n" % case + b"\n".join(center[j:j + 70] for j in range(0, clen, 70)) + b"\n"
        recs = []
        nq = [30, 40, 10, 25, 40][case]
        for i in range(nq):
            kind = i % 5
            if kind == 0:     # substitutions only: no insertion
                q = mutate_codes(rng, cc, 0.02, 0.0)
            elif kind == 1:   # deletions only
                q = cc.copy()
                for _ in range(3):
                    a = int(rng.integers(0, len(q) - 10))
                    q = np.concatenate([q[:a], q[a + int(rng.integers(1, 8)):]])
            else:
                q = mutate_codes(rng, cc, 0.01, 0.004)
            if i % 7 == 3:    # a part of the center
                a = int(rng.integers(0, clen // 2))
                q = q[a:a + int(rng.integers(200, clen // 2))]
            qs = codes_to_ascii(q)
            if i % 9 == 4:
                qs = qs.lower()
            name = b"q%d" % i
            if i == 5:
                name = b"@at%d" % i
            desc = b" len=%d" % len(qs) if i % 4 == 0 else b""
            recs.append((name, desc, qs))
        recs.append((b"short", b"", b"ACGTACGTAC"))
        recs.append((b"empty", b"", b""))
        if case == 3:   # FASTQ input
            text = b"".join(b"@" + n + dsc + b"\n" + q + b"\n+\n" + b"I" * len(q) + b"\n" for n, dsc, q in recs)
        else:
            text = b"".join(b">" + n + dsc + b"\n" + b"\n".join(q[j:j + 60] for j in range(0, len(q), 60)) +
                            b"\n" for n, dsc, q in recs)
        cpath, ipath, opath = (os.path.join(d, x + "%d" % case) for x in ("c.fa", "q.fa", "out.fa"))
        open(cpath, "wb").write(cfa)
        open(ipath, "wb").write(text)
        rc = H.run(ipath, opath, os.path.join(d, "wd%d" % case), center_path=cpath)
        assert rc == 0, (case, rc, H.last_log[-2000:])
        centers.append(np.frombuffer(cfa, np.uint8))
        inputs.append(np.frombuffer(text, np.uint8))
        outputs.append(np.frombuffer(open(opath, "rb").read(), np.uint8))
    cd, co = concat(centers, np.uint8)
    idd, io = concat(inputs, np.uint8)
    od, oo = concat(outputs, np.uint8)
    save("msa_keep_length.npz", center=cd, center_off=co, input=idd, input_off=io, output=od, output_off=oo)


def mash_cases(R, rng):
    """mash::sketchFromSequence (src/mash/mash.cpp:32-146) over k, sketch size,
    canonical/forward codes and seeds, on sequences with N, lower case and U;
    and RefAligner's choice of a reference by mash::jaccard (first strict
    maximum, src/align/ref_aligner.cpp:228-244)."""
    alpha = np.frombuffer(b"ACGTACGTACGTacgtNUu", np.uint8)
    seqs, par, sk = [], [], []
    for i in range(160):
        size = [2000, 64, 7, 1, 0][i % 5]
        L = int(rng.integers(0, 700 if size == 2000 else 3000))
        seq = alpha[rng.integers(0, len(alpha), L)].tobytes()
        k = [15, 5, 21, 32, 1, 31, 0, 33][i % 8]
        nc = int(bool(i % 3))
        seed = [42, 0, 123456789][i % 3]
        if i in (5, 60):   # full 2,000-hash sketches, RefAligner's parameters
            seq = ACGT[rng.integers(0, 4, 2600)].tobytes()
            k, size, nc, seed = 15, 2000, int(i == 5), 42
        seqs.append(np.frombuffer(seq, np.uint8))
        par.append((k, size, nc, seed))
        sk.append(R.mash_sketch(seq, k, size, nc, seed))
    # reference choice: queries against 5 references, sketches as RefAligner makes them
    base = ACGT[rng.integers(0, 4, 3000)]
    refs = []
    for r in range(5):
        b = base.copy()
        b[rng.integers(0, 3000, 60 * (r + 1))] = ACGT[rng.integers(0, 4, 60 * (r + 1))]
        refs.append(b.tobytes())
    refs[4] = refs[1]   # a tie: the first of two equal references wins
    qs = []
    for i in range(40):
        src = np.frombuffer(refs[i % 5], np.uint8).copy()
        src[rng.integers(0, 3000, 90)] = ACGT[rng.integers(0, 4, 90)]
        a = int(rng.integers(0, 1500))
        qs.append(src[a:a + int(rng.integers(0, 1500))].tobytes())
    qs.append(b"")
    qs.append(b"ACGTN" * 5)
    rsk = [R.mash_sketch(r) for r in refs]
    best, jac = [], []
    for q in qs:
        qsk = R.mash_sketch(q)
        bj, bi = -1.0, 0
        for r, x in enumerate(rsk):
            j = R.mash_jaccard(qsk, x)
            if j > bj:
                bj, bi = j, r
        best.append(bi)
        jac.append(bj)
    sd, so = concat(seqs, np.uint8)
    kd, ko = concat(sk, np.uint64)
    rd, ro = concat([np.frombuffer(r, np.uint8) for r in refs], np.uint8)
    qd, qo = concat([np.frombuffer(q, np.uint8) for q in qs], np.uint8)
    save("mash_cases.npz", seq=sd, seq_off=so, params=np.array(par, np.int64), sketch=kd, sketch_off=ko,
         ref=rd, ref_off=ro, qry=qd, qry_off=qo, best=np.array(best, np.int32), jaccard=np.array(jac, np.float64))


def msa_refs(H, rng):
    """halign4 -c refs --keep-length with several references, one thread,
    run as the reference program itself: each query goes to the reference with the
    largest sketch Jaccard, its row spread over that reference's aligned
    record. Case 0: three equal-length references, the copy command; case 1:
    four references of different lengths and an aligned file with gap columns
    (what an MSA command would write); case 2: FASTQ queries and two identical
    references (a Jaccard tie)."""
    d = tempfile.mkdtemp()
    out = {k: [] for k in ("center", "aligned", "input", "output")}
    has_aligned = []
    for case in range(3):
        W = [3000, 4000, 2500][case]
        base = ACGT[rng.integers(0, 4, W)]
        nref = [3, 4, 3][case]
        refs, rows = [], []
        for r in range(nref):
            b = base.copy()
            b[rng.integers(0, W, 40 * (r + 1))] = ACGT[rng.integers(0, 4, 40 * (r + 1))]
            present = np.ones(W, bool)
            if case == 1:   # each reference lacks some columns
                for _ in range(3):
                    a = int(rng.integers(0, W - 60))
                    present[a:a + int(rng.integers(1, 50))] = False
            refs.append(b[present].tobytes())
            rows.append(np.where(present, b, ord("-")).astype(np.uint8).tobytes())
        if case == 2:
            refs[2] = refs[0]
            rows[2] = rows[0]
        names = [b"ref%d" % r for r in range(nref)]
        cfa = b"".join(b">" + n + b" r\n" + x + b"\n" for n, x in zip(names, refs))
        afa = b"".join(b">" + n + b"\n" + x + b"\n" for n, x in zip(names, rows)) if case == 1 else None
        recs = []
        for i in range(36):
            lut = np.zeros(256, np.uint8)
            lut[ACGT] = np.arange(4, dtype=np.uint8)
            src = lut[np.frombuffer(refs[i % nref], np.uint8)]
            q = mutate_codes(rng, src, 0.01, 0.0 if i % 4 == 0 else 0.004)
            if i % 6 == 5:
                a = int(rng.integers(0, len(q) // 2))
                q = q[a:a + int(rng.integers(300, len(q) // 2))]
            qs = codes_to_ascii(q)
            name = b"@q%d" % i if i == 7 else b"q%d" % i
            recs.append((name, b" d" if i % 5 == 0 else b"", qs))
        recs.append((b"empty", b"", b""))
        recs.append((b"short", b"", b"ACGTACGTAC"))
        if case == 2:
            text = b"".join(b"@" + n + dsc + b"\n" + q + b"\n+\n" + b"I" * len(q) + b"\n" for n, dsc, q in recs)
        else:
            text = b"".join(b">" + n + dsc + b"\n" + q + b"\n" for n, dsc, q in recs)
        cp, ap, ip, op = (os.path.join(d, x + "%d" % case) for x in ("c.fa", "a.fa", "q.fa", "o.fa"))
        open(cp, "wb").write(cfa)
        open(ip, "wb").write(text)
        if afa is not None:
            open(ap, "wb").write(afa)
        rc = H.run(ip, op, os.path.join(d, "wd%d" % case), center_path=cp,
                   msa_cmd=msa_command(ap) if afa is not None else Halign4.COPY)
        assert rc == 0, (case, rc, H.last_log[-2000:])
        out["center"].append(np.frombuffer(cfa, np.uint8))
        out["aligned"].append(np.frombuffer(afa or b"", np.uint8))
        out["input"].append(np.frombuffer(text, np.uint8))
        out["output"].append(np.frombuffer(open(op, "rb").read(), np.uint8))
        has_aligned.append(afa is not None)
    arrays = {}
    for k, v in out.items():
        arrays[k], arrays[k + "_off"] = concat(v, np.uint8)
    save("msa_refs.npz", has_aligned=np.array(has_aligned, np.int32), **arrays)


def msa_command(aligned_path):
    """An MSA command template for the reference program whose output is the
    given aligned file for an input of several records, and the input itself
    for a single record (what any MSA tool writes for one sequence): the
    reference runs the command on the center file (RefAligner's constructor,
    src/align/ref_aligner.cpp:47-60) and, in the auto-center path, first on
    the Top-K file (src/halign4.cpp:119-134), then on the one-record
    consensus file."""
    return ("if [ \"$(grep -c '>' {input})\" = 1 ]; then cp {input} {output}; else cp %s {output}; fi"
            % aligned_path)


def auto_center(R, H, rng):
    """The auto-center path with --keep-length: preprocessInputFasta's Top-K
    consensus input (src/utils/preprocess.cpp:19-160), generateConsensusSequence's
    FASTA and JSON outputs (src/consensus/consensus.cpp:38-68, 611-703) and the
    whole run (src/halign4.cpp:87, 119-147) with the copy MSA command or a
    given aligned Top-K file."""
    d = tempfile.mkdtemp()
    out = {k: [] for k in ("topk_in", "topk_out", "cons_in", "cons_fa", "cons_json", "run_in", "run_aligned",
                           "run_out")}
    topk_k, cons_par, run_par = [], [], []
    alpha = np.frombuffer(b"ACGTACGTacgtNnRY", np.uint8)
    # Top-K: lengths with ties, cleaning of lower case and IUPAC
    for case, (n, k) in enumerate([(60, 7), (30, 30), (12, 50), (40, 1), (25, 0)]):
        lens = rng.integers(1, 400, n)
        lens[rng.integers(0, n, n // 3)] = 333   # ties
        recs = b"".join(b">s%d%s\n" % (i, b" x" if i % 3 == 0 else b"") + alpha[rng.integers(0, len(alpha), int(L))].tobytes()
                        + b"\n" for i, L in enumerate(lens))
        ip = os.path.join(d, "t%d.fa" % case)
        open(ip, "wb").write(recs)
        wd = os.path.join(d, "tw%d" % case)
        os.makedirs(wd)
        assert R.preprocess(ip, wd, k) == n
        out["topk_in"].append(np.frombuffer(recs, np.uint8))
        out["topk_out"].append(np.frombuffer(open(os.path.join(wd, "data", "clean_data", "consensus_unaligned.fasta"),
                                                   "rb").read(), np.uint8))
        topk_k.append(k)
    # consensus files: gaps, dots, lower case, the batch drop and seq_limit
    calpha = np.frombuffer(b"ACGTUN-.acgtuRYX", np.uint8)
    for case, (nrows, aln, batch, lim) in enumerate([(1, 7, 4096, 0), (5, 90, 4096, 0), (30, 33, 4, 0),
                                                     (30, 33, 4, 9), (4100, 5, 4096, 0), (17, 161, 3, 12)]):
        rows = calpha[rng.integers(0, len(calpha), (nrows, aln))]
        text = b"".join(b">r%d\n" % i + rows[i].tobytes() + b"\n" for i in range(nrows))
        fa = os.path.join(d, "c%d.fa" % case)
        open(fa, "wb").write(text)
        R.consensus_file(fa, fa + ".out.fa", fa + ".out.json", seq_limit=lim, batch=batch)
        out["cons_in"].append(np.frombuffer(text, np.uint8))
        out["cons_fa"].append(np.frombuffer(open(fa + ".out.fa", "rb").read(), np.uint8))
        out["cons_json"].append(np.frombuffer(open(fa + ".out.json", "rb").read(), np.uint8))
        cons_par.append((lim, batch))
    # whole runs: the cons_n longest queries share one length (copy command), or
    # an aligned Top-K file with gap columns
    for case in range(2):
        W = 2500
        base = ACGT[rng.integers(0, 4, W)]
        recs, aligned = [], []
        ktop = 5
        for i in range(ktop):
            b = base.copy()
            b[rng.integers(0, W, 30)] = ACGT[rng.integers(0, 4, 30)]
            present = np.ones(W, bool)
            if case == 1:
                a = int(rng.integers(0, W - 100))
                present[a:a + int(rng.integers(1, 60))] = False
            recs.append((b"long%d" % i, b[present].tobytes() if case == 1 else b.tobytes()))
            aligned.append(np.where(present, b, ord("-")).astype(np.uint8).tobytes())
        lut = np.zeros(256, np.uint8)
        lut[ACGT] = np.arange(4, dtype=np.uint8)
        for i in range(20):
            q = mutate_codes(rng, lut[base], 0.01, 0.004)
            L = int(rng.integers(300, W - 200))   # shorter than the Top-K records
            a = int(rng.integers(0, len(q) - L))
            recs.append((b"q%d" % i, codes_to_ascii(q[a:a + L])))
        if case == 1:   # the longest records must come first by length: keep them longest
            recs[ktop:] = [(n, x[:W - 300]) for n, x in recs[ktop:]]
        text = b"".join(b">" + n + b"\n" + x + b"\n" for n, x in recs)
        ip = os.path.join(d, "r%d.fa" % case)
        open(ip, "wb").write(text)
        ap = None
        atext = b""
        if case == 1:
            # the aligned Top-K in the order the Top-K file lists them (longest first)
            order = sorted(range(ktop), key=lambda i: (-len(recs[i][1]), i))
            atext = b"".join(b">" + recs[i][0] + b"\n" + aligned[i] + b"\n" for i in order)
            ap = os.path.join(d, "r%d.aligned.fa" % case)
            open(ap, "wb").write(atext)
        op = os.path.join(d, "r%d.out.fa" % case)
        rc = H.run(ip, op, os.path.join(d, "rw%d" % case), keep_length=True, cons_n=ktop,
                   msa_cmd=msa_command(ap) if ap else Halign4.COPY)
        assert rc == 0, (case, rc, H.last_log[-2000:])
        out["run_in"].append(np.frombuffer(text, np.uint8))
        out["run_aligned"].append(np.frombuffer(atext, np.uint8))
        out["run_out"].append(np.frombuffer(open(op, "rb").read(), np.uint8))
        run_par.append((ktop, int(ap is not None)))
    arrays = {}
    for k, v in out.items():
        arrays[k], arrays[k + "_off"] = concat(v, np.uint8)
    save("auto_center.npz", topk_k=np.array(topk_k, np.int32), cons_par=np.array(cons_par, np.int64),
         run_par=np.array(run_par, np.int32), **arrays)


def workload_lengths(cfg, first, count, chunk=50000):
    from halign import synth
    out = []
    for a in range(first, first + count, chunk):
        _, off = synth.queries(cfg, a, min(chunk, first + count - a))
        out.append(np.diff(off))
    return np.concatenate(out)


def sample_workload(cfg, rng):
    """Query indices of the sample of workload cfg: the first 64 (the start of
    every bench step), random ones over the whole workload, and the cases each
    workload exists to stress, forced in from a scan of its first queries."""
    from halign import synth
    total = synth.CONFIG_COUNTS[cfg]
    first = list(range(64))
    if cfg == 3:
        L = workload_lengths(3, 0, 200000)
        short = np.flatnonzero(L < 550)          # anchor-less whole pairs among them
        longest = np.argsort(L, kind="stable")[-200:]
        forced = list(rng.choice(short, 300, replace=False)) + list(longest)
        n_rand = 1436
    elif cfg == 4:                               # every query carries ~2 long indels
        forced, n_rand = [], 936
    else:
        L = workload_lengths(5, 0, 200000)
        tiny = np.flatnonzero(L < 20)            # no minimizers: whole-pair DP
        longer = np.flatnonzero(L > 29903)       # longer than the center
        edge = np.flatnonzero((L >= 14800) & (L <= 15100))   # auto band switches near 14,951
        forced = (list(rng.choice(tiny, 40, replace=False)) + list(rng.choice(longer, 60, replace=False)) +
                  list(rng.choice(edge, 40, replace=False)))
        n_rand = 1296
    rand = list(rng.choice(np.arange(64, total), n_rand, replace=False))
    idx, seen = [], set()
    for i in first + [int(x) for x in forced] + [int(x) for x in rand]:
        if i not in seen:
            seen.add(i)
            idx.append(i)
    return np.array(idx, np.int64)


def config_workload(H, cfg, rng):
    """RefAligner on a sample of workload cfg through the reference program
    (halign4 -c center --keep-length -t 1, copy command: src/halign4.cpp:
    96-147, src/align/ref_aligner.cpp:212-283, 286-394, 676-740): the CIGAR
    and SAM group (cigar::hasInsertion) of every query, read back from its
    thread-0 SAM files, and the SHA-256 of the final aligned FASTA."""
    import hashlib
    from halign import synth
    idx = sample_workload(cfg, rng)
    q, off = workload_queries(cfg, idx)
    d = tempfile.mkdtemp()
    cpath, ipath, opath = (os.path.join(d, x) for x in ("center.fa", "q.fa", "out.fa"))
    open(cpath, "wb").write(b">center\n" + synth.center(cfg) + b"\n")
    with open(ipath, "wb") as f:
        for k, i in enumerate(idx):
            f.write(b">q%d\n" % i + q[off[k]:off[k + 1]] + b"\n")
    rc = H.run(ipath, opath, os.path.join(d, "wd"), center_path=cpath)
    assert rc == 0, (cfg, rc, H.last_log[-2000:])
    recs = {name: (cig, ins) for name, _, cig, ins in Halign4.sam_records(os.path.join(d, "wd"))}
    assert len(recs) == len(idx)
    cigs, ins = [], []
    for i in idx:
        c, h = recs[b"q%d" % i]
        cigs.append(c)
        ins.append(h)
    cd, co = concat(cigs, np.uint32)
    digest = np.frombuffer(hashlib.sha256(open(opath, "rb").read()).digest(), np.uint8)
    save("config%d.npz" % cfg, idx=idx, cigar=cd, cigar_off=co, has_insertion=np.array(ins, np.uint8),
         fasta_sha256=digest, fasta_bytes=np.array([os.path.getsize(opath)], np.int64))


def sam_cases(H, rng):
    """The per-thread SAM files of RefAligner::alignQueryToRef with one thread
    (src/align/ref_aligner.cpp:286-394, routing :212-283), written by the
    reference program itself (halign4 -t 1 --save-workdir; result/thread0.sam
    and result/thread0_insertion.sam). Cases: one center with and without
    --keep-length (--cons-n 1 so RefAligner runs); three equal-length
    references (copy command) with and without --keep-length; four references
    with missing columns and their aligned file, with and without
    --keep-length. Without --keep-length and several references every query is
    aligned again to the consensus (:268-282); the queries include copies of
    full-width references, whose first alignment (to a reference that lacks
    some columns) has an insertion while the second (to the full-width
    consensus) may not. Queries also cover an '@' name, an empty and a short
    query, lower case and N. A run whose merge fails after the SAM files are
    written (the copy command leaves insertion records of several lengths)
    still counts: its rc is stored. Each reference set's inputs run with
    --keep-length (case 2k) and without (case 2k+1)."""
    d = tempfile.mkdtemp()
    out = {k: [] for k in ("center", "aligned", "input", "sam", "sam_ins")}
    par = []
    lut = np.zeros(256, np.uint8)
    lut[ACGT] = np.arange(4, dtype=np.uint8)
    for kind in range(3):          # 0: one center, 1: three references, 2: four with missing columns
        W = [2500, 3000, 2800][kind]
        base = ACGT[rng.integers(0, 4, W)]
        nref = [1, 3, 4][kind]
        refs, rows = [], []
        for r in range(nref):
            b = base.copy()
            if r:
                b[rng.integers(0, W, 30 * r)] = ACGT[rng.integers(0, 4, 30 * r)]
            present = np.ones(W, bool)
            if kind == 2 and r % 2 == 1:   # odd references lack some columns
                for _ in range(3):
                    a = int(rng.integers(0, W - 80))
                    present[a:a + int(rng.integers(5, 60))] = False
            refs.append(b[present].tobytes())
            rows.append(np.where(present, b, ord("-")).astype(np.uint8).tobytes())
        names = [b"ref%d" % r for r in range(nref)]
        cfa = b"".join(b">" + n + b" r\n" + x + b"\n" for n, x in zip(names, refs))
        afa = b"".join(b">" + n + b"\n" + x + b"\n" for n, x in zip(names, rows)) if kind == 2 else None
        recs = []
        for i in range(40):
            src = refs[i % nref]
            if kind == 2 and i % 5 == 0:
                # full width with a short reference's substitutions: its Jaccard picks that
                # reference (an insertion where it lacks columns); the consensus has them
                row = np.frombuffer(rows[1 + 2 * ((i // 5) % 2)], np.uint8)
                src = np.where(row == ord("-"), base, row).astype(np.uint8).tobytes()
            q = mutate_codes(rng, lut[np.frombuffer(src, np.uint8)], 0.01, 0.0 if i % 3 == 0 else 0.004)
            if i % 7 == 6:
                a = int(rng.integers(0, len(q) // 2))
                q = q[a:a + int(rng.integers(300, len(q) // 2))]
            qs = codes_to_ascii(q)
            if i % 11 == 4:
                qs = qs.lower()
            if i % 13 == 5:
                b = bytearray(qs)
                b[10:14] = b"NNNN"
                qs = bytes(b)
            name = b"@q%d" % i if i == 9 else b"q%d" % i
            recs.append((name, qs))
        recs.append((b"empty", b""))
        recs.append((b"short", b"ACGTACGTAC"))
        text = b"".join(b">" + n + b"\n" + q + b"\n" for n, q in recs)
        for keep in (True, False):
            case = 2 * kind + (0 if keep else 1)
            cp, ap, ip, op = (os.path.join(d, x + "%d" % case) for x in ("c.fa", "a.fa", "q.fa", "o.fa"))
            open(cp, "wb").write(cfa)
            open(ip, "wb").write(text)
            if afa is not None:
                open(ap, "wb").write(afa)
            wd = os.path.join(d, "wd%d" % case)
            rc = H.run(ip, op, wd, center_path=cp, keep_length=keep, cons_n=None if keep else 1,
                       msa_cmd=msa_command(ap) if afa is not None else Halign4.COPY)
            res = os.path.join(wd, "result")
            sam, sam_ins = (open(os.path.join(res, f), "rb").read() for f in ("thread0.sam", "thread0_insertion.sam"))
            assert sam.startswith(b"@HD") and sam_ins.startswith(b"@HD"), (case, rc, H.last_log[-1500:])
            out["center"].append(np.frombuffer(cfa, np.uint8))
            out["aligned"].append(np.frombuffer(afa or b"", np.uint8))
            out["input"].append(np.frombuffer(text, np.uint8))
            out["sam"].append(np.frombuffer(sam, np.uint8))
            out["sam_ins"].append(np.frombuffer(sam_ins, np.uint8))
            par.append((int(keep), int(afa is not None), nref, rc))
            nl = (sam.count(b"\n") - 1, sam_ins.count(b"\n") - 1)
            print(f"    sam case {case}: nref {nref} keep {keep} rc {rc} lines {nl[0]}+{nl[1]}")
    arrays = {}
    for k, v in out.items():
        arrays[k], arrays[k + "_off"] = concat(v, np.uint8)
    save("sam_cases.npz", par=np.array(par, np.int32), **arrays)


FIXTURES = ["ksw_cases", "global_align", "minimizers", "anchors_chain", "ref_global", "consensus",
            "whole_pairs", "fastx_cases", "msa_keep_length", "mash_cases", "msa_refs", "auto_center",
            "config3", "config4", "config5", "sam_cases"]


def main(which):
    R = RefLib()
    O = Oracle()
    H = Halign4()
    print("writing golden fixtures from", R.lib._name, "and", H.path, flush=True)
    run = {
        "ksw_cases": lambda: ksw_cases(R, O, np.random.default_rng(0x601D)),
        "global_align": lambda: global_align(R, np.random.default_rng(0x601E)),
        "minimizers": lambda: minimizers(R, O, np.random.default_rng(0x601F)),
        "anchors_chain": lambda: anchors_chain_mm2(R, np.random.default_rng(0x6020)),
        "ref_global": lambda: ref_global(R, np.random.default_rng(0x6021)),
        "consensus": lambda: consensus(R, np.random.default_rng(0x6022)),
        "whole_pairs": lambda: whole_pairs(R),
        "fastx_cases": lambda: fastx_cases(R, np.random.default_rng(0x6023)),
        "msa_keep_length": lambda: msa_keep_length(H, np.random.default_rng(0x6024)),
        "mash_cases": lambda: mash_cases(R, np.random.default_rng(0x6025)),
        "msa_refs": lambda: msa_refs(H, np.random.default_rng(0x6026)),
        "auto_center": lambda: auto_center(R, H, np.random.default_rng(0x6027)),
        "config3": lambda: config_workload(H, 3, np.random.default_rng(0x60C3)),
        "config4": lambda: config_workload(H, 4, np.random.default_rng(0x60C4)),
        "config5": lambda: config_workload(H, 5, np.random.default_rng(0x60C5)),
        "sam_cases": lambda: sam_cases(H, np.random.default_rng(0x6028)),
    }
    for name in which or FIXTURES:
        run[name]()


if __name__ == "__main__":
    main(sys.argv[1:])
