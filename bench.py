#!/usr/bin/env python3
"""Kernel-only pairwise DP benchmark on one MI355X.

Workload: 100,000 whole query-vs-center pairs (center 29,903 bp, queries
29.0-30.0 kb at 0.5% substitutions / 0.1% indels, seeded synthetic data),
globalAlignKSW2 semantics (auto band ~4,665, RIGHT-aligned gaps, score +
CIGAR on the device). Inputs are resident in HBM before timing starts.

A step is one batch of --batch pairs (the batch index walks through the
100,000-pair set). value = batched-persistent GCUPS over the timed steps,
cells = sum of the exact in-band cells of ksw2 (en0 - st0 + 1 per
anti-diagonal). The plain one-pair-per-launch form is timed on a few pairs
beside it, and the reference CPU path (compiled reference if present, else
the oracle port) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "halign-4_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "GCUPS + seqs/sec, batched vs plain GPU form, bit-exact; p99 latency serving"
N_PAIRS = 100000
OPS_PER_CELL = 24                 # algorithmic integer ops per DP cell (recurrence + H + max)
PEAK_VALU_TOPS = 78.6             # MI355X packed-16-bit integer VALU: 256 CU x 128 lanes x 2.4 GHz


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (a shared GPU host grants each GPU a
    share of its cores that way)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else None, "cpu_model": model}


def cpu_baseline(center, qbytes, off, sample_pairs, threads):
    """Time globalAlignKSW2 on `sample_pairs` pairs with `threads` host threads."""
    from oracle_lib import REF_SO, ORACLE_SO, Oracle, RefLib, FLAG_RIGHT
    kind = "reference" if os.path.exists(REF_SO) else "port"
    lib = RefLib() if kind == "reference" else Oracle()

    def one(i):
        q = qbytes[off[i]:off[i + 1]]
        lib.global_align_ksw2(center, q, FLAG_RIGHT)
        return 0

    # cells of the sample (closed form of the band, same as the kernel counts)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(one, range(sample_pairs)))
    dt = time.perf_counter() - t0
    return kind, dt


def verify_dp(ps, center, qbytes, off, last_range, k=0):
    """After timing: the score and CIGAR of k pairs of the last timed launch
    (spread over it, so pairs that reused a persistent workgroup's traceback
    slot are among them) against the compiled reference's globalAlignKSW2
    (the oracle port where the reference build is absent), on host threads.
    Returns "ok/checked" and the checker's kind."""
    from oracle_lib import REF_SO, FLAG_RIGHT, Oracle, RefLib
    kind = "reference" if os.path.exists(REF_SO) else "port"
    lib = RefLib() if kind == "reference" else Oracle()
    first, cnt = last_range
    k = k or 8
    idx = sorted(set(int(first + j) for j in np.linspace(0, cnt - 1, min(k, cnt))))

    got = {i: ps.result(i)[1] for i in idx}

    def one(i):
        exp = lib.global_align_ksw2(center, qbytes[off[i]:off[i + 1]], FLAG_RIGHT)
        exp = exp[0] if isinstance(exp, tuple) else exp
        return bool(np.array_equal(np.asarray(got[i], np.uint32), np.asarray(exp, np.uint32)))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(min(len(idx), host_threads())) as ex:
        ok = sum(ex.map(one, idx))
    log(f"verified {ok}/{len(idx)} pairs of the last launch against the {kind} ({time.perf_counter() - t0:.1f}s)")
    return {"ok": f"{ok}/{len(idx)}", "checker": kind, "what": "score+CIGAR of pairs spread over the last timed launch"}


def verify_msa(ctx, d, cpath, qv, off, n, refs):
    """The keep-length file run on the first n queries against the
    reference program's own run (oracle/_ref/halign4 -c --keep-length -t 1,
    copy command) on the same files: the final FASTA byte for byte."""
    import halign
    from oracle_lib import REF_BIN, Halign4
    if not os.path.exists(REF_BIN):
        log("verify: no reference program here, msa output not checked")
        return None
    vd = os.path.join(d, "verify")
    os.makedirs(vd, exist_ok=True)
    ip = os.path.join(vd, "q.fa")
    with open(ip, "wb") as f:
        for i in range(n):
            f.write(b">q%d\n" % i + bytes(qv[off[i]:off[i + 1]]) + b"\n")
    halign.msa_keep_length(ctx, cpath, ip, os.path.join(vd, "gpu.fa"))
    rc = Halign4().run(ip, os.path.join(vd, "ref.fa"), os.path.join(vd, "w"), center_path=cpath)
    same = rc == 0 and open(os.path.join(vd, "gpu.fa"), "rb").read() == open(os.path.join(vd, "ref.fa"), "rb").read()
    log(f"verified msa run of {n} queries ({refs} reference(s)) against the reference program: "
        f"{'identical' if same else 'DIFFERENT'}")
    return {"ok": f"{int(same)}/1", "checker": "reference program",
            "what": f"final FASTA of a {n}-query run, byte for byte"}


def verify_queries(center, qb, off, items, what):
    """RefAligner::globalAlign CIGARs of the GPU (items: (query index, CIGAR))
    against the compiled reference (center minimizers computed once, as
    RefAligner does), or the oracle port where the reference build is absent."""
    from oracle_lib import REF_SO, Oracle, RefLib
    kind = "reference" if os.path.exists(REF_SO) else "port"
    lib = RefLib() if kind == "reference" else Oracle()
    hits = lib.minimizers(center) if kind == "reference" else None

    def one(it):
        i, got = it
        q = qb[off[i]:off[i + 1]]
        exp = lib.ref_global_align_pre(center, hits, q) if kind == "reference" else lib.ref_global_align(center, q)[0]
        return bool(np.array_equal(np.asarray(got, np.uint32), np.asarray(exp, np.uint32)))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(host_threads()) as ex:
        ok = sum(ex.map(one, items))
    log(f"verified {ok}/{len(items)} {what} against the {kind} ({time.perf_counter() - t0:.1f}s)")
    return {"ok": f"{ok}/{len(items)}", "checker": kind, "what": what}


def band_cells(qlen, tlen, w):
    """sum_r (en0 - st0 + 1) of ksw2's band (src/align/ksw2_extz2_sse.c:107-110)."""
    if w < 0:
        w = max(qlen, tlen)
    r = np.arange(qlen + tlen - 1, dtype=np.int64)
    st = np.maximum.reduce([np.zeros_like(r), r - qlen + 1, (r - w + 1) >> 1])
    en = np.minimum.reduce([np.full_like(r, tlen - 1), r, (r + w) >> 1])
    bad = np.nonzero(st > en)[0]
    if len(bad):
        st, en = st[:bad[0]], en[:bad[0]]
    return int((en - st + 1).sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--plain-pairs", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=0, help="pairs for the CPU baseline (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true", help="skip the check of sampled outputs after timing")
    ap.add_argument("--verify", type=int, default=0,
                    help="outputs checked after timing (0 = the workload default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 times), to scale HBM counters")
    ap.add_argument("--workload", default="dp", choices=["dp", "consensus", "e2e", "serve", "msa"],
                    help="dp: the headline pairwise DP; consensus: column counting over rows in HBM; "
                         "e2e: the whole per-query path (--config 1, 3 or 4); serve: config 5 serving")
    ap.add_argument("--config", type=int, default=1, help="e2e workload configuration (1, 3, 4)")
    ap.add_argument("--queries", type=int, default=0, help="e2e: queries per step (0 = the config default)")
    ap.add_argument("--rows", type=int, default=100000, help="consensus workload: rows of 29,903 columns")
    ap.add_argument("--requests", type=int, default=0, help="serve: requests per load point (0 = --serve-queries)")
    ap.add_argument("--serve-queries", type=int, default=2000000, help="serve: queries per load point")
    ap.add_argument("--pool", type=int, default=262144, help="serve: distinct queries the requests draw on")
    ap.add_argument("--workers", type=int, default=3,
                    help="serve: concurrent batches (3: the most capacity and the lowest tail latency at equal load "
                         "of 2, 3 and 4 measured)")
    ap.add_argument("--batch-queries", type=int, default=8192, help="serve: queries per batch at most")
    ap.add_argument("--loads", type=float, nargs="*", default=[0.5, 0.65, 0.8, 0.95],
                    help="serve: offered load as fractions of the measured capacity")
    ap.add_argument("--refs", type=int, default=1,
                    help="msa: reference sequences in the center file (>1: several references, chosen by Jaccard)")
    ap.add_argument("--msa-batch", type=int, default=0,
                    help="msa: queries per chunk (0 = the library default, alignQueryToRef's 25600)")
    ap.add_argument("--kernel", default="auto",
                    help="auto | multiwave | WxS (ring kernel with W waves and S slots per lane; "
                         "WxSd: its two-per-SIMD form)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "the path shards by query ranges with no collective, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import halign
        halign.lib()
        n = args.copy_calibration << 20
        a, b = DeviceBuffer(n), DeviceBuffer(n)
        a.upload(np.ones(n, np.uint8))
        for _ in range(3):
            b.copy_from(a)
        a.free()
        b.free()
        print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
        return

    if args.workload == "consensus":
        return consensus_bench(args)
    if args.workload == "e2e":
        return e2e_bench(args)
    if args.workload == "serve":
        return serve_bench(args)
    if args.workload == "msa":
        return msa_bench(args)

    import halign
    from halign import synth

    need = (args.warmup + args.steps) * args.batch
    n_res = min(N_PAIRS, max(need, args.batch))
    log(f"generating config-2 workload: center 29903 bp, {n_res} of {N_PAIRS} queries resident")
    t0 = time.time()
    center = synth.center(2)
    qbytes, off = synth.queries(2, 0, n_res)
    log(f"  generated {len(qbytes) / 1e9:.2f} GB in {time.time() - t0:.1f}s")
    ctx = halign.Context(0)
    if args.kernel == "auto":
        policy = 0
    elif args.kernel == "multiwave":
        policy = 1
    else:
        k = args.kernel.lower()
        wv, sl = (int(x) for x in k.rstrip("d").split("x"))
        policy = halign.ring_policy(wv, sl, dense=k.endswith("d"))
    ctx.set_kernel_policy(policy)
    ps = halign.PairSet(ctx, center, qbytes, off)

    # plain form: one pair per launch
    plain_cells = 0
    plain_ms = 0.0
    for i in range(args.plain_pairs):
        ms, cells = ps.run(i, 1, plain=True)
        plain_ms += ms
        plain_cells += cells
    plain_gcups = plain_cells / (plain_ms * 1e-3) / 1e9 if plain_ms > 0 else 0.0
    log(f"plain form: {args.plain_pairs} pairs, {plain_gcups:.2f} GCUPS")

    for s in range(args.warmup):
        ps.run((s * args.batch) % n_res, min(args.batch, n_res), plain=False)
    log("warmup done")
    tot_cells = 0
    tot_kernel_ms = 0.0
    pairs = 0
    ranges = []
    t_start = time.perf_counter()
    for s in range(args.steps):
        first = ((args.warmup + s) * args.batch) % n_res
        cnt = min(args.batch, n_res - first)
        ms, cells = ps.run(first, cnt, plain=False)
        tot_cells += cells
        tot_kernel_ms += ms
        pairs += cnt
        ranges.append((first, cnt))
        log(f"step {s}: {cnt} pairs, kernel {ms:.1f} ms, {cells / (ms * 1e-3) / 1e9:.1f} GCUPS")
    t_wall = time.perf_counter() - t_start
    verified = None if args.no_verify else verify_dp(ps, center, qbytes, off, ranges[-1], args.verify)
    gcups = tot_cells / t_wall / 1e9
    kernel_gcups = tot_cells / (tot_kernel_ms * 1e-3) / 1e9
    achieved_tops = tot_cells * OPS_PER_CELL / (tot_kernel_ms * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu:
        threads = host_threads()
        sample = args.cpu_sample or threads * 32
        kind, dt = cpu_baseline(center, qbytes, off, sample, threads)
        w = [halign_auto_band(len(center), int(off[i + 1] - off[i])) for i in range(sample)]
        cells = sum(band_cells(int(off[i + 1] - off[i]), len(center), w[i]) for i in range(sample))
        cpu = {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": threads, "kind": kind,
               "sample": f"{sample} config-2 whole pairs (globalAlignKSW2), {dt:.1f}s on {threads} threads",
               "host": host_info()}
        log(f"cpu baseline ({kind}): {cpu['value']} GCUPS")

    out = {
        "metric": METRIC,
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic (seeded, csrc/synth.c config 2)",
        "config": {"workload": "config 2: kernel-only pairwise DP, 100000 pairs, center 29903 bp, "
                               "queries 29.0-30.0 kb, auto band, batched persistent form",
                   "pairs_per_step": args.batch, "pairs_timed": pairs, "kernel": args.kernel},
        "verified": verified,
        "seqs_per_s": round(pairs / t_wall, 2),
        "kernel_gcups": round(kernel_gcups, 3),
        "plain_gcups": round(plain_gcups, 4),
        "batched_vs_plain": round(kernel_gcups / plain_gcups, 1) if plain_gcups else None,
        "roofline": {"bound": "valu", "achieved": round(achieved_tops, 3), "peak": PEAK_VALU_TOPS,
                     "unit": "TOPS", "frac": round(achieved_tops / PEAK_VALU_TOPS, 4),
                     "traffic": traffic_per_launch(tot_cells / args.steps)},
        "cpu_baseline": cpu,
        "hip_graph": dict(zip(("captures", "replays"), ctx.graph_counts())),
    }
    print(json.dumps(out), flush=True)


E2E = {1: ("plain method, small: center 16569 bp, 1000 queries at 1% sub / 0.2% indel", 1000, 1000),
       3: ("uneven batch: center 20000 bp, 1000000 lognormal-length queries (median 1.5 kb)", 65536, 1000000),
       4: ("divergent stress: center 10000 bp, 200000 queries at 5% sub / 1% indel + long indels", 32768, 200000)}


def e2e_bench(args):
    """RefAligner::globalAlign per query (minimizers, anchors, chain, MM2 with
    its segment DPs) against a resident center, one C-ABI call per step. The
    queries are passed as host buffers: the wall time includes their upload
    and the CIGAR download. cpu_baseline: the compiled reference's
    RefAligner::globalAlign (center minimizers computed once) on a sample."""
    import halign
    from halign import synth
    cfg = args.config
    label, per_step, total = E2E[cfg]
    per = args.queries or per_step
    need = min(total, per * args.steps)
    center = synth.center(cfg)
    qb, off = synth.queries(cfg, 0, need)
    log(f"e2e config {cfg}: {need} queries generated, {off[-1] / 1e6:.1f} Mbases")
    ctx = halign.Context(0)
    c = halign.Center(ctx, center)
    ranges = [((s * per) % need, min(need, (s * per) % need + per)) for s in range(args.steps)]
    # warmup: every timed range once per warmup pass (device buffers and the
    # traceback pool reach their size, launch plans are captured)
    for w in range(args.warmup):
        for a, b in ranges:
            c.align_flat(qb, off[a:b + 1])
    steps = []
    for s, (a, b) in enumerate(ranges):
        t0 = time.perf_counter()
        cig, coff, st, ms = c.align_flat(qb, off[a:b + 1])
        dt = time.perf_counter() - t0
        last = (a, b, cig, coff)
        steps.append((b - a, int(off[b] - off[a]), dt, ms))
        log(f"step {s}: {b - a} queries, {dt * 1e3:.1f} ms wall, {ms:.1f} ms device")
    nq = sum(x[0] for x in steps)
    nb = sum(x[1] for x in steps)
    tw = sum(x[2] for x in steps)
    # one more step (not timed) with every DP launch timed on its own
    buckets = dp_bucket_report(ctx, lambda: c.align_flat(qb, off[0:min(per, need) + 1]))
    verified = None
    if not args.no_verify:
        a, b, cig, coff = last
        picks = sorted(set(int(j) for j in np.linspace(0, b - a - 1, min(args.verify or 64, b - a))))
        verified = verify_queries(center, qb, off, [(a + j, cig[coff[j]:coff[j + 1]]) for j in picks],
                                  "queries spread over the last timed step")
    cpu = None
    if not args.no_cpu:
        from oracle_lib import REF_SO, Oracle, RefLib
        threads = host_threads()
        kind = "reference" if os.path.exists(REF_SO) else "port"
        lib = RefLib() if kind == "reference" else Oracle()
        if kind == "reference":
            hits = lib.minimizers(center)
            fn = lambda i: lib.ref_global_align_pre(center, hits, qb[off[i]:off[i + 1]])
        else:
            fn = lambda i: lib.ref_global_align(center, qb[off[i]:off[i + 1]])
        sample, dt = cpu_timed(fn, need, threads, args.cpu_sample)
        cpu = {"value": round(sample / dt, 2), "unit": "seqs/s", "cores": threads, "kind": kind,
               "sample": f"{sample} config-{cfg} queries (RefAligner::globalAlign, the generated queries "
                         f"in order, repeated if fewer), {dt:.1f}s on {threads} threads"}
        log(f"cpu baseline ({kind}): {cpu['value']} seqs/s")
    out = {"metric": "seqs/s end-to-end per-query path (RefAligner::globalAlign)", "value": round(nq / tw, 2),
           "unit": "seqs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(tw * 1e3 / len(steps), 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int8", "data": f"synthetic (seeded, csrc/synth.c config {cfg})",
           "config": {"workload": f"config {cfg}: {label}", "queries_per_step": per},
           "verified": verified,
           "bases_per_s": round(nb / tw, 1), "device_ms_per_step": round(sum(x[3] for x in steps) / len(steps), 3),
           "dp_buckets": buckets, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def dp_bucket_report(ctx, run):
    """Run one step with per-launch DP timing (halign_set_bucket_timing) and
    report each DP launch: shape, jobs, cells, kernel ms alone on its stream,
    GCUPS and the fraction of the 78.6 T packed-16-bit VALU peak at 24 int
    ops per cell. Buckets on different streams overlap, so their times are
    each one's own span, not shares of the step."""
    import halign
    names = halign.shape_names()
    ctx.set_bucket_timing(True)
    try:
        run()
        rows = ctx.last_buckets()
    finally:
        ctx.set_bucket_timing(False)
    out = []
    for b in rows:
        fam = ("planned", "small segments", "multi-wave", "combined")[b["family"]]
        name = (names[b["shape"]] if b["family"] < 2 else f"multiwave{b['shape']}" if b["family"] == 2
                else "q-parts")
        g = b["cells"] / (b["kernel_ms"] * 1e6) if b["kernel_ms"] > 0 else 0.0
        out.append({"kind": fam, "shape": name, "right": b["right"], "stream": b["stream"], "jobs": b["jobs"],
                    "cells": b["cells"], "ms": round(b["kernel_ms"], 3), "est_ms": round(b["est_ms"], 3),
                    "gcups": round(g, 1), "roofline_frac": round(g * 24 / 78.6e3, 4)})
        log(f"  dp {fam:14s} {name:10s} rt {b['right']} stream {b['stream']}: {b['jobs']:7d} jobs "
            f"{b['cells'] / 1e9:7.2f} Gcells {b['kernel_ms']:8.2f} ms (est {b['est_ms']:7.2f}) {g:7.1f} GCUPS")
    return out


def msa_bench(args):
    """halign4 -c center.fa --keep-length (one center, the copy MSA command),
    file to file: read the query FASTA, align every query on the GPU, build
    the keep-length rows and FASTA records on the device, write the final
    FASTA (halign_msa_keep_length). A step is one whole run over the config's
    query file (config 1: the reference's 1,000-query case; config 3: the
    first --queries of the uneven batch). cpu_baseline: the same run by the
    reference's own functions (oracle/_ref ref_keep_length_msa: readers,
    RefAligner::globalAlign, SAM and FASTA writers, merge), one per thread
    over a time-bounded share of the queries."""
    import tempfile
    import halign
    from halign import synth
    cfg = args.config
    label, per_step, total = E2E[cfg]
    need = min(total, args.queries or (1000 if cfg == 1 else 65536))
    center = synth.center(cfg)
    qb, off = synth.queries(cfg, 0, need)
    d = tempfile.mkdtemp(prefix="halign_msa_")
    cpath = os.path.join(d, "center.fa")
    ipath = os.path.join(d, "queries.fa")
    opath = os.path.join(d, "out.fa")
    with open(cpath, "wb") as f:
        if args.refs <= 1:
            f.write(b">center\n" + bytes(center) + b"\n")
        else:   # variants of the center at 2% substitutions: one width, the copy command aligns them
            rng = np.random.default_rng(0xC0FFEE)
            base = np.frombuffer(bytes(center), np.uint8)
            for r in range(args.refs):
                v = base.copy()
                if r:
                    pos = rng.integers(0, len(v), len(v) // 50)
                    v[pos] = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, len(pos))]
                f.write(b">ref%d\n" % r + v.tobytes() + b"\n")
    qv = memoryview(qb if isinstance(qb, (bytes, bytearray)) else bytes(qb))
    with open(ipath, "wb") as f:
        for i in range(need):
            f.write(b">q%d\n" % i)
            f.write(qv[off[i]:off[i + 1]])
            f.write(b"\n")
    in_bytes = os.path.getsize(ipath)
    log(f"msa config {cfg}: {need} queries, {args.refs} reference(s), input {in_bytes / 1e6:.1f} MB in {d}")
    ctx = halign.Context(0)
    steps = []
    for s in range(args.warmup + args.steps):
        if os.path.exists(opath):   # a fresh output file, as a run normally writes
            os.remove(opath)
        t0 = time.perf_counter()
        st = halign.msa_keep_length(ctx, cpath, ipath, opath, batch=args.msa_batch)
        dt = time.perf_counter() - t0
        if s >= args.warmup:
            steps.append((dt, st))
            log(f"step {s - args.warmup}: {st['records'] - 1} rows, {st['bytes'] / 1e6:.1f} MB written, "
                f"{dt * 1e3:.1f} ms wall ({st['seconds'] * 1e3:.1f} in the call), "
                f"{st['gpu_seconds'] * 1e3:.1f} ms in GPU calls, "
                f"{st['input_wait_seconds'] * 1e3:.1f} ms waiting for input")
    tw = sum(x[0] for x in steps)
    out_bytes = steps[-1][1]["bytes"]
    verified = None
    if not args.no_verify:
        verified = verify_msa(ctx, d, cpath, qv, off, min(need, args.verify or 300), args.refs)
    cpu = None
    if not args.no_cpu:
        from oracle_lib import REF_BIN, Halign4
        if os.path.exists(REF_BIN):
            H = Halign4()
            threads = host_threads()
            # calibrate: one run on 8 queries, then a share per thread for ~10 s
            def run_share(k, a, b):
                wd = os.path.join(d, f"ref{k}")
                os.makedirs(wd, exist_ok=True)
                ip = os.path.join(wd, "q.fa")
                with open(ip, "wb") as f:
                    for i in range(a, b):
                        f.write(b">q%d\n" % i + bytes(qv[off[i]:off[i + 1]]) + b"\n")
                t0 = time.perf_counter()
                import shutil
                shutil.rmtree(os.path.join(wd, "w"), ignore_errors=True)
                rc = H.run(ip, os.path.join(wd, "out.fa"), os.path.join(wd, "w"), center_path=cpath)
                assert rc == 0, rc
                return time.perf_counter() - t0
            t8 = run_share(0, 0, min(8, need))
            per = max(1, min(need, int(10.0 * min(8, need) / max(t8, 1e-6))))
            shares = [((k * per) % need, min((k * per) % need + per, need)) for k in range(threads)]
            t0 = time.perf_counter()
            with ThreadPoolExecutor(threads) as ex:
                list(ex.map(lambda k: run_share(k, *shares[k]), range(threads)))
            dt = time.perf_counter() - t0
            done = sum(b - a for a, b in shares)
            cpu = {"value": round(done / dt, 2), "unit": "seqs/s", "cores": threads, "kind": "reference",
                   "sample": f"{done} config-{cfg} queries in {threads} independent keep-length runs of "
                             f"up to {per} queries (one run per thread), {dt:.1f}s"}
            log(f"cpu baseline (reference): {cpu['value']} seqs/s")
    out = {"metric": "seqs/s end-to-end keep-length MSA, FASTA file to FASTA file (halign4 -c --keep-length)",
           "value": round(need * len(steps) / tw, 2), "unit": "seqs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(tw * 1e3 / len(steps), 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int8",
           "data": f"synthetic (seeded, csrc/synth.c config {cfg}), written to FASTA files",
           "config": {"workload": f"config {cfg}: {label}; keep-length FASTA file to file"
                                  + (f", {args.refs} references (variants at 2% substitutions)" if args.refs > 1 else ""),
                      "queries": need, "references": args.refs, "chunk_queries": args.msa_batch or 25600},
           "output_mb_per_s": round(out_bytes * len(steps) / tw / 1e6, 1), "input_bytes": in_bytes,
           "output_bytes": out_bytes,
           "gpu_s_per_step": round(sum(x[1]["gpu_seconds"] for x in steps) / len(steps), 4),
           "input_wait_s_per_step": round(sum(x[1]["input_wait_seconds"] for x in steps) / len(steps), 4),
           "verified": verified, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    import shutil
    shutil.rmtree(d, ignore_errors=True)


def cpu_timed(fn, n, threads, fixed=0, target_s=12.0):
    """Run fn over items 0, 1, ... (cycling through n) on `threads` threads in
    rounds of 4 x threads items until target_s seconds have passed (or exactly
    `fixed` items). Returns (items done, seconds)."""
    done = 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while True:
            m = (fixed - done) if fixed else 4 * threads
            if m <= 0:
                break
            list(ex.map(lambda k: fn((done + k) % n), range(m)))
            done += m
            if not fixed and time.perf_counter() - t0 >= target_s:
                break
    return done, time.perf_counter() - t0


def serve_bench(args):
    """Config 5 serving: a continuous-batching server (halign_server_*) with the
    29,903 bp center and its index resident, requests of n ~ U{1..4096} queries
    (lengths U{1..30000}, 1% divergence), --serve-queries queries in all
    (2,000,000 by default). The queries of request i are a contiguous range of
    a seeded pool of --pool distinct queries. First all requests at once
    (closed loop: the capacity in seqs/s), then Poisson arrivals at fractions
    of that capacity; latency = submit -> last CIGAR, measured in the
    library. The server's footprint estimate of its largest batch and the
    device memory it was seen to use are reported beside the budget."""
    import threading
    import halign
    from halign import synth
    center = synth.center(5)
    rng = np.random.default_rng(0xC5A)
    sizes = []
    total = 0
    while total < args.serve_queries and (args.requests == 0 or len(sizes) < args.requests):
        n = int(rng.integers(1, 4097))
        sizes.append(n)
        total += n
    sizes = np.array(sizes, np.int64)
    R = len(sizes)
    need = int(sizes.sum())
    P = max(8192, min(args.pool, need))
    qb, off = synth.queries(5, 0, P)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]) % (P - 4096)   # request i: pool[starts[i] : +sizes[i]]
    bases = int(sum(off[s + n] - off[s] for s, n in zip(starts, sizes)))
    log(f"serve: {R} requests, {need} queries ({P} distinct), {bases / 1e9:.2f} Gbases per pass")
    srv = halign.Server(center, workers=args.workers, max_batch_queries=args.batch_queries)

    kept = []

    def run(arrivals, label, R=R):
        kept.clear()
        ids = [None] * R
        lat = np.zeros(R)
        done_t = np.zeros(R)
        t0 = time.perf_counter()

        def collect():
            for i in range(R):
                while ids[i] is None:
                    time.sleep(0.0005)
                cig, coff, _, timing = srv.result(ids[i])
                lat[i] = timing[1]
                done_t[i] = time.perf_counter() - t0
                if i % 16 == 0:   # one query of every sixteenth request is checked after the run
                    j = int(sizes[i]) // 2
                    kept.append((int(starts[i]) + j, cig[coff[j]:coff[j + 1]].copy()))
        th = threading.Thread(target=collect)
        th.start()
        for i in range(R):
            if arrivals is not None:
                dt = arrivals[i] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
            ids[i] = srv.submit_flat(qb, off[starts[i]:starts[i] + sizes[i] + 1])
        th.join()
        span = float(done_t.max())
        nq = int(sizes[:R].sum())
        log(f"{label}: {nq / span:.0f} seqs/s, p50 {np.percentile(lat, 50):.1f} ms, "
            f"p99 {np.percentile(lat, 99):.1f} ms, span {span:.2f} s")
        return lat, span
    try:
        run(None, "warmup", R=min(R, 64))
        _, span = run(None, "capacity (all at once)")
        cap = need / span
        points = []
        mean_n = float(sizes.mean())
        for f in args.loads:
            lam = f * cap / mean_n   # requests per second
            arr = np.cumsum(np.random.default_rng(0xC5A + int(f * 100)).exponential(1.0 / lam, R))
            lat, sp = run(arr, f"load {f:.2f}")
            points.append({"load": f, "offered_req_per_s": round(lam, 3), "seqs_per_s": round(need / sp, 1),
                           "p50_ms": round(float(np.percentile(lat, 50)), 1),
                           "p99_ms": round(float(np.percentile(lat, 99)), 1)})
        memory = srv.stats()
        memory["device_free_after"], _ = hip_mem_info()
        log(f"serve memory: {memory['batches']} batches; resident {memory['resident'] / 2**30:.2f} GiB, "
            f"traceback reserved {memory['traceback_reserved'] / 2**30:.2f} GiB; working buffers: estimate "
            f"{memory['sum_worker_estimates'] / 2**30:.2f} GiB (largest batch per worker, summed), measured "
            f"{memory['working_measured'] / 2**30:.2f} GiB")
    finally:
        srv.close()
    verified = None if args.no_verify else verify_queries(
        center, qb, off, list(kept), "served queries (the middle one of every sixteenth request at the last load)")
    cpu = None
    if not args.no_cpu:
        from oracle_lib import REF_SO, Oracle, RefLib
        threads = host_threads()
        kind = "reference" if os.path.exists(REF_SO) else "port"
        lib = RefLib() if kind == "reference" else Oracle()
        if kind == "reference":
            hits = lib.minimizers(center)
            fn = lambda i: lib.ref_global_align_pre(center, hits, qb[off[i]:off[i + 1]])
        else:
            fn = lambda i: lib.ref_global_align(center, qb[off[i]:off[i + 1]])
        sample, dt = cpu_timed(fn, need, threads, args.cpu_sample)
        cpu = {"value": round(sample / dt, 2), "unit": "seqs/s", "cores": threads, "kind": kind,
               "sample": f"{sample} config-5 queries (RefAligner::globalAlign, in request order), "
                         f"{dt:.1f}s on {threads} threads"}
        log(f"cpu baseline ({kind}): {cpu['value']} seqs/s")
    hi = points[-1] if points else {}
    out = {"metric": "serving seqs/s and p50/p99 request latency (continuous batching)",
           "value": round(cap, 1), "unit": "seqs/s", "n_gpus": 1, "steps": len(args.loads) + 1, "warmup": 1,
           "ms_per_step": round(span * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "int8", "data": "synthetic (seeded, csrc/synth.c config 5)",
           "config": {"workload": "config 5: serving, center 29903 bp resident, requests of U{1..4096} queries "
                                  "of U{1..30000} bp at 1% divergence", "requests": R, "queries": need,
                      "distinct_queries": P, "workers": args.workers, "max_batch_queries": args.batch_queries},
           "p50_ms": hi.get("p50_ms"), "p99_ms": hi.get("p99_ms"), "load_points": points, "memory": memory,
           "verified": verified,
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def hip_mem_info():
    """(free, total) device bytes from the HIP runtime libhalign_hip.so links."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    f, t = C.c_size_t(0), C.c_size_t(0)
    if hip.hipMemGetInfo(C.byref(f), C.byref(t)) != 0:
        return None, None
    return f.value, t.value


class DeviceBuffer:
    """Device memory from the HIP runtime libhalign_hip.so links (loaded with
    it): plain hipMalloc / hipMemcpy through ctypes."""

    def __init__(self, nbytes):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        if self.hip.hipMalloc(C.byref(self.ptr), nbytes) != 0:
            raise MemoryError(f"hipMalloc({nbytes})")

    def upload(self, host, offset=0):
        host = np.ascontiguousarray(host)
        assert self.hip.hipMemcpy(self.ptr.value + offset, host.ctypes.data_as(self.C.c_void_p), host.nbytes, 1) == 0

    def copy_from(self, other):
        assert self.hip.hipMemcpy(self.ptr, other.ptr, min(self.nbytes, other.nbytes), 3) == 0   # device to device

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = self.C.c_void_p()


def consensus_bench(args):
    """Column counting + consensus over rows x 29,903 aligned columns already
    resident in HBM (stride rounded to 4 bytes). Algorithmic bytes: rows x
    columns read + 28 x columns written (7 uint32 counts)."""
    import halign
    aln = 29903
    stride = (aln + 3) & ~3
    alpha = np.frombuffer(b"ACGT-N", np.uint8)
    halign.lib()
    rows = DeviceBuffer(args.rows * stride)
    rng = np.random.default_rng(0xC5)
    for r0 in range(0, args.rows, 8192):
        r1 = min(args.rows, r0 + 8192)
        rows.upload(alpha[rng.integers(0, len(alpha), (r1 - r0, stride), dtype=np.uint8)], r0 * stride)
    ctx = halign.Context(0)
    ms_all = []
    for s in range(args.warmup + args.steps):
        counts, cons, n, ms = ctx.consensus((args.rows, aln), device_ptr=rows.ptr.value, row_stride=stride)
        if s >= args.warmup:
            ms_all.append(ms)
            log(f"step {s - args.warmup}: {args.rows} rows, kernel {ms:.3f} ms")
    assert int(counts.sum()) == n * aln
    t = sum(ms_all) / len(ms_all) * 1e-3
    algo = args.rows * aln + 28 * aln
    out = {"metric": "consensus column counting GB/s (rows in HBM)", "value": round(algo / t / 1e9, 1),
           "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(t * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u8", "data": "synthetic (numpy seeded, ACGT-N)",
           "config": {"workload": f"consensus counting, {args.rows} rows x {aln} columns"},
           "roofline": {"bound": "hbm", "achieved": round(algo / t / 1e9, 1), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(algo / t / 8e12, 4), "traffic": None}}
    print(json.dumps(out), flush=True)


def traffic_per_launch(cells_per_launch):
    """HBM bytes of one DP launch (one step): bytes per cell from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), times the cells
    of a step. None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "dp_ring_config2_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_cell:"):
            return int(float(line.split(":")[1]) * cells_per_launch)
    return None


def halign_auto_band(rlen, qlen):
    d = abs(rlen - qlen)
    if d / max(rlen, qlen) > 0.5:
        return -1
    return 200 + int(0.1 * (rlen + qlen // 2))


if __name__ == "__main__":
    main()
