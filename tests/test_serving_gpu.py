"""Continuous-batching server (halign_server_*): every request's CIGARs are
the same as those of the direct batch call (halign_center_align, itself pinned
to the reference fixtures and the oracle), whether the request runs alone or
mixed with others in shared batches on several workers."""
import sys
import threading

import numpy as np
import pytest

from seqgen import codes_to_ascii, mutate_codes, rand_codes

pytestmark = pytest.mark.gpu


def _requests(rng, center, nreq):
    reqs = []
    for r in range(nreq):
        qs = []
        for _ in range(int(rng.integers(1, 40))):
            a = int(rng.integers(0, len(center) - 50))
            L = int(rng.integers(1, len(center) - a))
            q = mutate_codes(rng, center[a:a + L], 0.01, 0.002)
            if rng.random() < 0.1:
                q = rand_codes(rng, int(rng.integers(0, 300)))   # no anchors: whole-pair DP
            qs.append(codes_to_ascii(q))
        reqs.append(qs)
    return reqs


def _split(cig, coff):
    return [cig[coff[i]:coff[i + 1]] for i in range(len(coff) - 1)]


def test_alone_vs_mixed():
    import halign
    rng = np.random.default_rng(9)
    center = rand_codes(rng, 6000)
    cstr = codes_to_ascii(center)
    reqs = _requests(rng, center, 24)
    print(f"serving: {len(reqs)} requests, {sum(len(r) for r in reqs)} queries", file=sys.stderr, flush=True)
    # direct batch calls: the expected CIGARs
    ctx = halign.Context(0)
    c = halign.Center(ctx, cstr)
    expect = []
    for qs in reqs:
        cigs, _, _ = c.align(qs)
        expect.append(cigs)
    ctx.close()

    srv = halign.Server(cstr, workers=2, max_batch_queries=64)
    try:
        # alone: one request at a time
        for i in (0, 5):
            rid = srv.submit(reqs[i])
            cig, coff, st, timing = srv.result(rid)
            assert timing[2] == len(reqs[i])
            got = _split(cig, coff)
            assert all(np.array_equal(a, b) for a, b in zip(got, expect[i])), f"request {i} alone"
        # mixed: all requests at once from two submitting threads, batches shared
        ids = [None] * len(reqs)

        def sub(lo, hi):
            for i in range(lo, hi):
                ids[i] = srv.submit(reqs[i])
        th = [threading.Thread(target=sub, args=(0, 12)), threading.Thread(target=sub, args=(12, 24))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        shared = 0
        for i, rid in enumerate(ids):
            cig, coff, st, timing = srv.result(rid)
            shared += timing[2] > len(reqs[i])
            got = _split(cig, coff)
            assert len(got) == len(expect[i])
            bad = [k for k, (a, b) in enumerate(zip(got, expect[i])) if not np.array_equal(a, b)]
            assert not bad, f"request {i} mixed: queries {bad[:5]} differ"
        assert shared > 0, "no request shared a batch"
    finally:
        srv.close()


def test_empty_request_and_errors():
    import halign
    rng = np.random.default_rng(3)
    cstr = codes_to_ascii(rand_codes(rng, 2000))
    srv = halign.Server(cstr, workers=1)
    try:
        rid = srv.submit([])
        cig, coff, st, timing = srv.result(rid)
        assert len(cig) == 0 and list(coff) == [0]
        with pytest.raises(halign.HalignError):
            srv.wait(12345)
    finally:
        srv.close()
