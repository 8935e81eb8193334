"""HIP-graph replay of the batched DP launches: the same pair range run with
direct launches, then captured, then replayed several times with other work
(different DP batches, host synchronisation) in between; every replay gives
the same scores and CIGARs as the direct run, and the context reports the
replays."""
import sys

import numpy as np
import pytest

from seqgen import codes_to_ascii, mutate_codes, rand_codes

pytestmark = pytest.mark.gpu


def _results(ps, n):
    out = []
    for i in range(n):
        score, cig = ps.result(i)
        out.append((score, np.asarray(cig).tobytes()))
    return out


def test_graph_replays_match_direct_launches():
    import halign
    rng = np.random.default_rng(5)
    center = rand_codes(rng, 3000)
    qs = [mutate_codes(rng, center[int(rng.integers(0, 200)):], 0.02, 0.005) for _ in range(96)]
    blob = b"".join(bytes(codes_to_ascii(q)) for q in qs)
    off = np.zeros(len(qs) + 1, np.int64)
    off[1:] = np.cumsum([len(q) for q in qs])
    ctx = halign.Context(0)
    try:
        ps = halign.PairSet(ctx, codes_to_ascii(center), blob, off)
        ctx.set_graph_mode(False)
        ps.run(0, len(qs))
        direct = _results(ps, len(qs))
        ctx.set_graph_mode(True)
        c0, r0 = ctx.graph_counts()
        other = [mutate_codes(rng, rand_codes(rng, 500), 0.05, 0.01) for _ in range(8)]
        for rep in range(4):
            ps.run(0, len(qs))
            got = _results(ps, len(qs))
            bad = [i for i, (a, b) in enumerate(zip(direct, got)) if a != b]
            assert not bad, f"replay {rep}: pairs {bad[:5]} differ"
            # other eager work between the replays
            ctx.global_align_batch([codes_to_ascii(o) for o in other], [codes_to_ascii(o[::-1]) for o in other],
                                   [halign.FLAG_RIGHT] * len(other))
        c1, r1 = ctx.graph_counts()
        print(f"graph: {c1 - c0} captures, {r1 - r0} replays", file=sys.stderr, flush=True)
        assert r1 - r0 >= 3
        ps.close()
    finally:
        ctx.close()
