"""GPU parity of minimizer extraction (minimizer::extractMinimizer,
src/seed/minimizer.cpp:103-177) against the reference-written fixtures and the
oracle, bit-exact, for both device forms (the wave form and the exact
sequential form that takes over when the reference's ring queue overwrites
its front slot)."""
import os
import sys

import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu
ACGT = np.frombuffer(b"ACGT", np.uint8)


@pytest.fixture(scope="module", params=[0, 1], ids=["auto", "sequential"])
def ctx(request):
    import halign
    c = halign.Context(0)
    c.set_seed_policy(request.param)
    print(f"seed policy {request.param}", file=sys.stderr, flush=True)
    yield c
    c.close()


def ctx_policy(ctx):
    return getattr(ctx, "_seed_policy", None)


def test_minimizers_golden(ctx):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "minimizers.npz")
    with np.load(path, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    n = len(g["k"])
    # group by (k, w, non_canonical) so each group is one batched call
    groups = {}
    for i in range(n):
        groups.setdefault((int(g["k"][i]), int(g["w"][i]), int(g["non_canonical"][i])), []).append(i)
    print(f"minimizer golden: {n} sequences in {len(groups)} (k, w, canonical) groups", file=sys.stderr,
          flush=True)
    seq_form = 0
    for (k, w, nc), idx in groups.items():
        seqs = [g["seq"][g["seq_off"][i]:g["seq_off"][i + 1]].tobytes() for i in idx]
        hits, st = ctx.extract_minimizers(seqs, k, w, bool(nc))
        seq_form += int((st & 1).sum())
        for j, i in enumerate(idx):
            ex = g["x"][g["hit_off"][i]:g["hit_off"][i + 1]]
            ey = g["y"][g["hit_off"][i]:g["hit_off"][i + 1]]
            if not (np.array_equal(hits[j][0], ex) and np.array_equal(hits[j][1], ey)):
                gx, gy = hits[j]
                d = next((t for t in range(min(len(gx), len(ex))) if gx[t] != ex[t] or gy[t] != ey[t]),
                         min(len(gx), len(ex)))
                pytest.fail(f"case {i} k={k} w={w} nc={nc} status={st[j]}: {len(gx)} hits vs {len(ex)}, first "
                            f"difference at {d}: gpu {[(hex(int(a)), int(b) & 0xffffffff) for a, b in zip(gx[d:d + 3], gy[d:d + 3])]} "
                            f"expected {[(hex(int(a)), int(b) & 0xffffffff) for a, b in zip(ex[d:d + 3], ey[d:d + 3])]}")
            assert st[j] & 6 == 0
    assert seq_form > 0


def test_minimizers_random_vs_oracle(ctx):
    """Path parameters (k 15, w 10, non-canonical) on sequences of 0-40 kb with
    N runs, lowercase and U, plus increasing-hash stretches that make the
    reference's queue overwrite its front slot."""
    O = Oracle()
    rng = np.random.default_rng(4242)
    seqs = []
    for i in range(400):
        n = int(rng.integers(0, 3000)) if i % 10 else int(rng.integers(20000, 40000))
        s = ACGT[rng.integers(0, 4, n)].copy()
        if i % 3 == 0 and n:
            s[rng.random(n) < 0.002] = ord("N")
        if i % 5 == 0 and n > 100:
            a = int(rng.integers(0, n - 50))
            s[a:a + int(rng.integers(1, 40))] = ord("N")
        b = s.tobytes()
        if i % 7 == 0:
            b = b.lower()
        if i % 11 == 0:
            b = b.replace(b"T", b"U")
        seqs.append(b)
    print("minimizers random: 400 sequences", file=sys.stderr, flush=True)
    for w in (10, 1, 5, 32, 40):
        hits, st = ctx.extract_minimizers(seqs, 15, w, True)
        for j, s in enumerate(seqs):
            x, y, ub = O.minimizers(s, 15, w, True)
            if ub & 3:
                continue
            if not (np.array_equal(hits[j][0], x) and np.array_equal(hits[j][1], y)):
                gx, gy = hits[j]
                d = next((t for t in range(min(len(gx), len(x))) if gx[t] != x[t] or gy[t] != y[t]),
                         min(len(gx), len(x)))
                pytest.fail(f"seq {j} len {len(s)} w={w} status={st[j]}: {len(gx)} hits vs {len(x)}, first "
                            f"difference at {d}")
        if w == 10 and ctx_policy(ctx) == 0:
            assert int((st & 1).sum()) < len(seqs)   # the wave form carries most sequences
