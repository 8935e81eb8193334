"""The traceback pool is planned within what the device can still give.

A context's default budget is the HBM free at halign_init less a reserve, so
a context created before other allocations (another context's pool, a
framework's tensors) would plan past the memory that is left. The batch
runner therefore plans each call within min(budget, its pool + free HBM - 1
GiB): under that pressure the launches shrink (fewer workgroups, more rounds
of tickets) and the results stay those of an unconstrained run.

The memory held elsewhere comes from the HIP runtime the library links
(libamdhip64.so.7, already loaded with it), not from a framework: a second
runtime of the same soname loaded after the library's would bind to it."""
import ctypes as C
import sys

import numpy as np
import pytest

from seqgen import mutate_codes, rand_codes
from test_dp_gpu import _tasks

pytestmark = pytest.mark.gpu


def _jobs(n, tlen, seed):
    """n pairs of a tlen-base target and a 1%-mutated copy, band 4,665 (n_col_ 293:
    the one-wave ring of five slots, in either form)."""
    rng = np.random.default_rng(seed)
    base = rand_codes(rng, tlen + 64)
    cases = []
    for i in range(n):
        s = int(rng.integers(0, 64))
        t = base[s:s + tlen].copy()
        q = mutate_codes(rng, t, 0.008, 0.002)
        cases.append((q, t, 4665, 0x06))
    return cases


class Held:
    """Device memory held through the library's own HIP runtime."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()

    def free_bytes(self):
        free, total = C.c_size_t(0), C.c_size_t(0)
        assert self.hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
        return free.value

    def hold(self, nbytes):
        assert self.hip.hipMalloc(C.byref(self.ptr), nbytes) == 0, nbytes

    def release(self):
        if self.ptr:
            assert self.hip.hipFree(self.ptr) == 0
            self.ptr = C.c_void_p()


def test_pool_shrinks_under_memory_pressure():
    import halign
    cases = _jobs(2048, 8000, 99)
    tasks, qc, tc = _tasks(cases)
    ref = halign.Context(0)
    try:
        print("2,048 pairs of 8 kb, unconstrained", file=sys.stderr, flush=True)
        r0, c0, _ = ref.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
    finally:
        ref.close()
    ctx = halign.Context(0)   # budget taken now, with nearly all HBM free
    hog = Held()
    try:
        free = hog.free_bytes()
        left = 24 << 30
        hog.hold(free - left)
        print(f"holding {(free - left) / 2**30:.0f} GiB elsewhere, {left >> 30} GiB left", file=sys.stderr, flush=True)
        ctx.set_bucket_timing(True)
        r1, c1, _ = ctx.dp_batch(tasks, qc, tc, full_stats=False, plain=False)
        rows = ctx.last_buckets()
        ctx.set_bucket_timing(False)
    finally:
        ctx.close()
        hog.release()
    assert rows and all(b["jobs"] == 2048 for b in rows if b["family"] == 0), rows
    assert np.array_equal(r1["score"], r0["score"]) and np.array_equal(r1["n_cigar"], r0["n_cigar"])
    assert all(np.array_equal(a, b) for a, b in zip(c1, c0))
    assert (r1["status"] == 0).all()
