"""halign_consensus_noalloc (no counterpart in the reference: the
no-allocation form for callers that cannot allocate) against
halign_consensus on the same device rows, through the C ABI: two calls in a
row, the second with fewer rows and columns (the first call's counts must
not leak into it), and the column limit of its static scratch."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceRows:
    """Rows in device memory from the HIP runtime the library itself links
    (libamdhip64.so.7, already loaded with it), so no second runtime enters
    the process."""

    def __init__(self, host):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), host.nbytes) == 0
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), host.nbytes, 1) == 0   # host to device

    def data_ptr(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _device_rows(rng, nrows, aln):
    stride = (aln + 3) & ~3
    alpha = np.frombuffer(b"ACGTUN-.acgtRY", np.uint8)
    host = np.ascontiguousarray(alpha[rng.integers(0, len(alpha), (nrows, stride))])
    return DeviceRows(host), stride


def _noalloc(lib, ctx, rows, nrows, aln, stride, batch, lim):
    counts = np.zeros((aln, 7), np.uint32)
    cons = np.zeros(aln, np.uint8)
    n = C.c_uint64(0)
    rc = lib.halign_consensus_noalloc(ctx.h, rows.data_ptr(), nrows, aln, stride, batch, lim,
                                      counts.ctypes.data_as(C.c_void_p), cons.ctypes.data_as(C.c_void_p),
                                      C.byref(n))
    return rc, counts, cons.tobytes(), n.value


def test_noalloc_matches_consensus_twice():
    import halign
    lib = halign.lib()
    ctx = halign.Context(0)
    rng = np.random.default_rng(77)
    held = []
    try:
        for nrows, aln, batch, lim in ((9000, 5001, 4096, 0), (700, 3003, 64, 500)):
            print(f"no-alloc consensus: {nrows} rows x {aln} columns, batch {batch}, limit {lim}",
                  file=sys.stderr, flush=True)
            rows, stride = _device_rows(rng, nrows, aln)
            held.append(rows)
            exp_counts, exp_cons, exp_n, _ = ctx.consensus((nrows, aln), batch=batch, seq_limit=lim,
                                                           device_ptr=rows.data_ptr(), row_stride=stride)
            rc, counts, cons, n = _noalloc(lib, ctx, rows, nrows, aln, stride, batch, lim)
            assert rc == 0, lib.halign_last_error(ctx.h)
            assert np.array_equal(counts, exp_counts) and cons == exp_cons and n == exp_n
            assert int(counts.sum()) == n * aln
        # beyond the static scratch: refused before any device work
        counts = np.zeros(1, np.uint32)
        rc = lib.halign_consensus_noalloc(ctx.h, rows.data_ptr(), 1, (1 << 20) + 1, (1 << 20) + 4, 4096, 0,
                                          counts.ctypes.data_as(C.c_void_p), counts.ctypes.data_as(C.c_void_p), None)
        assert rc == -5
    finally:
        ctx.close()
        for r in held:
            r.free()
