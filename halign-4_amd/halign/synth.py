"""Seeded synthetic workloads of the five benchmark configurations
(csrc/synth.c): center of configuration c and its queries i."""
import ctypes as C
import os

import numpy as np

from ._lib import PKG_DIR

_SO = os.path.join(PKG_DIR, "lib", "libhalign_synth.so")
_lib = None

CONFIG_COUNTS = {1: 1000, 2: 100000, 3: 1000000, 4: 200000, 5: 2000000}


def _l():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            raise ImportError(f"{_SO} is not built; run __graft_entry__.build() first")
        _lib = C.CDLL(_SO)
        _lib.synth_center.argtypes = [C.c_int, C.c_char_p, C.c_int64]
        _lib.synth_center.restype = C.c_int64
        _lib.synth_center_len.argtypes = [C.c_int]
        _lib.synth_center_len.restype = C.c_int64
        _lib.synth_queries.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_void_p, C.c_int64,
                                       C.c_void_p]
        _lib.synth_queries.restype = C.c_int64
    return _lib


def center(cfg):
    n = _l().synth_center_len(cfg)
    buf = C.create_string_buffer(int(n))
    _l().synth_center(cfg, buf, n)
    return buf.raw[:n]


def queries(cfg, first, count):
    """Returns (bytes of concatenated queries, int64 offsets[count+1])."""
    off = np.zeros(count + 1, np.int64)
    need = _l().synth_queries(cfg, first, count, None, 0, off.ctypes.data_as(C.c_void_p))
    if need == 0 and count > 0:
        raise RuntimeError("synth_queries failed")
    cap = -need if need < 0 else need
    buf = np.zeros(max(1, cap), np.uint8)
    got = _l().synth_queries(cfg, first, count, buf.ctypes.data_as(C.c_void_p), cap,
                             off.ctypes.data_as(C.c_void_p))
    assert got == cap, (got, cap)
    return buf[:cap].tobytes(), off
