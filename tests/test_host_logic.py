"""Host-side logic of the C-ABI library on the CPU (no GPU compute): the band
geometry that decides the kernel mode (CIGAR-only vs full statistics), the
ring shape and the traceback size, against a direct restatement of
src/align/ksw2_extz2_sse.c:75-76,107-114 and against the oracle's cell count."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import FLAG_RIGHT, Oracle
from seqgen import rand_codes


@pytest.fixture(scope="module")
def lib():
    from halign import _lib
    L = _lib.lib()
    L.halign_band_info.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
    return L


def band_info(L, q, t, w):
    nc, ov, cells = C.c_int32(), C.c_int32(), C.c_int64()
    assert L.halign_band_info(q, t, w, C.byref(nc), C.byref(ov), C.byref(cells)) == 0
    return nc.value, ov.value, cells.value


def direct(q, t, w):
    if w < 0:
        w = max(q, t)
    nc = (min(q, t, w + 1) + 15) // 16 + 1
    cells = 0
    for r in range(q + t - 1):
        st = max(0, r - q + 1, (r - w + 1) >> 1)
        en = min(t - 1, r, (r + w) >> 1)
        if st > en:
            return nc, 1, cells
        cells += en - st + 1
    return nc, 0, cells


def test_band_info_exhaustive_small(lib):
    for q in range(1, 34):
        for t in range(1, 34):
            for w in range(-1, 40):
                assert band_info(lib, q, t, w) == direct(q, t, w), (q, t, w)


def test_band_info_matches_oracle_cells(lib):
    O = Oracle()
    rng = np.random.default_rng(8)
    for it in range(200):
        q = rand_codes(rng, int(rng.integers(1, 300)))
        t = rand_codes(rng, int(rng.integers(1, 300)))
        w = int(rng.integers(-1, 60))
        ez, _, cells = O.ksw(q, t, w, FLAG_RIGHT)
        nc, ov, c = band_info(lib, len(q), len(t), w)
        assert ov == ez[1] and c == cells, (len(q), len(t), w)


def test_band_info_config2_pair(lib):
    """A whole config-2 pair: auto band 4,665 -> n_col_ = 293 blocks, ~2.55e8 cells."""
    O = Oracle()
    w = O.auto_band(29903, 29500)
    nc, ov, cells = band_info(lib, 29500, 29903, w)
    assert (w, nc, ov, cells) == (4665, 293, 0, 255296602)


def test_band_info_rejects_empty(lib):
    nc, ov, cells = C.c_int32(), C.c_int32(), C.c_int64()
    assert lib.halign_band_info(0, 5, 3, C.byref(nc), C.byref(ov), C.byref(cells)) == -1
