"""The drop-in boundary through the C ABI, as a reference-side shim binds it.

- halign_ksw_extz2 with ksw_extz2_sse's own argument list (include/ksw2.h:
  64-65, called from src/align/align.cpp:49-55): one ksw_extz_t reused
  across all 3,000 reference-written cases (ksw_cases.npz), every field and
  the CIGAR compared, and m_cigar checked against ksw_push_cigar's growth
  (include/ksw2.h:113-123: 4 when empty, then doubling); the early returns
  for m <= 0, qlen = 0 or tlen = 0 (src/align/ksw2_extz2_sse.c:56-57) with a
  reset ez (ksw_reset_extz, include/ksw2.h:184-189) and the cigar buffer
  kept; every configuration outside the path refused with
  HALIGN_E_UNSUPPORTED.
- Lifetimes: centers and pair sets freed before or after their context.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NEG_INF = -0x40000000


@pytest.fixture(scope="module")
def ctx():
    import halign
    c = halign.Context(0)
    yield c
    c.close()


def _reset_fields(ez):
    return (ez.max_q, ez.max_t, ez.mqe_t, ez.mte_q, ez.max, ez.score, ez.mqe, ez.mte, ez.n_cigar, ez.zdropped,
            ez.reach_end) == (-1, -1, -1, -1, 0, NEG_INF, NEG_INF, NEG_INF, 0, 0, 0)


def test_ksw_extz2_drop_in_reused_ez(ctx):
    import halign
    with np.load(os.path.join(GOLD, "ksw_cases.npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    fields = [str(f) for f in g["ez_fields"]]
    n = len(g["w"])
    print(f"ksw_extz2 drop-in: {n} cases on one reused ez", file=sys.stderr, flush=True)
    ez = halign.ExtzT()
    bad = []
    try:
        for i in range(n):
            q = g["q"][g["q_off"][i]:g["q_off"][i + 1]]
            t = g["t"][g["t_off"][i]:g["t_off"][i + 1]]
            m_before = ez.m_cigar
            rc = ctx.ksw_extz2(q, t, int(g["w"][i]), int(g["flag"][i]), ez)
            assert rc == 0, (i, rc, ctx.lib.halign_last_error(ctx.h))
            exp = dict(zip(fields, (int(x) for x in g["ez"][i])))
            got = {f: int(getattr(ez, f)) for f in fields}
            cig = ez.cigar_array()
            exp_cig = g["cigar"][g["cigar_off"][i]:g["cigar_off"][i + 1]]
            m = m_before
            while m < ez.n_cigar:
                m = m << 1 if m else 4
            if got != exp or not np.array_equal(cig, exp_cig) or ez.m_cigar != m or ez.reach_end != 0:
                bad.append(i)
    finally:
        halign.free_extz(ez)
    assert not bad, f"{len(bad)} cases differ, first {bad[:5]}"


def test_ksw_extz2_empty_and_unsupported(ctx):
    import halign
    ez = halign.ExtzT()
    one = np.array([0, 1, 2, 3], np.uint8)
    try:
        assert ctx.ksw_extz2(one, one, -1, 0x06, ez) == 0
        cig_ptr = C.cast(ez.cigar, C.c_void_p).value
        m = ez.m_cigar
        assert ez.n_cigar == 1 and ez.cigar[0] == (4 << 4) and ez.score == 20
        # empty query, empty target, m <= 0: the reset ez, the cigar buffer kept
        for q, t, kw in ((np.zeros(0, np.uint8), one, {}), (one, np.zeros(0, np.uint8), {}), (one, one, {"m": 0})):
            assert ctx.ksw_extz2(q, t, -1, 0x06, ez, **kw) == 0
            assert _reset_fields(ez)
            assert C.cast(ez.cigar, C.c_void_p).value == cig_ptr and ez.m_cigar == m
        # outside the path's configuration: refused, never a wrong answer
        other = (C.c_int8 * 25)(*([2] * 25))
        for kw in ({"m": 4}, {"mat": other}, {"q": 5}, {"e": 1}, {"zdrop": 100}, {"end_bonus": 1}):
            assert ctx.ksw_extz2(one, one, -1, 0x06, ez, **kw) == -2, kw
        for flag in (0x02, 0x00, 0x06 | 0x40, 0x06 | 0x01, 0x06 | 0x80):
            assert ctx.ksw_extz2(one, one, -1, flag, ez) == -2, flag
    finally:
        halign.free_extz(ez)


def _make(lib, center):
    ctxh = C.c_void_p()
    assert lib.halign_init(0, C.byref(ctxh)) == 0
    ch = C.c_void_p()
    assert lib.halign_center_create(ctxh, center, len(center), C.byref(ch)) == 0
    ps = C.c_void_p()
    off = np.array([0, len(center)], np.int64)
    assert lib.halign_pairset_create(ctxh, center, len(center), 1, center, off.ctypes.data_as(C.c_void_p),
                                     C.byref(ps)) == 0
    return ctxh, ch, ps


def _align_one(lib, ch, q):
    off = np.array([0, len(q)], np.int64)
    cig = np.zeros(len(q) + 64, np.uint32)
    coff = np.zeros(2, np.int64)
    rc = lib.halign_center_align(ch, 1, q, off.ctypes.data_as(C.c_void_p), cig.ctypes.data_as(C.c_void_p), len(cig),
                                 coff.ctypes.data_as(C.c_void_p), None, None)
    return rc, cig[:coff[1]]


def test_children_freed_before_or_after_the_context():
    """The raw C ABI: a context freed first stays alive for its center and
    pair set, which still run; the last child's free releases it. The other
    order is the plain one. Both end without a fault."""
    import halign
    lib = halign.lib()
    rng = np.random.default_rng(7)
    center = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 3000)].tobytes()
    q = center[200:2500]
    # context first
    ctxh, ch, ps = _make(lib, center)
    lib.halign_free(ctxh)
    rc, cig = _align_one(lib, ch, q)
    assert rc == 0 and len(cig) > 0
    ms, cells = C.c_double(0), C.c_int64(0)
    assert lib.halign_pairset_run(ps, 0, 1, 0, C.byref(ms), C.byref(cells)) == 0
    lib.halign_center_free(ch)
    assert lib.halign_pairset_run(ps, 0, 1, 0, C.byref(ms), C.byref(cells)) == 0
    lib.halign_pairset_free(ps)
    # children first
    ctxh, ch, ps = _make(lib, center)
    rc2, cig2 = _align_one(lib, ch, q)
    assert rc2 == 0 and np.array_equal(cig, cig2)
    lib.halign_pairset_free(ps)
    lib.halign_center_free(ch)
    lib.halign_free(ctxh)
