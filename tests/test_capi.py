"""The C-ABI library on the CPU: it loads, exports every entry point that
include/*.h declares, and its host-only calls answer without a GPU. No
compute runs here (that is tests/test_dp_gpu.py, marked gpu)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(halign_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    from halign import _lib
    return _lib.lib()


def test_header_declares_the_path_entry_points():
    names = declared_functions()
    for must in ("halign_init", "halign_free", "halign_last_error", "halign_ksw_extz2",
                 "halign_dp_batch", "halign_global_align_batch", "halign_pairset_create",
                 "halign_pairset_run", "halign_pairset_result", "halign_pairset_free"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_names_the_target(lib):
    assert b"gfx950" in lib.halign_version()


def test_ring_shapes(lib):
    import halign
    shapes = halign.ring_shapes()
    assert len(shapes) >= 8
    for waves, lanes, ring in shapes:
        assert 0 <= waves <= 16 and ring % lanes == 0
        assert lanes == 64 * waves if waves else lanes in (4, 16)
    assert {g for w, g, _ in shapes if w == 0} == {4, 16}
    # the config-2 band (n_col_ = 293 blocks) has a shape that holds it
    assert any(ring - 2 >= 293 for _, _, ring in shapes)
    assert halign.ring_policy(5, 1) >= 2
    assert halign.ring_policy(0, 2) >= 2 and halign.ring_policy(0, 1, lanes=4) >= 2


def test_null_context_is_an_argument_error(lib):
    lib.halign_set_kernel_policy.argtypes = [C.c_void_p, C.c_int]
    assert lib.halign_set_kernel_policy(None, 0) == -1
    lib.halign_last_error.restype = C.c_char_p
    lib.halign_last_error.argtypes = [C.c_void_p]
    assert lib.halign_last_error(None) == b"no context"
