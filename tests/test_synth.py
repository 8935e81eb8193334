"""The seeded workloads of the five benchmark configurations are the same
bytes on every machine: SHA-256 prefixes of each center and of the first and
last 64 queries (bytes followed by the little-endian int64 offsets). The
golden fixtures of the configuration tests store query indices only, so they
rely on these."""
import hashlib
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "halign-4_amd"))
from halign import synth  # noqa: E402

CENTERS = {
    1: "ece3a3ea9c6c2eae2fcad3fb8d5318ec",
    2: "0dcdc6f1306c714ce2871c19b7aa8ed3",
    3: "970e27aeb1f1f7d7b85e6f42abe1164a",
    4: "895a3273f17d6adf659cddfda864c476",
    5: "a8d0b2fbedbb70da92ffd89820785bb3",
}
BLOCKS = {
    (1, 0): "cc4682611fce0241185c0bee2d66b1cc",
    (2, 0): "ae61673228e49b892ef646b5256e3e60",
    (3, 0): "493d141811b6685f7ec496cdae193b96",
    (4, 0): "b8580fc14e280dcc809d58436a986a2a",
    (5, 0): "85e7c8f3055eb7b9d87ce8c45e5ca542",
    (2, 99936): "8cc8e5f09ad6e89eb2256802d8bc5a7a",
    (3, 999936): "9884c72c93a2d233f0c2be93f6fbb743",
    (4, 199936): "08ba4b96331db95304c6b927b1ceeedb",
    (5, 1999936): "2cd33308ca67aa09ad3f90991c92942c",
}


def _digest(b):
    return hashlib.sha256(b).hexdigest()[:32]


@pytest.mark.parametrize("cfg", sorted(CENTERS))
def test_center_checksum(cfg):
    c = synth.center(cfg)
    assert len(c) == synth._l().synth_center_len(cfg)
    assert _digest(c) == CENTERS[cfg]


@pytest.mark.parametrize("cfg,first", sorted(BLOCKS))
def test_query_block_checksum(cfg, first):
    q, off = synth.queries(cfg, first, 64)
    assert _digest(q + off.astype("<i8").tobytes()) == BLOCKS[(cfg, first)]


def test_uneven_lengths_follow_the_quantile_table():
    """Config 3 draws its lengths from the committed lognormal quantile table
    (median 1,500 bp, sigma 0.8, clamped to [200, 20,000]) before mutation:
    the sample median and the share of short queries match the table's."""
    q, off = synth.queries(3, 0, 20000)
    L = np.diff(off)
    assert abs(np.median(L) - 1500) < 40
    assert 0.09 < (L < 550).mean() < 0.12
    assert L.min() >= 150 and L.max() <= 20500
