"""Several references (RefAligner with a multi-record center file) on the
CPU: the oracle's mash sketches, its Jaccard choice of a reference and its
keep-length merge against the reference-written fixtures mash_cases.npz and
msa_refs.npz (tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

from oracle_lib import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _seg(z, k, i):
    return z[k][z[k + "_off"][i]:z[k + "_off"][i + 1]]


def mash_fixture():
    z = np.load(os.path.join(GOLDEN, "mash_cases.npz"))
    cases = [(_seg(z, "seq", i).tobytes(), tuple(int(x) for x in z["params"][i]), _seg(z, "sketch", i))
             for i in range(len(z["params"]))]
    refs = [_seg(z, "ref", i).tobytes() for i in range(len(z["ref_off"]) - 1)]
    qs = [_seg(z, "qry", i).tobytes() for i in range(len(z["qry_off"]) - 1)]
    return cases, refs, qs, z["best"], z["jaccard"]


def msa_refs_fixture():
    z = np.load(os.path.join(GOLDEN, "msa_refs.npz"))
    return [(_seg(z, "center", i).tobytes(), _seg(z, "aligned", i).tobytes() if z["has_aligned"][i] else None,
             _seg(z, "input", i).tobytes(), _seg(z, "output", i).tobytes()) for i in range(len(z["has_aligned"]))]


def test_sketches_vs_reference():
    O = Oracle()
    cases, _, _, _, _ = mash_fixture()
    assert any(len(s) == 2000 for _, _, s in cases)   # full sketches are covered
    assert any(p[0] == 32 for _, p, _ in cases)        # and the k = 32 shift
    for i, (seq, (k, size, nc, seed), expect) in enumerate(cases):
        assert np.array_equal(O.mash_sketch(seq, k, size, nc, seed), expect), i


def test_hash_known_values():
    """XXH3-64 of 8 bytes, pinned against values the reference wrote: the
    one-hash sketches (sketch size 1) hold the smallest hash of their k-mers,
    so a wrong secret word, rotation or multiplier changes them."""
    O = Oracle()
    cases, _, _, _, _ = mash_fixture()
    ones = [(seq, p, s) for seq, p, s in cases if p[1] == 1 and len(s) == 1]
    assert len(ones) > 10
    for seq, (k, size, nc, seed), s in ones:
        assert O.mash_sketch(seq, k, 1, nc, seed)[0] == s[0]


def test_reference_choice_vs_reference():
    O = Oracle()
    _, refs, qs, best, jac = mash_fixture()
    rsk = [O.mash_sketch(r) for r in refs]
    for i, q in enumerate(qs):
        qsk = O.mash_sketch(q)
        js = [O.mash_jaccard(qsk, x) for x in rsk]
        b = max(range(len(js)), key=lambda r: (js[r], -r))
        assert b == best[i] and js[b] == jac[i], i
    assert 4 not in set(best.tolist())   # reference 4 equals reference 1: never the first maximum


def oracle_keep_length_refs(O, center_text, aligned_text, input_text):
    """The multi-reference run restated with the oracle: the reference by
    Jaccard, its CIGAR (index from its own sequence), the row projected and
    spread over the reference's aligned record, the merge order."""
    refs = O.fastx_parse(center_text)
    aligned = O.fastx_parse(aligned_text if aligned_text is not None else center_text)
    by_id = {}
    for i, (n, _, _) in enumerate(aligned):
        by_id[n] = i
    rsk = [O.mash_sketch(s) for _, _, s in refs]
    hits = [O.minimizers(s)[:2] for _, _, s in refs]
    head = b"".join(b">" + n + (b" " + d if d else b"") + b"\n" +
                    bytes(b if b in b"ACGTU-" else ord("N") for b in s.upper()) + b"\n"
                    for n, d, s in aligned)
    groups = ([], [])
    for name, _, q in O.fastx_parse(input_text):
        qsk = O.mash_sketch(q)
        js = [O.mash_jaccard(qsk, x) for x in rsk]
        r = max(range(len(js)), key=lambda t: (js[t], -t))
        ref = refs[r][2]
        cig = O.ref_global_align(ref, q, ref_hits=hits[r] if len(hits[r][0]) else None)[0]
        row = O.del_query_to_ref(q, cig)
        arow = aligned[by_id[refs[r][0]]][2]
        cols = [c for c, b in enumerate(arow) if b != ord("-")]
        full = bytearray(b"-" * len(arow))
        for j, c in enumerate(cols):
            full[c] = row[j]
        if name.startswith(b"@"):
            continue
        ins = any((int(x) & 0xF) == 1 for x in cig)
        groups[0 if ins else 1].append(b">" + name + b"\n" + bytes(full) + b"\n")
    return head + b"".join(groups[0]) + b"".join(groups[1])


@pytest.mark.parametrize("case", range(3))
def test_keep_length_refs_oracle_vs_reference(case):
    O = Oracle()
    center, aligned, inp, expect = msa_refs_fixture()[case]
    assert oracle_keep_length_refs(O, center, aligned, inp) == expect
