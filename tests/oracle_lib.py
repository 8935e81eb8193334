"""ctypes bindings for the checker libraries.

- ``Oracle``: oracle/liboracle.so, the plain-C restatement (test checker).
- ``RefLib``: oracle/_ref/libhalign_ref.so, the real reference path compiled
  from its own sources (only present where the reference tree was available
  at build time).
- ``Halign4``: oracle/_ref/halign4, the reference's own command-line program
  built from the same sources, run as a child process.

Both expose the same call shapes so a test can run one input through both.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libhalign_ref.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "halign4")

RIGHT = 0x02
GENERIC = 0x04
FLAG_RIGHT = GENERIC | RIGHT
FLAG_LEFT = GENERIC


class Ez(C.Structure):
    _fields_ = [("max", C.c_uint32), ("zdropped", C.c_int32), ("max_q", C.c_int32),
                ("max_t", C.c_int32), ("mqe", C.c_int32), ("mqe_t", C.c_int32),
                ("mte", C.c_int32), ("mte_q", C.c_int32), ("score", C.c_int32),
                ("n_cigar", C.c_int32)]

    def as_tuple(self):
        return tuple(getattr(self, f) for f, _ in self._fields_)


class SeedParams(C.Structure):
    _fields_ = [("f_top_frac", C.c_double), ("u_floor", C.c_uint64), ("u_ceil", C.c_uint64),
                ("q_occ_frac", C.c_double), ("sample_every_bp", C.c_uint64)]


class ChainParams(C.Structure):
    _fields_ = [("max_dist_x", C.c_int32), ("max_dist_y", C.c_int32), ("bw", C.c_int32),
                ("max_skip", C.c_int32), ("max_iter", C.c_int32), ("min_cnt", C.c_int32),
                ("min_score", C.c_int32), ("gap_penalty", C.c_float),
                ("skip_penalty", C.c_float)]


def pack_hit(h56, pos, rid=0, strand=False, span=15):
    """MinimizerHit (include/seed.h:206-265): x = h56<<8 | span, y = (rid | strand<<31)<<32 | pos."""
    x = (int(h56) << 8) | (int(span) & 0xff)
    y = ((int(rid) | (int(bool(strand)) << 31)) << 32) | int(pos)
    return x, y


def pack_hits(hits):
    xs, ys = zip(*[pack_hit(*h) for h in hits]) if hits else ((), ())
    return np.array(xs, np.uint64), np.array(ys, np.uint64)


ANCHOR_DTYPE = np.dtype([("hash", "<u8"), ("rid_ref", "<u4"), ("pos_ref", "<u4"),
                         ("rid_qry", "<u4"), ("pos_qry", "<u4"), ("span", "<u4"),
                         ("is_rev", "<u4")])

_P = C.c_void_p


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def encode_codes(s):
    """ScoreChar2Idx (include/align.h:72-89) for a bytes/str sequence."""
    if isinstance(s, str):
        s = s.encode()
    lut = np.full(256, 4, np.uint8)
    for ch, v in ((b"A", 0), (b"C", 1), (b"G", 2), (b"T", 3)):
        lut[ch[0]] = v
        lut[ch.lower()[0]] = v
    return lut[np.frombuffer(s, np.uint8)]


def cigar_str(cig):
    ops = "MIDNSHP=X"
    return "".join(f"{int(c) >> 4}{ops[int(c) & 15]}" for c in cig)


class _Base:
    def __init__(self, path):
        self.lib = C.CDLL(path)


class Oracle(_Base):
    def __init__(self, path=ORACLE_SO):
        super().__init__(path)
        L = self.lib
        L.or_ksw_extz2.argtypes = [C.c_int, _P, C.c_int, _P, C.c_int, C.c_int, C.POINTER(Ez), _P,
                                   C.c_int, _P]
        L.or_global_align_ksw2.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_int, _P,
                                           C.c_int, _P]
        L.or_extract_minimizer.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P,
                                           C.c_int, _P]
        L.or_collect_anchors.argtypes = [_P, _P, C.c_int, _P, _P, C.c_int, _P, C.c_int]
        L.or_chain_anchors.argtypes = [_P, C.c_int, _P, _P]
        L.or_collect_anchors_ex.argtypes = [_P, _P, C.c_int, _P, _P, C.c_int, _P, _P, C.c_int]
        L.or_chain_anchors_ex.argtypes = [_P, C.c_int, _P, _P, _P]
        L.or_global_align_mm2.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, _P, C.c_int,
                                          _P, _P, C.c_int, _P, _P]
        L.or_ref_global_align.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, _P, _P,
                                          C.c_int, _P, _P, C.c_int, _P, _P, _P]
        L.or_del_query_to_ref.argtypes = [C.c_char_p, C.c_int, _P, C.c_int, _P, C.c_int]
        L.or_consensus.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.c_uint64, _P,
                                   C.c_char_p]
        L.or_consensus.restype = C.c_int64
        L.or_auto_band.argtypes = [C.c_int, C.c_int]
        L.or_fastx_parse.argtypes = [C.c_char_p, C.c_int64, C.c_int, _P, _P, _P, _P, _P, _P, C.c_int64]
        L.or_keep_length_fasta.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_int64,
                                           C.c_int, C.c_char_p, _P, C.c_char_p, _P, _P, _P, _P, C.c_int64]
        L.or_keep_length_fasta.restype = C.c_int64
        L.or_hash2bit.argtypes = [C.c_uint64, C.c_uint32]
        L.or_hash2bit.restype = C.c_uint64
        L.or_mash_sketch.argtypes = [C.c_char_p, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_uint32, _P, C.c_int]
        L.or_mash_jaccard.argtypes = [_P, C.c_int, _P, C.c_int]
        L.or_mash_jaccard.restype = C.c_double
        L.or_fill_log2_lut.argtypes = [_P]
        self.lut = np.zeros(502, np.float32)
        L.or_fill_log2_lut(_ptr(self.lut))

    def auto_band(self, rlen, qlen):
        return self.lib.or_auto_band(rlen, qlen)

    def ksw(self, query_codes, target_codes, w, flag):
        """Returns (ez tuple, cigar array, cells)."""
        q = np.ascontiguousarray(query_codes, np.uint8)
        t = np.ascontiguousarray(target_codes, np.uint8)
        ez = Ez()
        cells = np.zeros(1, np.int64)
        cap = len(q) + len(t) + 4
        cig = np.zeros(cap, np.uint32)
        n = self.lib.or_ksw_extz2(len(q), _ptr(q), len(t), _ptr(t), w, flag, C.byref(ez),
                                  _ptr(cig), cap, _ptr(cells))
        assert n >= 0
        return ez.as_tuple(), cig[:n].copy(), int(cells[0])

    def global_align_ksw2(self, ref, qry, flag=FLAG_RIGHT):
        ref = ref.encode() if isinstance(ref, str) else ref
        qry = qry.encode() if isinstance(qry, str) else qry
        cap = len(ref) + len(qry) + 4
        cig = np.zeros(cap, np.uint32)
        cells = np.zeros(1, np.int64)
        n = self.lib.or_global_align_ksw2(ref, len(ref), qry, len(qry), flag, _ptr(cig), cap,
                                          _ptr(cells))
        assert n >= 0
        return cig[:n].copy(), int(cells[0])

    def minimizers(self, seq, k=15, w=10, non_canonical=True):
        seq = seq.encode() if isinstance(seq, str) else seq
        cap = len(seq) + 1
        x = np.zeros(cap, np.uint64)
        y = np.zeros(cap, np.uint64)
        ub = np.zeros(1, np.int32)
        n = self.lib.or_extract_minimizer(seq, len(seq), k, w, int(non_canonical), _ptr(x),
                                          _ptr(y), cap, _ptr(ub))
        assert n >= 0
        return x[:n].copy(), y[:n].copy(), int(ub[0])

    def collect_anchors(self, rx, ry, qx, qy):
        cap = max(16, len(qx) * 64)
        out = np.zeros(cap, ANCHOR_DTYPE)
        n = self.lib.or_collect_anchors(_ptr(rx), _ptr(ry), len(rx), _ptr(qx), _ptr(qy), len(qx),
                                        _ptr(out), cap)
        assert n >= 0
        return out[:n].copy()

    def chain(self, anchors):
        a = np.ascontiguousarray(anchors.copy())
        out = np.zeros(max(1, len(a)), ANCHOR_DTYPE)
        n = self.lib.or_chain_anchors(_ptr(a), len(a), _ptr(self.lut), _ptr(out))
        return a, out[:n].copy()

    def collect_anchors_ex(self, rx, ry, qx, qy, f_top_frac=2e-4, u_floor=10, u_ceil=1000000,
                           q_occ_frac=0.01, sample_every_bp=500):
        """collect_anchors with explicit SeedFilterParams (include/anchor.h:65-71)."""
        p = SeedParams(f_top_frac, u_floor, u_ceil, q_occ_frac, sample_every_bp)
        cap = max(16, len(qx) * max(1, len(rx)))
        out = np.zeros(cap, ANCHOR_DTYPE)
        n = self.lib.or_collect_anchors_ex(_ptr(rx), _ptr(ry), len(rx), _ptr(qx), _ptr(qy),
                                           len(qx), C.cast(C.pointer(p), _P), _ptr(out), cap)
        assert n >= 0
        return out[:n].copy()

    def chain_ex(self, anchors, **kw):
        """chainAnchors with explicit ChainParams (include/anchor.h:116-126)."""
        prm = dict(max_dist_x=5000, max_dist_y=5000, bw=500, max_skip=25, max_iter=5000,
                   min_cnt=3, min_score=40, gap_penalty=0.01, skip_penalty=0.01)
        prm.update(kw)
        p = ChainParams(**prm)
        a = np.ascontiguousarray(anchors.copy())
        out = np.zeros(max(1, len(a)), ANCHOR_DTYPE)
        n = self.lib.or_chain_anchors_ex(_ptr(a), len(a), C.cast(C.pointer(p), _P),
                                         _ptr(self.lut), _ptr(out))
        assert n >= 0
        return a, out[:n].copy()

    def mm2(self, ref, qry, anchors):
        ref = ref.encode() if isinstance(ref, str) else ref
        qry = qry.encode() if isinstance(qry, str) else qry
        a = np.ascontiguousarray(anchors)
        cap = len(ref) + len(qry) + 4
        cig = np.zeros(cap, np.uint32)
        cells = np.zeros(1, np.int64)
        st = np.zeros(4, np.int32)
        n = self.lib.or_global_align_mm2(ref, len(ref), qry, len(qry), _ptr(a), len(a),
                                         _ptr(self.lut), _ptr(cig), cap, _ptr(cells), _ptr(st))
        assert n >= 0
        return cig[:n].copy(), int(cells[0]), st.copy()

    def ref_global_align(self, ref, qry, ref_hits=None):
        ref = ref.encode() if isinstance(ref, str) else ref
        qry = qry.encode() if isinstance(qry, str) else qry
        rx = ry = None
        nref = 0
        if ref_hits is not None:
            rx, ry = ref_hits
            nref = len(rx)
        cap = len(ref) + len(qry) + 4
        cig = np.zeros(cap, np.uint32)
        cells = np.zeros(1, np.int64)
        st = np.zeros(4, np.int32)
        ub = np.zeros(1, np.int32)
        n = self.lib.or_ref_global_align(ref, len(ref), qry, len(qry), _ptr(rx), _ptr(ry), nref,
                                         _ptr(self.lut), _ptr(cig), cap, _ptr(cells), _ptr(st),
                                         _ptr(ub))
        assert n >= 0
        return cig[:n].copy(), int(cells[0]), st.copy(), int(ub[0])

    def del_query_to_ref(self, qry, cig):
        qry = qry.encode() if isinstance(qry, str) else qry
        c = np.ascontiguousarray(cig, np.uint32)
        cap = int(sum(int(x) >> 4 for x in c)) + len(qry) + 1
        out = np.zeros(cap, np.uint8)
        n = self.lib.or_del_query_to_ref(qry, len(qry), _ptr(c), len(c), _ptr(out), cap)
        assert n >= 0
        return out[:n].tobytes()

    def consensus(self, rows, batch=4096, seq_limit=0):
        """rows: (nrows, aln_len) uint8 array. Returns (counts[aln,7], consensus bytes, n)."""
        rows = np.ascontiguousarray(rows, np.uint8)
        nrows, aln = rows.shape
        counts = np.zeros((aln, 7), np.uint32)
        cons = C.create_string_buffer(aln)
        n = self.lib.or_consensus(rows.tobytes(), nrows, aln, batch, seq_limit, _ptr(counts), cons)
        return counts, cons.raw[:aln], int(n)


    def fastx_parse(self, data, max_records=1 << 30):
        """Records (id, desc, seq) of an in-memory FASTA/FASTQ file, or an int
        error code (-2: malformed FASTQ)."""
        cap = len(data) + 1
        bufs = [np.zeros(cap, np.uint8) for _ in range(3)]
        offs = [np.zeros(len(data) + 2, np.int64) for _ in range(3)]
        n = self.lib.or_fastx_parse(data, len(data), min(max_records, len(data) + 1), _ptr(bufs[0]),
                                    _ptr(offs[0]), _ptr(bufs[1]), _ptr(offs[1]), _ptr(bufs[2]),
                                    _ptr(offs[2]), cap)
        if n < 0:
            return n
        return [tuple(bufs[k][offs[k][i]:offs[k][i + 1]].tobytes() for k in range(3)) for i in range(n)]

    def mash_sketch(self, seq, k=15, size=2000, noncanonical=True, seed=42):
        """mash::sketchFromSequence: sorted uint64 hashes."""
        return _mash_sketch(self.lib.or_mash_sketch, seq, k, size, noncanonical, seed)

    def mash_jaccard(self, a, b):
        a = np.ascontiguousarray(a, np.uint64)
        b = np.ascontiguousarray(b, np.uint64)
        return self.lib.or_mash_jaccard(_ptr(a), len(a), _ptr(b), len(b))

    def keep_length_fasta(self, center_id, center_desc, center_seq, queries, names, cigars):
        """The merged keep-length FASTA bytes (or_keep_length_fasta)."""
        qb, qo = _concat_bytes(queries)
        nb, no = _concat_bytes(names)
        cb = np.concatenate([np.asarray(c, np.uint32) for c in cigars]) if cigars else np.zeros(0, np.uint32)
        co = np.zeros(len(cigars) + 1, np.int64)
        co[1:] = np.cumsum([len(c) for c in cigars])
        cb = np.ascontiguousarray(cb, np.uint32)
        cap = (len(center_id) + len(center_desc) + len(center_seq) + 4 +
               sum(len(n) + 3 for n in names) + sum(len(q) for q in queries) +
               sum(int(sum(int(x) >> 4 for x in c)) for c in cigars) + 16)
        out = np.zeros(cap, np.uint8)
        n = self.lib.or_keep_length_fasta(center_id, len(center_id), center_desc, len(center_desc), center_seq,
                                          len(center_seq), len(queries), qb, _ptr(qo), nb, _ptr(no), _ptr(cb),
                                          _ptr(co), _ptr(out), cap)
        assert n >= 0, n
        return out[:n].tobytes()


def _mash_sketch(fn, seq, k, size, noncanonical, seed):
    seq = seq.encode() if isinstance(seq, str) else bytes(seq)
    cap = max(1, min(size, len(seq) + 1))
    out = np.zeros(cap, np.uint64)
    n = fn(seq, len(seq), k, size, int(noncanonical), seed, _ptr(out), cap)
    assert n >= 0, n
    return out[:n].copy()


def _concat_bytes(items):
    off = np.zeros(len(items) + 1, np.int64)
    off[1:] = np.cumsum([len(x) for x in items]) if items else []
    return b"".join(items), off


class RefLib(_Base):
    """The compiled reference (same call shapes as Oracle)."""

    def __init__(self, path=REF_SO):
        super().__init__(path)
        L = self.lib
        L.ref_ksw_extz2.argtypes = [C.c_int, _P, C.c_int, _P, C.c_int, C.c_int, C.POINTER(Ez), _P,
                                    C.c_int]
        L.ref_global_align_ksw2.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_int,
                                            _P, C.c_int]
        L.ref_extract_minimizer.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, _P,
                                            _P, C.c_int]
        L.ref_collect_anchors.argtypes = [_P, _P, C.c_int, _P, _P, C.c_int, _P, C.c_int]
        L.ref_chain_anchors.argtypes = [_P, C.c_int, _P]
        L.ref_global_align_mm2.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, _P, C.c_int,
                                           _P, C.c_int]
        L.ref_ref_global_align.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, _P, C.c_int]
        L.ref_ref_global_align_pre.argtypes = [C.c_char_p, C.c_int, _P, _P, C.c_int, C.c_char_p, C.c_int,
                                               _P, C.c_int]
        L.ref_consensus.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint64, C.c_int,
                                    C.c_int, C.c_char_p, C.c_int]

    def ksw(self, query_codes, target_codes, w, flag):
        q = np.ascontiguousarray(query_codes, np.uint8)
        t = np.ascontiguousarray(target_codes, np.uint8)
        ez = Ez()
        cap = len(q) + len(t) + 4
        cig = np.zeros(cap, np.uint32)
        n = self.lib.ref_ksw_extz2(len(q), _ptr(q), len(t), _ptr(t), w, flag, C.byref(ez),
                                   _ptr(cig), cap)
        assert n >= 0
        return ez.as_tuple(), cig[:n].copy()

    def global_align_ksw2(self, ref, qry, flag=FLAG_RIGHT):
        ref = ref.encode() if isinstance(ref, str) else ref
        qry = qry.encode() if isinstance(qry, str) else qry
        cap = len(ref) + len(qry) + 4
        cig = np.zeros(cap, np.uint32)
        n = self.lib.ref_global_align_ksw2(ref, len(ref), qry, len(qry), flag, _ptr(cig), cap)
        assert n >= 0
        return cig[:n].copy()

    def minimizers(self, seq, k=15, w=10, non_canonical=True):
        seq = seq.encode() if isinstance(seq, str) else seq
        cap = len(seq) + 1
        x = np.zeros(cap, np.uint64)
        y = np.zeros(cap, np.uint64)
        n = self.lib.ref_extract_minimizer(seq, len(seq), k, w, int(non_canonical), _ptr(x),
                                           _ptr(y), cap)
        assert n >= 0
        return x[:n].copy(), y[:n].copy()

    def collect_anchors(self, rx, ry, qx, qy):
        cap = max(16, len(qx) * 64)
        out = np.zeros(cap, ANCHOR_DTYPE)
        n = self.lib.ref_collect_anchors(_ptr(rx), _ptr(ry), len(rx), _ptr(qx), _ptr(qy),
                                         len(qx), _ptr(out), cap)
        assert n >= 0
        return out[:n].copy()

    def chain(self, anchors):
        a = np.ascontiguousarray(anchors.copy())
        out = np.zeros(max(1, len(a)), ANCHOR_DTYPE)
        n = self.lib.ref_chain_anchors(_ptr(a), len(a), _ptr(out))
        return a, out[:n].copy()

    def mm2(self, ref, qry, anchors):
        ref = ref.encode() if isinstance(ref, str) else ref
        qry = qry.encode() if isinstance(qry, str) else qry
        a = np.ascontiguousarray(anchors)
        cap = len(ref) + len(qry) + 4
        cig = np.zeros(cap, np.uint32)
        n = self.lib.ref_global_align_mm2(ref, len(ref), qry, len(qry), _ptr(a), len(a),
                                          _ptr(cig), cap)
        assert n >= 0
        return cig[:n].copy()

    def ref_global_align(self, ref, qry):
        ref = ref.encode() if isinstance(ref, str) else ref
        qry = qry.encode() if isinstance(qry, str) else qry
        cap = len(ref) + len(qry) + 4
        cig = np.zeros(cap, np.uint32)
        n = self.lib.ref_ref_global_align(ref, len(ref), qry, len(qry), _ptr(cig), cap)
        assert n >= 0
        return cig[:n].copy()

    def ref_global_align_pre(self, ref, ref_hits, qry):
        """RefAligner::globalAlign with the center's minimizers given (x, y arrays)."""
        ref = ref.encode() if isinstance(ref, str) else ref
        qry = qry.encode() if isinstance(qry, str) else qry
        rx = np.ascontiguousarray(ref_hits[0], np.uint64)
        ry = np.ascontiguousarray(ref_hits[1], np.uint64)
        cap = len(ref) + len(qry) + 4
        cig = np.zeros(cap, np.uint32)
        n = self.lib.ref_ref_global_align_pre(ref, len(ref), _ptr(rx), _ptr(ry), len(rx), qry, len(qry),
                                              _ptr(cig), cap)
        assert n >= 0
        return cig[:n].copy()

    def read_fastx(self, path, max_records=1 << 30):
        """seq_io::KseqReader records (id, desc, seq), or -2 when it throws."""
        size = max(os.path.getsize(path), 1)
        if path.endswith(".gz"):
            import gzip
            size = max(len(gzip.open(path).read()), 1)
        cap = size + 1
        bufs = [np.zeros(cap, np.uint8) for _ in range(3)]
        offs = [np.zeros(size + 2, np.int64) for _ in range(3)]
        L = self.lib
        L.ref_read_fastx.argtypes = [C.c_char_p, C.c_int, _P, _P, _P, _P, _P, _P, C.c_int64]
        n = L.ref_read_fastx(path.encode(), min(max_records, size + 1), _ptr(bufs[0]), _ptr(offs[0]),
                             _ptr(bufs[1]), _ptr(offs[1]), _ptr(bufs[2]), _ptr(offs[2]), cap)
        if n < 0:
            return n
        return [tuple(bufs[k][offs[k][i]:offs[k][i + 1]].tobytes() for k in range(3)) for i in range(n)]

    def mash_sketch(self, seq, k=15, size=2000, noncanonical=True, seed=42):
        L = self.lib
        L.ref_mash_sketch.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P, C.c_int]
        return _mash_sketch(L.ref_mash_sketch, seq, k, size, noncanonical, seed)

    def mash_jaccard(self, a, b, k=15):
        L = self.lib
        L.ref_mash_jaccard.argtypes = [_P, C.c_int, _P, C.c_int, C.c_int]
        L.ref_mash_jaccard.restype = C.c_double
        a = np.ascontiguousarray(a, np.uint64)
        b = np.ascontiguousarray(b, np.uint64)
        return L.ref_mash_jaccard(_ptr(a), len(a), _ptr(b), len(b), k)

    def hash2bit(self, code, seed):
        L = self.lib
        L.ref_hash2bit.argtypes = [C.c_uint64, C.c_uint32]
        L.ref_hash2bit.restype = C.c_uint64
        return L.ref_hash2bit(code, seed)

    def preprocess(self, input_path, work_dir, cons_n):
        L = self.lib
        L.ref_preprocess.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.ref_preprocess.restype = C.c_int64
        return L.ref_preprocess(input_path.encode(), work_dir.encode(), cons_n)

    def consensus_file(self, fasta, out_fa, out_json, seq_limit=0, batch=4096, threads=1,
                       aln_len=1 << 20):
        cons = C.create_string_buffer(aln_len)
        n = self.lib.ref_consensus(fasta.encode(), out_fa.encode(), out_json.encode(), seq_limit,
                                   threads, batch, cons, aln_len)
        assert n >= 0
        return cons.raw[:n]


def ref_available():
    return os.path.exists(REF_SO)


def parse_cigar_string(s):
    """cigar::stringToCigar (src/align/cigar.cpp) of a SAM CIGAR field: uint32
    len << 4 | op, ops M I D N S H P = X as BAM numbers them."""
    ops = {b"M": 0, b"I": 1, b"D": 2, b"N": 3, b"S": 4, b"H": 5, b"P": 6, b"=": 7, b"X": 8}
    out, num = [], 0
    for ch in s:
        c = bytes([ch])
        if c.isdigit():
            num = num * 10 + ch - 48
        else:
            out.append(num << 4 | ops[c])
            num = 0
    return np.array(out, np.uint32)


class Halign4:
    """The reference's own program (oracle/_ref/halign4: src/halign4.cpp main
    with RefAligner, src/align/ref_aligner.cpp), run as a child process with
    --save-workdir so its per-thread SAM files can be read back."""

    COPY = "cp {input} {output}"

    def __init__(self, path=REF_BIN):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path

    def run(self, input_path, output_path, work_dir, center_path=None, keep_length=True, msa_cmd=COPY,
            threads=1, cons_n=None, timeout=3600):
        """Runs halign4 and returns its exit code; work_dir must not exist
        (the reference wants an empty working directory)."""
        import subprocess
        args = [self.path, "-i", input_path, "-o", output_path, "-w", work_dir, "-p", msa_cmd, "-t",
                str(threads), "--save-workdir"]
        if center_path:
            args += ["-c", center_path]
        if keep_length:
            args.append("--keep-length")
        if cons_n is not None:
            args += ["--cons-n", str(cons_n)]
        r = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=timeout,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
        self.last_log = r.stdout
        return r.returncode

    @staticmethod
    def sam_records(work_dir, thread=0):
        """The records of one thread's two SAM files (result/thread<t>.sam and
        result/thread<t>_insertion.sam): a list of (qname, rname, CIGAR words,
        has_insertion) in file order, the plain file first."""
        out = []
        res = os.path.join(work_dir, "result")
        for name, ins in ((f"thread{thread}.sam", False), (f"thread{thread}_insertion.sam", True)):
            p = os.path.join(res, name)
            if not os.path.exists(p):
                continue
            for line in open(p, "rb").read().split(b"\n"):
                if not line or line.startswith(b"@"):
                    continue
                f = line.split(b"\t")
                out.append((f[0], f[2], parse_cigar_string(f[5]) if f[5] != b"*" else np.zeros(0, np.uint32), ins))
        return out
