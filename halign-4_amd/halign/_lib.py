"""ctypes binding of libhalign_hip.so (C ABI in include/halign_hip.h).

The product path runs only through this library; if it is missing the import
fails loudly (there is no CPU fallback).
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libhalign_hip.so")

HALIGN_OK = 0
ERRORS = {-1: "invalid argument", -2: "unsupported configuration", -3: "HIP runtime error",
          -4: "out of memory", -5: "output buffer too small", -6: "kernel status", -7: "I/O or format error"}

_P = C.c_void_p


class DpTask(C.Structure):
    _fields_ = [("qlen", C.c_int32), ("tlen", C.c_int32), ("w", C.c_int32), ("flag", C.c_int32),
                ("query_off", C.c_int64), ("target_off", C.c_int64)]


DP_RESULT_DTYPE = np.dtype([("max", "<u4"), ("zdropped", "<i4"), ("max_q", "<i4"),
                            ("max_t", "<i4"), ("mqe", "<i4"), ("mqe_t", "<i4"), ("mte", "<i4"),
                            ("mte_q", "<i4"), ("score", "<i4"), ("n_cigar", "<i4"),
                            ("status", "<i4"), ("pad_", "<i4"), ("cells", "<i8"),
                            ("cigar_off", "<i8")])
DP_TASK_DTYPE = np.dtype([("qlen", "<i4"), ("tlen", "<i4"), ("w", "<i4"), ("flag", "<i4"),
                          ("query_off", "<i8"), ("target_off", "<i8")])


class HalignError(RuntimeError):
    pass


HALIGN_PENDING = 1


class MsaStats(C.Structure):
    _fields_ = [("records", C.c_int64), ("insertion_records", C.c_int64), ("bytes", C.c_int64),
                ("seconds", C.c_double), ("gpu_seconds", C.c_double), ("input_wait_seconds", C.c_double)]


class ExtzT(C.Structure):
    """halign_extz_t: the field layout of ksw2's ksw_extz_t (include/ksw2.h:33-42)."""
    _fields_ = [("max", C.c_uint32, 31), ("zdropped", C.c_uint32, 1), ("max_q", C.c_int), ("max_t", C.c_int),
                ("mqe", C.c_int), ("mqe_t", C.c_int), ("mte", C.c_int), ("mte_q", C.c_int), ("score", C.c_int),
                ("m_cigar", C.c_int), ("n_cigar", C.c_int), ("reach_end", C.c_int),
                ("cigar", C.POINTER(C.c_uint32))]

    def cigar_array(self):
        return np.array([self.cigar[i] for i in range(self.n_cigar)], np.uint32)


# dna5 scoring matrix of the path (reference include/align.h:109-116)
DNA5_MAT = (C.c_int8 * 25)(5, -4, -4, -4, 0, -4, 5, -4, -4, 0, -4, -4, 5, -4, 0, -4, -4, -4, 5, 0, 0, 0, 0, 0, 0)
_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


def free_extz(ez):
    """Releases ez.cigar with free(), as the reference caller does (src/align/align.cpp:63)."""
    if ez.cigar:
        _libc.free(C.cast(ez.cigar, C.c_void_p))
    ez.cigar = C.POINTER(C.c_uint32)()
    ez.m_cigar = ez.n_cigar = 0


class TuneEntry(C.Structure):
    _fields_ = [("waves", C.c_int32), ("lanes", C.c_int32), ("ring", C.c_int32), ("uniform", C.c_int32),
                ("jobs_in_flight", C.c_int32), ("npoints", C.c_int32), ("ncol", C.c_int32 * 4),
                ("ns_diag", C.c_double * 4), ("ns_job", C.c_double)]


class BucketStat(C.Structure):
    _fields_ = [("family", C.c_int32), ("shape", C.c_int32), ("stream", C.c_int32), ("right", C.c_int32),
                ("jobs", C.c_int64), ("cells", C.c_int64), ("kernel_ms", C.c_double), ("est_ms", C.c_double)]


class ServerMemory(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("batches", "max_batch_estimate", "sum_worker_estimates", "batch_budget",
                                         "resident", "traceback_reserved", "device_used", "working_measured")]


class ServerConfig(C.Structure):
    _fields_ = [("workers", C.c_int32), ("max_batch_queries", C.c_int32), ("hbm_budget", C.c_int64),
                ("chunk_queries", C.c_int32), ("pad_", C.c_int32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    lib.halign_init.argtypes = [C.c_int, C.POINTER(_P)]
    lib.halign_free.argtypes = [_P]
    lib.halign_free.restype = None
    lib.halign_last_error.argtypes = [_P]
    lib.halign_last_error.restype = C.c_char_p
    lib.halign_version.restype = C.c_char_p
    lib.halign_set_kernel_policy.argtypes = [_P, C.c_int]
    lib.halign_ring_shapes.argtypes = [_P, _P, _P, C.c_int32]
    lib.halign_qfix_shapes.argtypes = [_P, _P, C.c_int32]
    lib.halign_ksw_extz2.argtypes = [_P, _P, C.c_int, _P, C.c_int, _P, C.c_int8, _P, C.c_int8, C.c_int8,
                                     C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(ExtzT)]
    lib.halign_dp_batch.argtypes = [_P, C.c_int32, _P, _P, _P, C.c_int32, C.c_int32, _P, _P,
                                    C.c_int64, C.POINTER(C.c_double)]
    lib.halign_global_align_batch.argtypes = [_P, C.c_int32, _P, _P, _P, _P, _P, _P, C.c_int64,
                                              _P, _P]
    lib.halign_pairset_create.argtypes = [_P, C.c_char_p, C.c_int32, C.c_int32, C.c_char_p, _P,
                                          C.POINTER(_P)]
    lib.halign_pairset_run.argtypes = [_P, C.c_int32, C.c_int32, C.c_int32,
                                       C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    lib.halign_pairset_result.argtypes = [_P, C.c_int32, C.POINTER(C.c_int32), _P, C.c_int32,
                                          C.POINTER(C.c_int32)]
    lib.halign_consensus.argtypes = [_P, _P, C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_int64,
                                     C.c_uint64, _P, _P, C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_double)]
    lib.halign_refset_create.argtypes = [_P, C.c_int32, _P, _P, C.c_char_p, C.c_int32, C.POINTER(_P)]
    lib.halign_refset_align.argtypes = [_P, C.c_int32, C.c_char_p, _P, C.c_int32, _P, _P, C.c_int64, _P, _P,
                                        C.c_int64, _P, _P, _P]
    lib.halign_refset_free.argtypes = [_P]
    lib.halign_refset_free.restype = None
    lib.halign_align_sam.argtypes = [_P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32,
                                     C.c_int32, C.POINTER(MsaStats)]
    lib.halign_consensus_noalloc.argtypes = [_P, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                             C.c_uint64, _P, _P, C.POINTER(C.c_uint64)]
    lib.halign_set_seed_policy.argtypes = [_P, C.c_int]
    lib.halign_extract_minimizers.argtypes = [_P, C.c_int32, C.c_char_p, _P, C.c_int32, C.c_int32,
                                              C.c_int32, _P, _P, C.c_int64, _P, _P]
    lib.halign_center_create.argtypes = [_P, C.c_char_p, C.c_int32, C.POINTER(_P)]
    lib.halign_center_create_indexed.argtypes = [_P, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32,
                                                 C.POINTER(_P)]
    lib.halign_center_align_fasta.argtypes = [_P, C.c_int32, C.c_char_p, _P, C.c_char_p, _P, _P, C.c_int64,
                                              C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                              C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    lib.halign_fastx_open.argtypes = [C.c_char_p, C.POINTER(_P)]
    lib.halign_fastx_next.argtypes = [_P, C.c_int32, C.POINTER(C.c_int32)] + [C.POINTER(_P)] * 6
    lib.halign_fastx_error.argtypes = [_P]
    lib.halign_fastx_error.restype = C.c_char_p
    lib.halign_fastx_close.argtypes = [_P]
    lib.halign_msa_keep_length.argtypes = [_P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32,
                                           C.POINTER(MsaStats)]
    lib.halign_msa_keep_length_refs.argtypes = [_P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32,
                                                C.POINTER(MsaStats)]
    lib.halign_mash_sketch.argtypes = [_P, C.c_int32, C.c_char_p, _P, C.c_int32, C.c_int32, C.c_int32, C.c_uint32,
                                       _P, C.c_int64, _P]
    lib.halign_mash_best.argtypes = [_P, C.c_int32, _P, _P, C.c_int32, _P, _P, _P, _P]
    lib.halign_topk_longest.argtypes = [_P, C.c_char_p, C.c_int32, C.c_char_p, C.POINTER(C.c_int64)]
    lib.halign_consensus_file.argtypes = [_P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint64, C.c_int64, _P,
                                          C.c_int64, C.POINTER(C.c_int64)]
    lib.halign_msa_auto_center.argtypes = [_P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32, C.c_int32,
                                           C.POINTER(MsaStats)]
    lib.halign_center_free.argtypes = [_P]
    lib.halign_center_free.restype = None
    lib.halign_center_anchors.argtypes = [_P, C.c_int32, C.c_char_p, _P, _P, C.c_int64, _P]
    lib.halign_center_chains.argtypes = [_P, C.c_int32, C.c_char_p, _P, _P, C.c_int64, _P]
    lib.halign_center_align.argtypes = [_P, C.c_int32, C.c_char_p, _P, _P, C.c_int64, _P, _P,
                                        C.POINTER(C.c_double)]
    lib.halign_center_align_rows.argtypes = [_P, C.c_int32, C.c_char_p, _P, _P, C.c_int64, _P, _P,
                                             C.c_int64, _P, _P, _P, C.POINTER(C.c_double)]
    lib.halign_pairset_free.argtypes = [_P]
    lib.halign_pairset_free.restype = None
    lib.halign_set_traceback_budget.argtypes = [_P, C.c_int64]
    lib.halign_tune.argtypes = [_P]
    lib.halign_set_bucket_timing.argtypes = [_P, C.c_int32]
    lib.halign_last_buckets.argtypes = [_P, C.POINTER(BucketStat), C.c_int32, C.POINTER(C.c_int32)]
    lib.halign_tune_table.argtypes = [_P, C.POINTER(TuneEntry), C.c_int32, C.POINTER(C.c_int32)]
    lib.halign_plan_shape.argtypes = [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    lib.halign_set_graph_mode.argtypes = [_P, C.c_int32]
    lib.halign_graph_counts.argtypes = [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.halign_server_create.argtypes = [C.c_int, C.c_char_p, C.c_int32, C.POINTER(ServerConfig),
                                         C.POINTER(_P)]
    lib.halign_server_submit.argtypes = [_P, C.c_int32, C.c_char_p, _P, C.POINTER(C.c_int64)]
    lib.halign_server_wait.argtypes = [_P, C.c_int64, C.c_int32]
    lib.halign_server_result.argtypes = [_P, C.c_int64, _P, C.c_int64, _P, _P, _P,
                                         C.POINTER(C.c_int64)]
    lib.halign_server_stats.argtypes = [_P, C.POINTER(ServerMemory)]
    lib.halign_server_last_error.argtypes = [_P]
    lib.halign_server_last_error.restype = C.c_char_p
    lib.halign_server_free.argtypes = [_P]
    lib.halign_server_free.restype = None
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def ring_shapes():
    """[(waves per workgroup, lanes per job, ring size in blocks)] of the
    ring-form kernel, by shape index; waves 0 is a sub-wave shape (groups of 4
    or 16 lanes, one job each)."""
    n = lib().halign_ring_shapes(None, None, None, 0)
    w = np.zeros(n, np.int32)
    g = np.zeros(n, np.int32)
    r = np.zeros(n, np.int32)
    lib().halign_ring_shapes(_ptr(w), _ptr(g), _ptr(r), n)
    return [(int(a), int(b), int(c)) for a, b, c in zip(w, g, r)]


def qfix_shapes():
    """[(lanes per job, slots per lane)] of the query-fixed kernel, by index:
    full-band jobs whose query is the short side, 16 * lanes * slots query
    bases at most."""
    n = lib().halign_qfix_shapes(None, None, 0)
    g = np.zeros(n, np.int32)
    s = np.zeros(n, np.int32)
    lib().halign_qfix_shapes(_ptr(g), _ptr(s), n)
    return [(int(a), int(b)) for a, b in zip(g, s)]


def qfix_policy(lanes, slots):
    """Kernel policy that forces the query-fixed shape (lanes, slots) on the
    jobs it holds (the automatic choice for the others)."""
    hits = [q for q, (g, s) in enumerate(qfix_shapes()) if g == lanes and s == slots]
    if not hits:
        raise ValueError(f"no query-fixed shape with {lanes} lanes and {slots} slots")
    return 3 + len(ring_shapes()) + hits[0]


def shape_names():
    """Names of the ring shapes (WxS, WxSd for the two-per-SIMD form, subLxS,
    subLxSu for the uniform-ticket form) followed by the query-fixed shapes
    (qLxS), by the index the planner and the bucket statistics use."""
    seen, names = set(), []
    for w, l, r in ring_shapes():
        u = (w, l, r) in seen
        seen.add((w, l, r))
        names.append(f"{w}x{r // l}{'d' if u else ''}" if w else f"sub{l}x{r // l}{'u' if u else ''}")
    return names + [f"q{g}x{s}" for g, s in qfix_shapes()]


def wide_policy():
    """Kernel policy that runs every job on the wide-band kernel."""
    return 2 + len(ring_shapes())


def ring_policy(waves, slots, lanes=None, uniform=False, dense=False):
    """Kernel policy that forces the ring shape (waves, slots per lane); waves 0
    selects the sub-wave shape with groups of `lanes` lanes (default 16), in its
    per-group form or (uniform=True) its uniform-ticket form, listed second;
    dense=True selects a ring shape's two-per-SIMD form, listed second."""
    lanes = 64 * waves if waves else (lanes or 16)
    hits = [2 + i for i, (w, g, r) in enumerate(ring_shapes()) if w == waves and g == lanes and r == lanes * slots]
    want = 1 if ((uniform and not waves) or (dense and waves)) else 0
    if len(hits) > want:
        return hits[want]
    raise ValueError(f"no ring shape with {waves} waves, {lanes} lanes and {slots} slots "
                     f"(uniform={uniform}, dense={dense})")


class Context:
    """One HIP device + stream (halign_init)."""

    def __init__(self, device=0):
        self.lib = lib()
        h = _P()
        rc = self.lib.halign_init(device, C.byref(h))
        if rc != HALIGN_OK:
            raise HalignError(f"halign_init failed: {ERRORS.get(rc, rc)}")
        self.h = h
        self._children = []   # centers and pair sets, freed before the context

    def close(self):
        if self.h:
            for ch in self._children:
                ch.close()
            self._children = []
            self.lib.halign_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != HALIGN_OK:
            msg = self.lib.halign_last_error(self.h).decode()
            raise HalignError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")

    def tune(self):
        """Measure every ring shape on this device (halign_tune) and plan with it."""
        self._check(self.lib.halign_tune(self.h), "halign_tune")

    def tune_table(self):
        """The planner's cost table: a list of dicts (halign_tune_entry fields)."""
        n = C.c_int32(0)
        self._check(self.lib.halign_tune_table(self.h, None, 0, C.byref(n)), "halign_tune_table")
        arr = (TuneEntry * max(1, n.value))()
        self._check(self.lib.halign_tune_table(self.h, arr, n.value, C.byref(n)), "halign_tune_table")
        out = []
        for i in range(n.value):
            e = arr[i]
            out.append({"waves": e.waves, "lanes": e.lanes, "ring": e.ring, "uniform": e.uniform,
                        "jobs_in_flight": e.jobs_in_flight, "ncol": list(e.ncol[:e.npoints]),
                        "ns_diag": list(e.ns_diag[:e.npoints]), "ns_job": e.ns_job})
        return out

    def set_bucket_timing(self, on):
        """Diagnostic mode: time every DP launch on its own (halign_set_bucket_timing)."""
        self._check(self.lib.halign_set_bucket_timing(self.h, 1 if on else 0), "halign_set_bucket_timing")

    def last_buckets(self):
        """DP launches timed since the last call: a list of dicts (halign_bucket_stat)."""
        n = C.c_int32(0)
        self._check(self.lib.halign_last_buckets(self.h, None, 0, C.byref(n)), "halign_last_buckets")
        arr = (BucketStat * max(1, n.value))()
        self._check(self.lib.halign_last_buckets(self.h, arr, n.value, C.byref(n)), "halign_last_buckets")
        return [{k: getattr(arr[i], k) for k, _ in BucketStat._fields_} for i in range(n.value)]

    def plan_shape(self, n, qlen, tlen, w):
        """Ring shape index the planner gives n equal jobs (-1: another kernel family)."""
        out = C.c_int32(-1)
        self._check(self.lib.halign_plan_shape(self.h, n, qlen, tlen, w, C.byref(out)), "halign_plan_shape")
        return out.value

    def set_kernel_policy(self, policy):
        """0 = automatic, 1 = multi-wave kernels only, 2 + i = ring shape i (see ring_policy)."""
        self._check(self.lib.halign_set_kernel_policy(self.h, int(policy)), "set_kernel_policy")

    def set_graph_mode(self, on):
        """Capture the batched DP launches into a replayed HIP graph (default on)."""
        self._check(self.lib.halign_set_graph_mode(self.h, int(bool(on))), "set_graph_mode")

    def graph_counts(self):
        """(graphs captured, graph replays) on this context."""
        a, b = C.c_int64(0), C.c_int64(0)
        self._check(self.lib.halign_graph_counts(self.h, C.byref(a), C.byref(b)), "graph_counts")
        return a.value, b.value

    def ksw_extz2(self, query_codes, target_codes, w, flag, ez, m=5, mat=DNA5_MAT, q=6, e=2, zdrop=-1,
                  end_bonus=0):
        """halign_ksw_extz2 with ksw_extz2_sse's argument list (include/ksw2.h:64-65):
        fills ez (an ExtzT reused across calls; its cigar grows by realloc and
        is released by free_extz). Returns the status code."""
        qc = np.ascontiguousarray(query_codes, np.uint8)
        tc = np.ascontiguousarray(target_codes, np.uint8)
        return self.lib.halign_ksw_extz2(self.h, None, len(qc), _ptr(qc), len(tc), _ptr(tc), m, mat, q, e, w, zdrop,
                                         end_bonus, flag, C.byref(ez))

    def dp_batch(self, tasks, query_codes, target_codes, full_stats=True, plain=False):
        """tasks: structured array DP_TASK_DTYPE. Returns (results, cigars list, kernel_ms)."""
        tasks = np.ascontiguousarray(tasks, DP_TASK_DTYPE)
        n = len(tasks)
        qc = np.ascontiguousarray(query_codes, np.uint8)
        tc = np.ascontiguousarray(target_codes, np.uint8)
        res = np.zeros(n, DP_RESULT_DTYPE)
        cap = int((tasks["qlen"].astype(np.int64) + tasks["tlen"] + 4).sum()) + 1
        cig = np.zeros(cap, np.uint32)
        ms = C.c_double(0)
        rc = self.lib.halign_dp_batch(self.h, n, _ptr(tasks), _ptr(qc) if len(qc) else None,
                                      _ptr(tc) if len(tc) else None, int(full_stats), int(plain),
                                      _ptr(res), _ptr(cig), cap, C.byref(ms))
        self._check(rc, "halign_dp_batch")
        cigs = [cig[r["cigar_off"]:r["cigar_off"] + r["n_cigar"]].copy() for r in res]
        return res, cigs, ms.value

    def global_align_batch(self, refs, queries, flags):
        """globalAlignKSW2 over ASCII pairs. Returns (list of CIGAR arrays, scores)."""
        n = len(refs)
        rb = [r.encode() if isinstance(r, str) else bytes(r) for r in refs]
        qb = [q.encode() if isinstance(q, str) else bytes(q) for q in queries]
        rbufs = [C.create_string_buffer(x, len(x) + 1) for x in rb]
        qbufs = [C.create_string_buffer(x, len(x) + 1) for x in qb]
        rptr = (C.c_char_p * n)(*[C.cast(b, C.c_char_p) for b in rbufs])
        qptr = (C.c_char_p * n)(*[C.cast(b, C.c_char_p) for b in qbufs])
        rl = np.array([len(x) for x in rb], np.int32)
        ql = np.array([len(x) for x in qb], np.int32)
        fl = np.ascontiguousarray(flags, np.int32)
        cap = int(rl.astype(np.int64).sum() + ql.sum() + 4 * n + 4)
        cig = np.zeros(cap, np.uint32)
        off = np.zeros(n + 1, np.int64)
        sc = np.zeros(n, np.int32)
        rc = self.lib.halign_global_align_batch(self.h, n, C.cast(rptr, _P), _ptr(rl),
                                                C.cast(qptr, _P), _ptr(ql), _ptr(fl), _ptr(cig),
                                                cap, _ptr(off), _ptr(sc))
        self._check(rc, "halign_global_align_batch")
        return [cig[off[i]:off[i + 1]].copy() for i in range(n)], sc


    def consensus(self, rows, batch=4096, seq_limit=0, device_ptr=None, row_stride=None):
        """Column counts (aln_len x 7: A C G T U N -), consensus bytes and the
        number of rows counted (generateConsensusSequence's counting core).
        rows: (nrows, aln_len) uint8 host array, or its shape when device_ptr
        (an integer device address) holds the rows with row_stride bytes."""
        if device_ptr is None:
            rows = np.ascontiguousarray(rows, np.uint8)
            nrows, aln = rows.shape
            ptr, stride, on_dev = _ptr(rows), aln, 0
        else:
            nrows, aln = rows
            ptr, stride, on_dev = C.c_void_p(device_ptr), row_stride, 1
        counts = np.zeros((aln, 7), np.uint32)
        cons = np.zeros(aln, np.uint8)
        n = C.c_uint64(0)
        ms = C.c_double(0)
        rc = self.lib.halign_consensus(self.h, ptr, nrows, aln, stride, on_dev, batch, seq_limit,
                                       _ptr(counts), _ptr(cons), C.byref(n), C.byref(ms))
        self._check(rc, "halign_consensus")
        return counts, cons.tobytes(), n.value, ms.value


    def mash_sketch(self, seqs, k=15, size=2000, non_canonical=True, seed=42):
        """mash::sketchFromSequence per sequence: list of sorted uint64 arrays."""
        blob, off = _blob(seqs)
        n = len(off) - 1
        cap = max(1, n * max(size, 0))
        out = np.zeros(cap, np.uint64)
        soff = np.zeros(n + 1, np.int64)
        self._check(self.lib.halign_mash_sketch(self.h, n, blob, _ptr(off), k, size, int(non_canonical), seed,
                                                _ptr(out), cap, _ptr(soff)), "halign_mash_sketch")
        return [out[soff[i]:soff[i + 1]].copy() for i in range(n)]

    def mash_best(self, query_sketches, ref_sketches):
        """(best reference index, Jaccard) per query sketch."""
        def flat(sk):
            off = np.zeros(len(sk) + 1, np.int64)
            off[1:] = np.cumsum([len(x) for x in sk])
            data = np.concatenate([np.asarray(x, np.uint64) for x in sk]) if off[-1] else np.zeros(1, np.uint64)
            return np.ascontiguousarray(data, np.uint64), off
        qd, qo = flat(query_sketches)
        rd, ro = flat(ref_sketches)
        nq = len(query_sketches)
        best = np.zeros(max(nq, 1), np.int32)
        jac = np.zeros(max(nq, 1), np.float64)
        self._check(self.lib.halign_mash_best(self.h, nq, _ptr(qd), _ptr(qo), len(ref_sketches), _ptr(rd), _ptr(ro),
                                              _ptr(best), _ptr(jac)), "halign_mash_best")
        return best[:nq], jac[:nq]

    def set_seed_policy(self, policy):
        """0 = minimizer wave form where it applies, 1 = sequential form only."""
        self._check(self.lib.halign_set_seed_policy(self.h, int(policy)), "set_seed_policy")
        self._seed_policy = int(policy)

    def extract_minimizers(self, seqs, k=15, w=10, non_canonical=True):
        """extractMinimizer over a list of byte strings. Returns a list of
        (x, y) uint64 arrays (MinimizerHit records) and the per-sequence status."""
        n = len(seqs)
        blob = b"".join(seqs)
        off = np.zeros(n + 1, np.int64)
        off[1:] = np.cumsum([len(x) for x in seqs])
        cap = max(1, int(sum(max(len(x) - k + 1, 0) for x in seqs)))
        hx = np.zeros(cap, np.uint64)
        hy = np.zeros(cap, np.uint64)
        hoff = np.zeros(n + 1, np.int64)
        st = np.zeros(max(n, 1), np.int32)
        rc = self.lib.halign_extract_minimizers(self.h, n, blob, _ptr(off), k, w, int(non_canonical),
                                                _ptr(hx), _ptr(hy), cap, _ptr(hoff), _ptr(st))
        self._check(rc, "halign_extract_minimizers")
        return [(hx[hoff[i]:hoff[i + 1]].copy(), hy[hoff[i]:hoff[i + 1]].copy()) for i in range(n)], st[:n]


ANCHOR_DTYPE = np.dtype([("hash", "<u8"), ("rid_ref", "<u4"), ("pos_ref", "<u4"),
                         ("rid_qry", "<u4"), ("pos_qry", "<u4"), ("span", "<u4"), ("is_rev", "<u4")])


def _blob(queries):
    qs = [q.encode() if isinstance(q, str) else bytes(q) for q in queries]
    off = np.zeros(len(qs) + 1, np.int64)
    off[1:] = np.cumsum([len(q) for q in qs])
    return b"".join(qs), off


class Center:
    """A resident center for the per-query path (halign_center_*)."""

    def __init__(self, ctx, center, index_seq=None):
        """index_seq: build the minimizer index from this sequence instead
        (halign_center_create_indexed)."""
        self.ctx = ctx
        center = center.encode() if isinstance(center, str) else bytes(center)
        self.rlen = len(center)
        h = _P()
        if index_seq is None:
            ctx._check(ctx.lib.halign_center_create(ctx.h, center, len(center), C.byref(h)),
                       "halign_center_create")
        else:
            index_seq = index_seq.encode() if isinstance(index_seq, str) else bytes(index_seq)
            ctx._check(ctx.lib.halign_center_create_indexed(ctx.h, center, len(center), index_seq,
                                                            len(index_seq), C.byref(h)),
                       "halign_center_create_indexed")
        self.h = h
        ctx._children.append(self)

    def close(self):
        if self.h:
            if self.ctx.h:
                self.ctx.lib.halign_center_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def anchors(self, queries):
        blob, off = _blob(queries)
        cap = max(16, 64 * (len(blob) + 1))
        out = np.zeros(cap, ANCHOR_DTYPE)
        aoff = np.zeros(len(off), np.int64)
        self.ctx._check(self.ctx.lib.halign_center_anchors(self.h, len(off) - 1, blob, _ptr(off),
                                                           _ptr(out), cap, _ptr(aoff)),
                        "halign_center_anchors")
        return [out[aoff[i]:aoff[i + 1]].copy() for i in range(len(off) - 1)]

    def chains(self, queries):
        blob, off = _blob(queries)
        cap = max(16, len(blob) + 1)
        out = np.zeros(cap, ANCHOR_DTYPE)
        coff = np.zeros(len(off), np.int64)
        self.ctx._check(self.ctx.lib.halign_center_chains(self.h, len(off) - 1, blob, _ptr(off),
                                                          _ptr(out), cap, _ptr(coff)),
                        "halign_center_chains")
        return [out[coff[i]:coff[i + 1]].copy() for i in range(len(off) - 1)]

    def align(self, queries):
        """RefAligner::globalAlign per query: (list of CIGAR arrays, stats[n,4], ms)."""
        blob, off = _blob(queries)
        cig, coff, st, ms = self.align_flat(blob, off)
        return [cig[coff[i]:coff[i + 1]] for i in range(len(off) - 1)], st, ms

    def align_rows(self, queries, row_stride=None):
        """RefAligner::globalAlign plus the keep-length rows (delQueryToRefByCigar)
        and hasInsertion per query: (list of CIGAR arrays, list of row bytes,
        has_insertion bool array)."""
        blob, off = _blob(queries)
        off = np.ascontiguousarray(off, np.int64)
        n = len(off) - 1
        stride = row_stride or max(self.rlen, 1)
        cap = int(off[-1] - off[0]) + n * (self.rlen + 1) + 1
        cig = np.empty(cap, np.uint32)
        coff = np.zeros(n + 1, np.int64)
        rows = np.zeros((max(n, 1), stride), np.uint8)
        rlen = np.zeros(max(n, 1), np.int64)
        hins = np.zeros(max(n, 1), np.uint8)
        st = np.zeros((max(n, 1), 4), np.int32)
        ms = C.c_double(0)
        self.ctx._check(self.ctx.lib.halign_center_align_rows(
            self.h, n, blob, _ptr(off), _ptr(cig), cap, _ptr(coff), _ptr(rows), stride, _ptr(rlen),
            _ptr(hins), _ptr(st), C.byref(ms)), "halign_center_align_rows")
        cigs = [cig[coff[i]:coff[i + 1]].copy() for i in range(n)]
        rws = [rows[i, :min(int(rlen[i]), stride)].tobytes() for i in range(n)]
        return cigs, rws, hins[:n].astype(bool)

    def align_fasta(self, queries, names):
        """The keep-length MSA records of the batch (halign_center_align_fasta):
        (record bytes, bytes of the insertion group, queries in it, ms)."""
        blob, off = _blob(queries)
        nb, noff = _blob(names)
        n = len(off) - 1
        cap = int(noff[-1]) + n * (self.rlen + 3) + 1
        out = np.empty(cap, np.uint8)
        out_len, ins, nins, ms = C.c_int64(0), C.c_int64(0), C.c_int32(0), C.c_double(0)
        self.ctx._check(self.ctx.lib.halign_center_align_fasta(
            self.h, n, blob, _ptr(off), nb, _ptr(noff), _ptr(out), cap, C.byref(out_len), C.byref(ins),
            C.byref(nins), C.byref(ms)), "halign_center_align_fasta")
        return out[:out_len.value].tobytes(), ins.value, nins.value, ms.value

    def align_flat(self, blob, off):
        """The same on concatenated queries (bytes or uint8 array, int64
        offsets): (flat CIGAR array, CIGAR offsets, stats[n,4], ms)."""
        keep = None
        if not isinstance(blob, (bytes, bytearray)):   # a uint8 array is passed in place
            keep = np.ascontiguousarray(blob, np.uint8)
            blob = C.cast(_ptr(keep), C.c_char_p)
        off = np.ascontiguousarray(off, np.int64)
        n = len(off) - 1
        cap = int(off[-1] - off[0]) + n * (self.rlen + 1) + 1
        cig = np.empty(cap, np.uint32)   # only the first coff[n] words are written
        coff = np.zeros(n + 1, np.int64)
        st = np.zeros((max(n, 1), 4), np.int32)
        ms = C.c_double(0)
        self.ctx._check(self.ctx.lib.halign_center_align(self.h, n, blob, _ptr(off), _ptr(cig), cap,
                                                         _ptr(coff), _ptr(st), C.byref(ms)),
                        "halign_center_align")
        return cig, coff, st[:n], ms.value


class RefSet:
    """RefAligner's reference set (halign_refset_*): reference records and the
    consensus of their MSA."""

    def __init__(self, ctx, refs, consensus):
        self.ctx = ctx
        rb = [r.encode() if isinstance(r, str) else bytes(r) for r in refs]
        self._keep = [C.create_string_buffer(r, len(r) + 1) for r in rb]
        ptrs = (C.c_char_p * len(rb))(*[C.cast(b, C.c_char_p) for b in self._keep])
        lens = np.array([len(r) for r in rb], np.int32)
        cons = consensus.encode() if isinstance(consensus, str) else bytes(consensus)
        self.maxlen = max([len(cons)] + [len(r) for r in rb])
        h = _P()
        ctx._check(ctx.lib.halign_refset_create(ctx.h, len(rb), C.cast(ptrs, _P), _ptr(lens), cons, len(cons),
                                                C.byref(h)), "halign_refset_create")
        self.h = h

    def close(self):
        if self.h:
            self.ctx.lib.halign_refset_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def align(self, queries, keep_length):
        """(best_ref, first CIGARs, result CIGARs, has_insertion, result reference (-1: consensus))."""
        blob, off = _blob(queries)
        n = len(off) - 1
        cap = int(off[-1]) + n * (self.maxlen + 1) + 1
        ic, fc = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
        io, fo = np.zeros(n + 1, np.int64), np.zeros(n + 1, np.int64)
        best, fr = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32)
        hi = np.zeros(max(n, 1), np.uint8)
        self.ctx._check(self.ctx.lib.halign_refset_align(self.h, n, blob, _ptr(off), int(keep_length), _ptr(best),
                                                         _ptr(ic), cap, _ptr(io), _ptr(fc), cap, _ptr(fo), _ptr(hi),
                                                         _ptr(fr)), "halign_refset_align")
        return (best[:n].copy(), [ic[io[i]:io[i + 1]].copy() for i in range(n)],
                [fc[fo[i]:fo[i + 1]].copy() for i in range(n)], hi[:n].astype(bool), fr[:n].copy())


def align_sam(ctx, center_path, input_path, sam_path, insertion_sam_path, keep_length=True, batch=0,
              aligned_path=None):
    """alignQueryToRef's two SAM files with one thread (halign_align_sam)."""
    st = MsaStats()
    ctx._check(ctx.lib.halign_align_sam(ctx.h, center_path.encode(), aligned_path.encode() if aligned_path else None,
                                        input_path.encode(), sam_path.encode(), insertion_sam_path.encode(),
                                        int(keep_length), int(batch), C.byref(st)), "halign_align_sam")
    return {f: getattr(st, f) for f, _ in MsaStats._fields_}


class PairSet:
    """Center + queries resident in HBM (halign_pairset_*)."""

    def __init__(self, ctx, center, queries_concat, offsets):
        self.ctx = ctx
        center = center.encode() if isinstance(center, str) else bytes(center)
        qc = queries_concat if isinstance(queries_concat, (bytes, bytearray)) else bytes(queries_concat)
        off = np.ascontiguousarray(offsets, np.int64)
        h = _P()
        rc = ctx.lib.halign_pairset_create(ctx.h, center, len(center), len(off) - 1, qc, _ptr(off),
                                           C.byref(h))
        ctx._check(rc, "halign_pairset_create")
        self.h = h
        ctx._children.append(self)
        self.n = len(off) - 1

    def run(self, first, count, plain=False):
        ms = C.c_double(0)
        cells = C.c_int64(0)
        rc = self.ctx.lib.halign_pairset_run(self.h, first, count, int(plain), C.byref(ms),
                                             C.byref(cells))
        self.ctx._check(rc, "halign_pairset_run")
        return ms.value, cells.value

    def result(self, idx):
        sc = C.c_int32(0)
        nc = C.c_int32(0)
        self.ctx._check(self.ctx.lib.halign_pairset_result(self.h, idx, C.byref(sc), None, 0,
                                                            C.byref(nc)), "pairset_result")
        cig = np.zeros(max(1, nc.value), np.uint32)
        self.ctx._check(self.ctx.lib.halign_pairset_result(self.h, idx, C.byref(sc), _ptr(cig),
                                                            len(cig), C.byref(nc)),
                        "pairset_result")
        return sc.value, cig[:nc.value].copy()

    def close(self):
        if self.h:
            if self.ctx.h:
                self.ctx.lib.halign_pairset_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server over a resident center (halign_server_*):
    submit() queues a request of queries, wait()/result() collect it."""

    def __init__(self, center, device=0, workers=2, max_batch_queries=4096, hbm_budget=0, chunk_queries=0):
        self.lib = lib()
        center = center.encode() if isinstance(center, str) else bytes(center)
        cfg = ServerConfig(workers, max_batch_queries, hbm_budget, chunk_queries, 0)
        h = _P()
        rc = self.lib.halign_server_create(device, center, len(center), C.byref(cfg), C.byref(h))
        if rc != HALIGN_OK:
            raise HalignError(f"halign_server_create failed: {ERRORS.get(rc, rc)}")
        self.h = h
        self._nq = {}

    def close(self):
        if self.h:
            self.lib.halign_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != HALIGN_OK:
            msg = self.lib.halign_server_last_error(self.h).decode()
            raise HalignError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")

    def stats(self):
        """Device memory of the server, estimated and measured: a dict of the
        halign_server_memory fields (bytes; batches run)."""
        m = ServerMemory()
        self._check(self.lib.halign_server_stats(self.h, C.byref(m)), "halign_server_stats")
        return {k: getattr(m, k) for k, _ in ServerMemory._fields_}

    def submit_flat(self, blob, off):
        """Queue queries blob[off[i]:off[i+1]] (bytes or uint8 array, int64 offsets)."""
        keep = None
        if not isinstance(blob, (bytes, bytearray)):
            keep = np.ascontiguousarray(blob, np.uint8)
            blob = C.cast(_ptr(keep), C.c_char_p)
        off = np.ascontiguousarray(off, np.int64)
        rid = C.c_int64(0)
        self._check(self.lib.halign_server_submit(self.h, len(off) - 1, blob, _ptr(off), C.byref(rid)),
                    "halign_server_submit")
        self._nq[rid.value] = len(off) - 1
        return rid.value

    def submit(self, queries):
        qb = [q.encode() if isinstance(q, str) else bytes(q) for q in queries]
        off = np.zeros(len(qb) + 1, np.int64)
        off[1:] = np.cumsum([len(q) for q in qb])
        return self.submit_flat(b"".join(qb), off)

    def wait(self, rid, timeout_ms=-1):
        """True once the request is done (False if timeout_ms passed first)."""
        rc = self.lib.halign_server_wait(self.h, rid, int(timeout_ms))
        if rc == HALIGN_PENDING:
            return False
        self._check(rc, "halign_server_wait")
        return True

    def result(self, rid):
        """(flat CIGAR array, CIGAR offsets, stats[n,4], (queue ms, total ms, batch
        queries)); waits for the request and releases it."""
        self.wait(rid)
        n = C.c_int64(0)
        timing = np.zeros(3, np.float64)
        self._check(self.lib.halign_server_result(self.h, rid, None, 0, None, None, _ptr(timing),
                                                  C.byref(n)), "halign_server_result")
        nq = self._nq.pop(rid)
        cig = np.zeros(max(1, n.value), np.uint32)
        coff = np.zeros(nq + 1, np.int64)
        st = np.zeros((max(1, nq), 4), np.int32)
        self._check(self.lib.halign_server_result(self.h, rid, _ptr(cig), len(cig), _ptr(coff), _ptr(st),
                                                  _ptr(timing), C.byref(n)), "halign_server_result")
        return cig[:n.value], coff, st[:nq], tuple(float(x) for x in timing)


def read_fastx(path, chunk=4096):
    """Yield (id, desc, seq) bytes triples of a FASTA/FASTQ file, gzip or
    plain (halign_fastx_*, seq_io::KseqReader's rules)."""
    L = lib()
    h = _P()
    rc = L.halign_fastx_open(path.encode(), C.byref(h))
    if rc != HALIGN_OK:
        raise HalignError(f"halign_fastx_open: {ERRORS.get(rc, rc)}: {path}")
    try:
        while True:
            n = C.c_int32(0)
            ptrs = [_P() for _ in range(6)]
            rc = L.halign_fastx_next(h, chunk, C.byref(n), *[C.byref(p) for p in ptrs])
            if rc != HALIGN_OK:
                raise HalignError(f"halign_fastx_next: {ERRORS.get(rc, rc)}: {L.halign_fastx_error(h).decode()}")
            if n.value == 0:
                return
            seqs, soff, ids, ioff, descs, doff = ptrs
            so = np.ctypeslib.as_array(C.cast(soff, C.POINTER(C.c_int64)), (n.value + 1,)).copy()
            io = np.ctypeslib.as_array(C.cast(ioff, C.POINTER(C.c_int64)), (n.value + 1,)).copy()
            do = np.ctypeslib.as_array(C.cast(doff, C.POINTER(C.c_int64)), (n.value + 1,)).copy()
            sb = C.string_at(seqs, int(so[-1])) if so[-1] else b""
            ib = C.string_at(ids, int(io[-1])) if io[-1] else b""
            db = C.string_at(descs, int(do[-1])) if do[-1] else b""
            for i in range(n.value):
                yield ib[io[i]:io[i + 1]], db[do[i]:do[i + 1]], sb[so[i]:so[i + 1]]
    finally:
        L.halign_fastx_close(h)


def msa_keep_length(ctx, center_path, input_path, output_path, batch=0, aligned_path=None):
    """halign4 -c center --keep-length, file to file (halign_msa_keep_length,
    or halign_msa_keep_length_refs with the MSA command's output for a center
    file of several references): returns the run's statistics as a dict."""
    st = MsaStats()
    if aligned_path is None:
        rc = ctx.lib.halign_msa_keep_length(ctx.h, center_path.encode(), input_path.encode(),
                                            output_path.encode(), int(batch), C.byref(st))
    else:
        rc = ctx.lib.halign_msa_keep_length_refs(ctx.h, center_path.encode(), aligned_path.encode(),
                                                 input_path.encode(), output_path.encode(), int(batch), C.byref(st))
    ctx._check(rc, "halign_msa_keep_length")
    return {f: getattr(st, f) for f, _ in MsaStats._fields_}


def topk_longest(input_path, k, output_path, ctx=None):
    """preprocessInputFasta's Top-K consensus input (halign_topk_longest, host
    only): returns the number of records read."""
    L = lib()
    n = C.c_int64(0)
    rc = L.halign_topk_longest(ctx.h if ctx else None, input_path.encode(), int(k), output_path.encode(),
                               C.byref(n))
    if rc != HALIGN_OK:
        msg = L.halign_last_error(ctx.h).decode() if ctx else ""
        raise HalignError(f"halign_topk_longest: {ERRORS.get(rc, rc)} {msg}")
    return n.value


def consensus_file(ctx, aligned_path, out_fasta, out_json, seq_limit=0, batch=4096):
    """generateConsensusSequence on a file (halign_consensus_file): the consensus bytes."""
    n = C.c_int64(0)
    cap = max(16, os.path.getsize(aligned_path))
    buf = C.create_string_buffer(cap)
    ctx._check(ctx.lib.halign_consensus_file(ctx.h, aligned_path.encode(), out_fasta.encode(), out_json.encode(),
                                             int(seq_limit), int(batch), buf, cap, C.byref(n)),
               "halign_consensus_file")
    return buf.raw[:n.value]


def msa_auto_center(ctx, input_path, output_path, cons_n=1000, aligned_topk_path=None, batch=0):
    """halign4 --keep-length without a center (halign_msa_auto_center)."""
    st = MsaStats()
    ctx._check(ctx.lib.halign_msa_auto_center(ctx.h, input_path.encode(),
                                              aligned_topk_path.encode() if aligned_topk_path else None,
                                              output_path.encode(), int(cons_n), int(batch), C.byref(st)),
               "halign_msa_auto_center")
    return {f: getattr(st, f) for f, _ in MsaStats._fields_}
