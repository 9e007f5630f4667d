"""Python host binding of the MI355X batched Smith-Waterman engine (libgbsw.so).

Mirrors the reference's driver entry points (driver.hpp:32-37):
``kernel_driver_dna(reads, contigs, scores)`` and
``kernel_driver_aa(reads, contigs, scoring_matrix, open_gap, extend_gap)`` return
the five ``alignment_results`` fields (driver.hpp:21-27) as int32 numpy arrays.
Everything goes through the C ABI in include/gbsw.h; there is no fallback: if
the shared library is missing or no GPU is present, calls raise.
"""
import ctypes
import os

import numpy as np

from .scoring import AA_ORDER, BLOSUM62, encode_aa  # noqa: F401

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)  # gpu-bsw_amd/
LIB_PATH = os.path.join(ROOT_DIR, "libgbsw.so")

DNA, PROTEIN = 0, 1
PLAIN, BATCHED = 0, 1
MODES = {"plain": PLAIN, "batched": BATCHED}
FIELDS = ("score", "ref_begin", "ref_end", "query_begin", "query_end")

GBSW_OK = 0
ERRORS = {-1: "ARGUMENT", -2: "RESIDUE", -3: "SCORING", -4: "DEVICE", -5: "MEMORY", -6: "RANGE", -7: "INTERNAL",
          -8: "TIMEOUT"}
E_RESIDUE, E_RANGE, E_TIMEOUT = -2, -6, -8


class GbswError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gbsw error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code


class Scoring(ctypes.Structure):
    _fields_ = [
        ("alphabet", ctypes.c_int32),
        ("match", ctypes.c_int16),
        ("mismatch", ctypes.c_int16),
        ("gap_open", ctypes.c_int16),
        ("gap_extend", ctypes.c_int16),
        ("matrix", ctypes.c_int16 * 576),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("device_ms", ctypes.c_double),
        ("main_kernel_ms", ctypes.c_double),
        ("cells", ctypes.c_int64),
        ("launches", ctypes.c_int32),
        ("buckets", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class LaunchStat(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("columns_per_lane", ctypes.c_int32),
        ("workgroups", ctypes.c_int32),
        ("pairs", ctypes.c_int64),
        ("cells", ctypes.c_int64),
        ("ms", ctypes.c_double),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_int32),
        ("pairs_per_slot", ctypes.c_int32),
        ("residues_per_slot", ctypes.c_int64),
        ("boundary_bytes_per_slot", ctypes.c_int64),
        ("staging_bytes", ctypes.c_int64),
        ("dna_alphabet", ctypes.c_char * 32),
        ("device_staging_bytes", ctypes.c_int64),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("ticket", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("parts", ctypes.c_int32),
        ("pairs", ctypes.c_int64),
        ("cells", ctypes.c_int64),
        ("queue_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
    ]


class ServingStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_int64),
        ("completed", ctypes.c_int64),
        ("pending", ctypes.c_int64),
        ("micro_batches", ctypes.c_int64),
        ("pairs", ctypes.c_int64),
        ("cells", ctypes.c_int64),
        ("busy_ms", ctypes.c_double),
        ("device_bytes", ctypes.c_int64),
    ]


# Every symbol include/gbsw.h declares (checked by the CPU test suite).
EXPORTED = (
    "gbsw_context_create", "gbsw_context_destroy", "gbsw_last_error", "gbsw_batch_create",
    "gbsw_batch_destroy", "gbsw_batch_align", "gbsw_batch_results", "gbsw_batch_device_results",
    "gbsw_batch_stats", "gbsw_batch_launch_stats", "gbsw_batch_autotune", "gbsw_align", "gbsw_device_memory",
    "gbsw_version",
    "gbsw_server_config_default", "gbsw_server_create", "gbsw_server_destroy", "gbsw_server_submit",
    "gbsw_server_wait", "gbsw_server_poll", "gbsw_server_stats", "gbsw_host_alloc", "gbsw_host_free",
    "gbsw_context_set_option", "gbsw_batch_traceback", "gbsw_align_noalloc",
)
OPT_STREAMS, OPT_F16, OPT_GEOMETRY, OPT_SPLIT, OPT_GRAPHS = 1, 2, 3, 4, 5
# launch kinds (gbsw_launch_stat.kind): plain forward, plain reverse, batched
# int16, batched int32, batched f16; plus KIND_BYTE_PROFILE for int8 profiles
KIND_PLAIN_FORWARD, KIND_PLAIN_REVERSE, KIND_BATCHED, KIND_BATCHED_WIDE, KIND_BATCHED_F16 = 0, 1, 2, 3, 4
KIND_SPLIT, KIND_SPLIT_F16 = 5, 6
KIND_BYTE_PROFILE = 256

_lib = None


def library_path():
    return os.environ.get("GBSW_LIBRARY", LIB_PATH)


def load_library():
    """Load libgbsw.so (built in-tree by __graft_entry__.build() / make -C gpu-bsw_amd)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise FileNotFoundError("libgbsw.so not built: %s (run `make -C gpu-bsw_amd`)" % path)
    lib = ctypes.CDLL(path)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    lib.gbsw_context_create.argtypes = [i32, ctypes.POINTER(vp)]
    lib.gbsw_context_destroy.argtypes = [vp]
    lib.gbsw_context_set_option.argtypes = [vp, i32, i64]
    lib.gbsw_last_error.restype = ctypes.c_char_p
    lib.gbsw_version.restype = ctypes.c_char_p
    lib.gbsw_batch_create.argtypes = [vp, i32, i64, vp, vp, vp, vp, ctypes.POINTER(vp)]
    lib.gbsw_batch_destroy.argtypes = [vp]
    lib.gbsw_batch_align.argtypes = [vp, ctypes.POINTER(Scoring), i32]
    lib.gbsw_batch_results.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.gbsw_batch_device_results.argtypes = [vp, ctypes.POINTER(vp)]
    lib.gbsw_batch_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    lib.gbsw_batch_launch_stats.argtypes = [vp, i32, ctypes.POINTER(LaunchStat)]
    lib.gbsw_batch_traceback.argtypes = [vp, ctypes.POINTER(Scoring), i64, vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
    lib.gbsw_align.argtypes = [vp, ctypes.POINTER(Scoring), i32, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.gbsw_align_noalloc.argtypes = [vp, ctypes.POINTER(Scoring), i32, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.gbsw_device_memory.argtypes = [i32, ctypes.POINTER(ctypes.c_uint64)]
    u64 = ctypes.c_uint64
    lib.gbsw_server_config_default.argtypes = [ctypes.POINTER(ServerConfig)]
    lib.gbsw_server_config_default.restype = None
    lib.gbsw_server_create.argtypes = [i32, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    lib.gbsw_server_destroy.argtypes = [vp]
    lib.gbsw_server_submit.argtypes = [vp, ctypes.POINTER(Scoring), i64, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                       ctypes.POINTER(u64)]
    lib.gbsw_server_wait.argtypes = [vp, u64, ctypes.c_double, ctypes.POINTER(RequestInfo)]
    lib.gbsw_server_poll.argtypes = [vp, i32, ctypes.POINTER(RequestInfo), ctypes.POINTER(i32)]
    lib.gbsw_server_stats.argtypes = [vp, ctypes.POINTER(ServingStats)]
    lib.gbsw_host_alloc.argtypes = [u64, ctypes.POINTER(vp)]
    lib.gbsw_host_free.argtypes = [vp]
    for name in EXPORTED:
        if name not in ("gbsw_last_error", "gbsw_version", "gbsw_server_config_default"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _check(rc):
    if rc != GBSW_OK:
        raise GbswError(rc, load_library().gbsw_last_error().decode(errors="replace"))


def dna_scoring(match=3, mismatch=-3, gap_open=-6, gap_extend=-1):
    sc = Scoring()
    sc.alphabet, sc.match, sc.mismatch, sc.gap_open, sc.gap_extend = DNA, match, mismatch, gap_open, gap_extend
    return sc


def protein_scoring(matrix=None, gap_open=-6, gap_extend=-1):
    sc = Scoring()
    sc.alphabet, sc.gap_open, sc.gap_extend = PROTEIN, gap_open, gap_extend
    m = BLOSUM62 if matrix is None else list(matrix)
    if len(m) != 576:
        raise ValueError("protein scoring matrix needs 576 entries")
    for k, v in enumerate(m):
        sc.matrix[k] = int(v)
    return sc


def pack(seqs):
    """List of str/bytes -> (uint8 buffer, uint64 offsets of n+1)."""
    bs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    off = np.zeros(len(bs) + 1, dtype=np.uint64)
    if bs:
        off[1:] = np.cumsum([len(b) for b in bs], dtype=np.uint64)
    buf = np.frombuffer(b"".join(bs), dtype=np.uint8) if off[-1] else np.zeros(1, dtype=np.uint8)
    return np.ascontiguousarray(buf), off


class Context:
    """One HIP device, one stream (gbsw_context)."""

    def __init__(self, device=0):
        self._lib = load_library()
        self._h = ctypes.c_void_p()
        _check(self._lib.gbsw_context_create(device, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            self._lib.gbsw_context_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, option, value):
        """OPT_STREAMS (1..8), OPT_F16, OPT_GEOMETRY (0..5), OPT_SPLIT or OPT_GRAPHS (0/1); see include/gbsw.h."""
        _check(self._lib.gbsw_context_set_option(self._h, option, int(value)))

    def batch(self, alphabet, ref, ref_off, query, query_off):
        return Batch(self, alphabet, ref, ref_off, query, query_off)

    def align(self, refs, queries, scoring, mode="batched"):
        """Host lists in, dict of five int32 arrays out (gbsw_align)."""
        rb, ro = pack(refs)
        qb, qo = pack(queries)
        return self.align_packed(rb, ro, qb, qo, scoring, mode)

    def align_packed(self, rb, ro, qb, qo, scoring, mode="batched", noalloc=False):
        """Packed buffers and n+1 offsets in, dict of five int32 arrays out
        (gbsw_align, or gbsw_align_noalloc when noalloc is set)."""
        n = len(ro) - 1
        if len(qo) - 1 != n:
            raise ValueError("ref and query counts differ")
        out = np.zeros((5, max(n, 1)), dtype=np.int32)
        ptrs = [out[f].ctypes.data for f in range(5)]
        rb = np.ascontiguousarray(rb, dtype=np.uint8)
        qb = np.ascontiguousarray(qb, dtype=np.uint8)
        ro = np.ascontiguousarray(ro, dtype=np.uint64)
        qo = np.ascontiguousarray(qo, dtype=np.uint64)
        fn = self._lib.gbsw_align_noalloc if noalloc else self._lib.gbsw_align
        _check(fn(self._h, ctypes.byref(scoring), MODES.get(mode, mode), n, rb.ctypes.data, ro.ctypes.data,
                  qb.ctypes.data, qo.ctypes.data, *ptrs))
        return {f: out[k][:n].copy() for k, f in enumerate(FIELDS)}


class Batch:
    """A batch resident in device memory (gbsw_batch)."""

    def __init__(self, ctx, alphabet, ref, ref_off, query, query_off):
        self._ctx = ctx
        self._lib = ctx._lib
        self._h = ctypes.c_void_p()
        self._keep = [np.ascontiguousarray(x) for x in (ref, ref_off, query, query_off)]
        rb, ro, qb, qo = self._keep
        assert ro.dtype == np.uint64 and qo.dtype == np.uint64
        self.n = len(ro) - 1
        _check(self._lib.gbsw_batch_create(ctx._h, alphabet, self.n, rb.ctypes.data, ro.ctypes.data,
                                           qb.ctypes.data, qo.ctypes.data, ctypes.byref(self._h)))
        self._keep = None

    def align(self, scoring, mode="batched"):
        _check(self._lib.gbsw_batch_align(self._h, ctypes.byref(scoring), MODES.get(mode, mode)))

    def results(self):
        out = np.zeros((5, max(self.n, 1)), dtype=np.int32)
        _check(self._lib.gbsw_batch_results(self._h, *[out[f].ctypes.data for f in range(5)]))
        return {f: out[k][:self.n].copy() for k, f in enumerate(FIELDS)}

    def autotune(self, scoring, repeats=2):
        """gbsw_batch_autotune: (chosen geometry model, its device ms)."""
        m = ctypes.c_int32()
        ms = ctypes.c_double()
        _check(self._lib.gbsw_batch_autotune(self._h, ctypes.byref(scoring), int(repeats), ctypes.byref(m),
                                             ctypes.byref(ms)))
        return m.value, ms.value

    def traceback(self, scoring, capacity=None):
        """gbsw_batch_traceback: alignment paths of the last align.

        Returns (words, offsets, path_score, device_ms): pair k's BAM CIGAR words
        (length << 4 | op; op 0 = M, 1 = I, 2 = D) are words[offsets[k]:offsets[k+1]].
        """
        off = np.zeros(self.n + 1, dtype=np.int64)
        ps = np.zeros(max(self.n, 1), dtype=np.int32)
        ms = ctypes.c_double()
        if capacity is None:
            # rows + columns of every rectangle bound its runs
            r = self.results()
            capacity = int(np.sum(np.where(r["score"] > 0, (r["ref_end"] - r["ref_begin"]).astype(np.int64) +
                                           (r["query_end"] - r["query_begin"]), 0)))
        words = np.zeros(max(1, capacity), dtype=np.uint32)
        _check(self._lib.gbsw_batch_traceback(self._h, ctypes.byref(scoring), int(capacity), words.ctypes.data,
                                              off.ctypes.data, ps.ctypes.data, ctypes.byref(ms)))
        return words[:off[-1]].copy(), off, ps[:self.n].copy(), ms.value

    def stats(self):
        st = Stats()
        _check(self._lib.gbsw_batch_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def launches(self):
        """Per-launch statistics of the last align: kind, K, workgroups, pairs, cells, ms."""
        out = []
        for i in range(self.stats()["launches"]):
            ls = LaunchStat()
            _check(self._lib.gbsw_batch_launch_stats(self._h, i, ctypes.byref(ls)))
            out.append({k: getattr(ls, k) for k, _ in LaunchStat._fields_})
        return out

    def close(self):
        if self._h:
            self._lib.gbsw_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(int(os.environ.get("GBSW_DEVICE", "0")))
    return _default_ctx


def kernel_driver_dna(reads, contigs, scores, factor=1.0, mode="batched"):
    """driver.hpp:32-33.  reads[k] (query) is aligned with contigs[k] (ref);
    scores = (match, mismatch, gap_open, gap_extend)."""
    del factor  # memory fraction of the reference's chunking; batches are resident whole
    n = len(contigs)
    sc = dna_scoring(*[int(v) for v in scores])
    return default_context().align(list(contigs), list(reads)[:n], sc, mode)


def kernel_driver_aa(reads, contigs, scoring_matrix, open_gap, extend_gap, factor=1.0, mode="batched"):
    """driver.hpp:36-37 with the 24x24 matrix in ARNDCQEGHILKMFPSTWYVBZX* order."""
    del factor
    n = len(contigs)
    sc = protein_scoring(scoring_matrix, int(open_gap), int(extend_gap))
    return default_context().align(list(contigs), list(reads)[:n], sc, mode)


# ---------------------------------------------------------------------------
# Serving (gbsw_server_*): asynchronous requests with continuous batching.

class HostBuffer:
    """Pinned host memory from gbsw_host_alloc, viewed as a numpy uint8 array.
    Requests whose sequences lie in such a buffer are read by the device in place
    (no staging copy); the buffer must then stay unchanged until they complete."""

    def __init__(self, nbytes):
        self._lib = load_library()
        self._p = ctypes.c_void_p()
        _check(self._lib.gbsw_host_alloc(max(int(nbytes), 1), ctypes.byref(self._p)))
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * max(int(nbytes), 1)).from_address(self._p.value))

    def close(self):
        if self._p:
            self.array = None
            self._lib.gbsw_host_free(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def server_config(**kw):
    cfg = ServerConfig()
    load_library().gbsw_server_config_default(ctypes.byref(cfg))
    for k, v in kw.items():
        if k == "dna_alphabet":
            v = v.encode() if isinstance(v, str) else bytes(v)
        setattr(cfg, k, v)
    return cfg


def _info(ri):
    return {k: getattr(ri, k) for k, _ in RequestInfo._fields_}


class Server:
    """gbsw_server: submit() returns a ticket at once; wait()/poll() report
    completed requests.  Arrays handed to submit() are kept alive here until the
    request is reported."""

    def __init__(self, device=0, **config):
        self._lib = load_library()
        self._h = ctypes.c_void_p()
        self._keep = {}
        cfg = server_config(**config)
        _check(self._lib.gbsw_server_create(device, ctypes.byref(cfg), ctypes.byref(self._h)))

    def submit(self, scoring, ref, ref_off, query, query_off, out=None):
        """Queue the pairs; `out` (5 x n int32, optional) receives the results.
        Returns (ticket, out)."""
        n = len(ref_off) - 1
        if len(query_off) - 1 != n:
            raise ValueError("ref and query counts differ")
        ref_off = np.ascontiguousarray(ref_off, dtype=np.uint64)
        query_off = np.ascontiguousarray(query_off, dtype=np.uint64)
        ref = np.ascontiguousarray(ref, dtype=np.uint8)
        query = np.ascontiguousarray(query, dtype=np.uint8)
        if out is None:
            out = np.zeros((5, max(n, 1)), dtype=np.int32)
        assert out.dtype == np.int32 and out.flags["C_CONTIGUOUS"] and out.shape[0] == 5 and out.shape[1] >= n
        t = ctypes.c_uint64()
        base = out.ctypes.data
        stride = out.strides[0]
        _check(self._lib.gbsw_server_submit(self._h, ctypes.byref(scoring), n, ref.ctypes.data, ref_off.ctypes.data,
                                            query.ctypes.data, query_off.ctypes.data, base, base + stride,
                                            base + 2 * stride, base + 3 * stride, base + 4 * stride,
                                            ctypes.byref(t)))
        self._keep[t.value] = (ref, query, out)
        return t.value, out

    def wait(self, ticket, timeout=-1.0):
        ri = RequestInfo()
        _check(self._lib.gbsw_server_wait(self._h, ticket, float(timeout), ctypes.byref(ri)))
        self._keep.pop(ticket, None)
        return _info(ri)

    def poll(self, max_count=1024):
        infos = (RequestInfo * max_count)()
        cnt = ctypes.c_int32()
        _check(self._lib.gbsw_server_poll(self._h, max_count, infos, ctypes.byref(cnt)))
        res = [_info(infos[i]) for i in range(cnt.value)]
        for r in res:
            self._keep.pop(r["ticket"], None)
        return res

    def stats(self):
        st = ServingStats()
        _check(self._lib.gbsw_server_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServingStats._fields_}

    def close(self):
        if self._h:
            self._lib.gbsw_server_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


CIGAR_OPS = "MID"


def cigar_string(words):
    """BAM CIGAR words (length << 4 | op) -> text, e.g. '12M1I30M'."""
    return "".join("%d%s" % (int(w) >> 4, CIGAR_OPS[int(w) & 15]) for w in words)
