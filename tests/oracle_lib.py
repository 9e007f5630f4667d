"""ctypes binding of the CPU oracle (oracle/libgbsw_oracle.so) for tests.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load the oracle.
"""
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "libgbsw_oracle.so")

DNA, AA = 0, 1
AA_ORDER = "ARNDCQEGHILKMFPSTWYVBZX*"


class Scoring(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("match", ctypes.c_int32),
        ("mismatch", ctypes.c_int32),
        ("gap_open", ctypes.c_int32),
        ("gap_extend", ctypes.c_int32),
        ("matrix", ctypes.c_int16 * 576),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("score", ctypes.c_int32),
        ("ref_begin", ctypes.c_int32),
        ("ref_end", ctypes.c_int32),
        ("query_begin", ctypes.c_int32),
        ("query_end", ctypes.c_int32),
    ]


RESULT_DTYPE = np.dtype([("score", np.int32), ("ref_begin", np.int32), ("ref_end", np.int32),
                         ("query_begin", np.int32), ("query_end", np.int32)])

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    src = os.path.join(ORACLE_DIR, "gbsw_oracle.c")
    if not os.path.exists(ORACLE_SO) or os.path.getmtime(ORACLE_SO) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "libgbsw_oracle.so"], check=True)
    lib = ctypes.CDLL(ORACLE_SO)
    lib.gbsw_oracle_encode_aa.argtypes = [ctypes.c_int]
    lib.gbsw_oracle_encode_aa.restype = ctypes.c_int
    lib.gbsw_oracle_prio.argtypes = [ctypes.c_uint32]
    lib.gbsw_oracle_prio.restype = ctypes.c_uint32
    lib.gbsw_oracle_align.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32,
                                      ctypes.POINTER(Scoring), ctypes.POINTER(Result)]
    lib.gbsw_oracle_align.restype = ctypes.c_int
    lib.gbsw_oracle_reverse.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32,
                                        ctypes.POINTER(Scoring), ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    lib.gbsw_oracle_reverse.restype = ctypes.c_int
    lib.gbsw_oracle_align_batch.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Scoring),
                                            ctypes.c_void_p, ctypes.c_int]
    lib.gbsw_oracle_align_batch.restype = ctypes.c_int64
    lib.gbsw_oracle_traceback.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32,
                                          ctypes.POINTER(Scoring), ctypes.POINTER(Result), ctypes.c_void_p,
                                          ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    lib.gbsw_oracle_traceback.restype = ctypes.c_int
    lib.gbsw_oracle_traceback_batch.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.POINTER(Scoring), ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_int]
    lib.gbsw_oracle_traceback_batch.restype = ctypes.c_int64
    lib.gbsw_oracle_threads.argtypes = [ctypes.c_int]
    lib.gbsw_oracle_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def dna_scoring(match, mismatch, gap_open, gap_extend):
    sc = Scoring()
    sc.kind, sc.match, sc.mismatch, sc.gap_open, sc.gap_extend = DNA, match, mismatch, gap_open, gap_extend
    return sc


def aa_scoring(matrix, gap_open, gap_extend):
    sc = Scoring()
    sc.kind, sc.gap_open, sc.gap_extend = AA, gap_open, gap_extend
    for k, v in enumerate(matrix):
        sc.matrix[k] = int(v)
    return sc


def align(ref, query, sc):
    """One pair; returns (score, ref_begin, ref_end, query_begin, query_end)."""
    lib = load()
    r = Result()
    rb, qb = ref.encode() if isinstance(ref, str) else ref, query.encode() if isinstance(query, str) else query
    rc = lib.gbsw_oracle_align(rb, len(rb), qb, len(qb), ctypes.byref(sc), ctypes.byref(r))
    if rc:
        raise ValueError("oracle rejected the pair (rc=%d)" % rc)
    return (r.score, r.ref_begin, r.ref_end, r.query_begin, r.query_end)


def reverse_max(ref, query, ref_end, query_end, sc):
    lib = load()
    m, a, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    rb, qb = ref.encode(), query.encode()
    lib.gbsw_oracle_reverse(rb, ref_end, qb, query_end, ctypes.byref(sc), ctypes.byref(m),
                            ctypes.byref(a), ctypes.byref(b))
    return m.value


def align_batch(ref_bytes, ref_off, query_bytes, query_off, sc, nthreads=0):
    """Batch through the OpenMP driver.  Offsets are n+1 exclusive prefix sums."""
    lib = load()
    ref_bytes = np.ascontiguousarray(ref_bytes, dtype=np.uint8)
    query_bytes = np.ascontiguousarray(query_bytes, dtype=np.uint8)
    ref_off = np.ascontiguousarray(ref_off, dtype=np.uint64)
    query_off = np.ascontiguousarray(query_off, dtype=np.uint64)
    n = len(ref_off) - 1
    out = np.zeros(n, dtype=RESULT_DTYPE)
    rc = lib.gbsw_oracle_align_batch(n, ref_bytes.ctypes.data, ref_off.ctypes.data, query_bytes.ctypes.data,
                                     query_off.ctypes.data, ctypes.byref(sc), out.ctypes.data, nthreads)
    if rc:
        raise ValueError("oracle rejected pair %d" % (-rc - 1))
    return out


def prio(t):
    return load().gbsw_oracle_prio(t)


def encode_aa(byte):
    return load().gbsw_oracle_encode_aa(byte)


CIGAR_OPS = "MID"


def cigar_string(words):
    """BAM CIGAR words (length << 4 | op) -> text, e.g. '12M1I30M'."""
    return "".join("%d%s" % (int(w) >> 4, CIGAR_OPS[int(w) & 15]) for w in words)


def traceback(ref, query, sc, res=None):
    """Alignment path of one pair: (list of BAM words, path score).  res: the
    pair's (score, ref_begin, ref_end, query_begin, query_end), aligned here when
    omitted."""
    lib = load()
    rb, qb = ref.encode() if isinstance(ref, str) else ref, query.encode() if isinstance(query, str) else query
    if res is None:
        res = align(rb, qb, sc)
    r = Result(*[int(x) for x in res])
    cap = max(1, len(rb) + len(qb))
    ops = np.zeros(cap, dtype=np.uint32)
    nops, ps = ctypes.c_int32(), ctypes.c_int32()
    rc = lib.gbsw_oracle_traceback(rb, len(rb), qb, len(qb), ctypes.byref(sc), ctypes.byref(r), ops.ctypes.data, cap,
                                   ctypes.byref(nops), ctypes.byref(ps))
    if rc:
        raise ValueError("oracle traceback failed (rc=%d)" % rc)
    return [int(x) for x in ops[:nops.value]], ps.value


def traceback_batch(ref_bytes, ref_off, query_bytes, query_off, sc, res, nthreads=0):
    """Paths of a batch given its results (RESULT_DTYPE array): (words, offsets
    (n+1), path scores)."""
    lib = load()
    ref_bytes = np.ascontiguousarray(ref_bytes, dtype=np.uint8)
    query_bytes = np.ascontiguousarray(query_bytes, dtype=np.uint8)
    ref_off = np.ascontiguousarray(ref_off, dtype=np.uint64)
    query_off = np.ascontiguousarray(query_off, dtype=np.uint64)
    res = np.ascontiguousarray(res, dtype=RESULT_DTYPE)
    n = len(ref_off) - 1
    span = (res["ref_end"].astype(np.int64) - res["ref_begin"] + res["query_end"] - res["query_begin"])
    span = np.where(res["score"] > 0, span, 0)
    cap_off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(span, out=cap_off[1:])
    ops = np.zeros(max(1, int(cap_off[-1])), dtype=np.uint32)
    nops = np.zeros(n, dtype=np.int32)
    ps = np.zeros(n, dtype=np.int32)
    rc = lib.gbsw_oracle_traceback_batch(n, ref_bytes.ctypes.data, ref_off.ctypes.data, query_bytes.ctypes.data,
                                         query_off.ctypes.data, ctypes.byref(sc), res.ctypes.data, ops.ctypes.data,
                                         cap_off.ctypes.data, nops.ctypes.data, ps.ctypes.data, nthreads)
    if rc:
        raise ValueError("oracle traceback rejected pair %d" % (-rc - 1))
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(nops, out=off[1:])
    words = np.concatenate([ops[cap_off[k]:cap_off[k] + nops[k]] for k in range(n)]) if off[-1] else np.zeros(0, np.uint32)
    return words.astype(np.uint32), off, ps
