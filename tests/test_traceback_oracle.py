"""The oracle's alignment paths (CIGAR), pinned without the oracle's own DP.

The reference declares a traceback (kernel.hpp:36-39) but never implements it,
so the path is defined here: the best global alignment of the rectangle the
forward and reverse passes report, ties broken by the op sequence read from the
end (M < D < I).  These tests check the oracle against that definition by brute
force over every alignment of short strings, and check invariants on larger
random pairs.
"""
import random

import numpy as np
import pytest

import oracle_lib as ol
from gbsw.scoring import BLOSUM62

ORDER = {"M": 0, "D": 1, "I": 2}


def all_alignments(m, n):
    out = []

    def rec(i, j, ops):
        if i == m and j == n:
            out.append(tuple(ops))
            return
        if i < m and j < n:
            rec(i + 1, j + 1, ops + ["M"])
        if i < m:
            rec(i + 1, j, ops + ["D"])
        if j < n:
            rec(i, j + 1, ops + ["I"])

    rec(0, 0, [])
    return out


def op_score(ops, a, b, sub, gap_open, gap_extend):
    tot, i, j, prev = 0, 0, 0, None
    for o in ops:
        if o == "M":
            tot += sub(a[i], b[j])
            i += 1
            j += 1
        elif o == "D":
            tot += gap_extend if prev == "D" else gap_open
            i += 1
        else:
            tot += gap_extend if prev == "I" else gap_open
            j += 1
        prev = o
    return tot


def brute_path(a, b, sub, gap_open, gap_extend):
    best, pick = None, None
    for ops in all_alignments(len(a), len(b)):
        s = op_score(ops, a, b, sub, gap_open, gap_extend)
        key = [ORDER[o] for o in reversed(ops)]
        if best is None or s > best or (s == best and key < pick[0]):
            best, pick = s, (key, ops)
    return pick[1], best


def expand(words):
    return tuple(op for w in words for op in ol.CIGAR_OPS[w & 15] * (w >> 4))


def whole(a, b, sc):
    """Path of the whole of a against b (the rectangle forced to the full pair)."""
    return ol.traceback(a, b, sc, (1, 0, len(a), 0, len(b)))


@pytest.mark.parametrize("scoring", [(3, -3, -6, -1), (1, -1, -1, -1), (2, -1, -2, -2), (1, -3, -3, -1)])
def test_dna_paths_match_brute_force(scoring):
    match, mismatch, go, ge = scoring
    sc = ol.dna_scoring(match, mismatch, go, ge)
    sub = lambda x, y: match if x == y else mismatch  # noqa: E731
    rng = random.Random(hash(scoring) & 0xFFFF)
    cases = 0
    for m in range(1, 6):
        for n in range(1, 6):
            for _ in range(6):
                a = "".join(rng.choice("AC") for _ in range(m))
                b = "".join(rng.choice("AC") for _ in range(n))
                words, ps = whole(a, b, sc)
                ops, best = brute_path(a, b, sub, go, ge)
                assert ps == best, (a, b)
                assert expand(words) == ops, (a, b, ol.cigar_string(words))
                cases += 1
    assert cases == 150


def test_protein_paths_use_the_forward_orientation():
    # an asymmetric table: s(row, col) != s(col, row); rows are the longer full sequence
    mat = [0] * 576
    for x in range(24):
        for y in range(24):
            mat[x * 24 + y] = (3 if x == y else -2) + (1 if x < y else 0)
    sc = ol.aa_scoring(mat, -3, -1)
    enc = ol.encode_aa
    rng = random.Random(11)
    for _ in range(120):
        m, n = rng.randint(1, 5), rng.randint(1, 5)
        a = "".join(rng.choice("ARN") for _ in range(m))
        b = "".join(rng.choice("ARN") for _ in range(n))
        # rows = the longer of the two full sequences (ref on ties)
        if len(a) < len(b):
            sub = lambda x, y: mat[enc(ord(y)) * 24 + enc(ord(x))]  # noqa: E731
        else:
            sub = lambda x, y: mat[enc(ord(x)) * 24 + enc(ord(y))]  # noqa: E731
        words, ps = whole(a, b, sc)
        ops, best = brute_path(a, b, sub, -3, -1)
        assert ps == best and expand(words) == ops, (a, b, ol.cigar_string(words), ops)


def test_paths_on_aligned_pairs_span_the_rectangle_and_score_it():
    sc = ol.dna_scoring(3, -3, -6, -1)
    sub = lambda x, y: 3 if x == y else -3  # noqa: E731
    rng = random.Random(5)
    same = 0
    for _ in range(300):
        ref = "".join(rng.choice("ACGT") for _ in range(rng.randint(5, 90)))
        q = list(ref[rng.randint(0, 3):])
        for k in range(len(q)):
            if rng.random() < 0.1:
                q[k] = rng.choice("ACGT")
        q = "".join(q[: rng.randint(3, len(q))]) if len(q) >= 3 else "ACG"
        res = ol.align(ref, q, sc)
        words, ps = ol.traceback(ref, q, sc, res)
        if res[0] == 0:
            assert words == [] and ps == 0
            continue
        ops = expand(words)
        a, b = ref[res[1]:res[2]], q[res[3]:res[4]]
        assert sum(o != "I" for o in ops) == len(a) and sum(o != "D" for o in ops) == len(b)
        assert op_score(ops, a, b, sub, -6, -1) == ps
        # a local alignment from the reported begin to the reported end scores at
        # most the local score; when one reaches it, the path is such an alignment
        assert ps <= res[0]
        same += ps == res[0]
    assert same >= 280


def test_identical_strings_give_one_match_run():
    sc = ol.dna_scoring(3, -3, -6, -1)
    s = "ACGTTGCAAGGCTTAC" * 5
    res = ol.align(s, s, sc)
    words, ps = ol.traceback(s, s, sc, res)
    assert ol.cigar_string(words) == "%dM" % len(s) and ps == 3 * len(s)


def test_batch_form_equals_single_pairs():
    sc = ol.aa_scoring(BLOSUM62, -6, -1)
    rng = np.random.default_rng(3)
    alpha = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", dtype=np.uint8)
    refs = [alpha[rng.integers(0, 20, rng.integers(10, 80))] for _ in range(40)]
    qs = [np.concatenate([r[rng.integers(0, 5):], alpha[rng.integers(0, 20, 4)]]) for r in refs]
    ref_off = np.zeros(41, np.uint64)
    np.cumsum([len(r) for r in refs], out=ref_off[1:])
    q_off = np.zeros(41, np.uint64)
    np.cumsum([len(q) for q in qs], out=q_off[1:])
    rb, qb = np.concatenate(refs), np.concatenate(qs)
    res = ol.align_batch(rb, ref_off, qb, q_off, sc)
    words, off, ps = ol.traceback_batch(rb, ref_off, qb, q_off, sc, res)
    for k in range(40):
        w1, p1 = ol.traceback(refs[k].tobytes(), qs[k].tobytes(), sc, tuple(res[k]))
        assert list(words[off[k]:off[k + 1]]) == w1 and ps[k] == p1
