"""Parity at full size: every pair of the 1M-pair DNA mix against the oracle,
1,000 long-protein pairs of the 50,000-pair set aligned as part of the whole
set, and served requests of up to 4,096 pairs.

Bar (BASELINE "Bit-exact" convention): all five int outputs equal the CPU
oracle's on every pair checked.
"""
import numpy as np
import pytest

import oracle_lib as ol
import gbsw
from gbsw import synth
from gbsw.scoring import BLOSUM62

pytestmark = pytest.mark.gpu

FIELDS = gbsw.FIELDS


@pytest.fixture(scope="module")
def ctx():
    c = gbsw.Context(0)
    yield c
    c.close()


def oracle_scoring(w):
    return ol.dna_scoring(*w.scoring) if w.alphabet == gbsw.DNA else ol.aa_scoring(BLOSUM62, *w.scoring)


def gbsw_scoring(w):
    return gbsw.dna_scoring(*w.scoring) if w.alphabet == gbsw.DNA else gbsw.protein_scoring(BLOSUM62, *w.scoring)


def assert_same(got, exp, label, idx=None):
    for f in FIELDS:
        g = got[f] if idx is None else got[f][idx]
        bad = np.nonzero(g != exp[f])[0]
        if len(bad):
            k = int(bad[0])
            detail = {h: (int((got[h] if idx is None else got[h][idx])[k]), int(exp[h][k])) for h in FIELDS}
            raise AssertionError("%s: %d pairs differ in %s; first pair %d (got, want) %s" %
                                 (label, len(bad), f, k if idx is None else int(idx[k]), detail))


def test_config2_every_pair_of_1m(ctx):
    """All 1,000,000 pairs of the batched DNA mix, batched and plain forms, all
    five outputs against the oracle."""
    w = synth.batched_dna()
    print("case config2 1M: oracle", flush=True)
    exp = ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, oracle_scoring(w))
    b = ctx.batch(w.alphabet, w.ref, w.ref_off, w.query, w.query_off)
    try:
        for mode in ("batched", "plain"):
            print("case config2 1M:", mode, flush=True)
            b.align(gbsw_scoring(w), mode)
            assert_same(b.results(), exp, "config2 1M " + mode)
    finally:
        b.close()


def test_config2_paths_across_launches(ctx):
    """Alignment paths of all 1,000,000 pairs of the batched DNA mix in one call:
    its scratch is cut into several launches run as two DP/walk pipelines (the
    smaller path tests fit one launch).  Every pair's path score equals its local
    score (the begin and end of every pair of this mix belong to one best local
    alignment), and every 20th pair's path (50,000 pairs across all launches)
    equals the oracle's."""
    w = synth.batched_dna()
    b = ctx.batch(w.alphabet, w.ref, w.ref_off, w.query, w.query_off)
    try:
        print("case config2 1M paths: align", flush=True)
        b.align(gbsw_scoring(w), "batched")
        got = b.results()
        print("case config2 1M paths: traceback", flush=True)
        words, off, ps, ms = b.traceback(gbsw_scoring(w))
    finally:
        b.close()
    assert np.array_equal(ps, got["score"]), "a path score differs from its local score"
    idx = np.arange(0, w.n, 20)
    ro = w.ref_off.astype(np.int64)
    qo = w.query_off.astype(np.int64)
    rl, ql = ro[idx + 1] - ro[idx], qo[idx + 1] - qo[idx]
    sro = np.zeros(len(idx) + 1, np.uint64)
    sqo = np.zeros(len(idx) + 1, np.uint64)
    np.cumsum(rl, out=sro[1:])
    np.cumsum(ql, out=sqo[1:])
    srb = np.concatenate([w.ref[ro[k]:ro[k + 1]] for k in idx])
    sqb = np.concatenate([w.query[qo[k]:qo[k + 1]] for k in idx])
    res = np.zeros(len(idx), dtype=ol.RESULT_DTYPE)
    for f in FIELDS:
        res[f] = got[f][idx]
    print("case config2 1M paths: oracle on", len(idx), "pairs", flush=True)
    ew, eoff, eps = ol.traceback_batch(srb, sro, sqb, sqo, oracle_scoring(w), res)
    assert np.array_equal(eps, ps[idx])
    for j, k in enumerate(idx):
        g = words[off[k]:off[k + 1]]
        e = ew[eoff[j]:eoff[j + 1]]
        if len(g) != len(e) or not np.array_equal(g, e):
            raise AssertionError("pair %d: got %s want %s" % (int(k), gbsw.cigar_string(g), ol.cigar_string(e)))


def test_config3_sample_of_the_full_set(ctx):
    """The whole 50,000-pair long-protein set is aligned at once (as the bench
    runs it); every 50th pair (1,000 pairs over the whole length range) is
    checked against the oracle, and the same pairs aligned as their own batch
    (split form) too."""
    w = synth.long_protein()
    idx = np.arange(0, w.n, 50)
    sub = w.subset(idx)
    lr, lq = sub.lengths()
    assert lq.min() < 1200 and lq.max() > 3900 and lr.min() < 5000 and lr.max() > 15500
    print("case config3: oracle on %d pairs" % len(idx), flush=True)
    exp = ol.align_batch(sub.ref, sub.ref_off, sub.query, sub.query_off, oracle_scoring(sub))
    print("case config3: 50k set batched", flush=True)
    b = ctx.batch(w.alphabet, w.ref, w.ref_off, w.query, w.query_off)
    try:
        b.align(gbsw_scoring(w), "batched")
        assert_same(b.results(), exp, "config3 in the full set", idx)
    finally:
        b.close()
    print("case config3: sample alone", flush=True)
    got = ctx.align_packed(sub.ref, sub.ref_off, sub.query, sub.query_off, gbsw_scoring(sub), "batched")
    assert_same(got, exp, "config3 sample alone")


def test_serving_requests_up_to_4096_pairs(ctx):
    """Served requests of the serving configuration's full size range (1-4,096
    pairs, 70 % DNA, 30 % protein), including requests of exactly 4,096 pairs
    of each alphabet, submitted at once: every request equals the oracle."""
    reqs = synth.serving_requests(40, seed=0x5EED0410, max_pairs=4096)
    reqs.append(synth.batched_dna(4096, seed=0x5EED0411))
    reqs.append(synth.plain_protein(4096, seed=0x5EED0412))
    assert max(w.n for w in reqs) == 4096
    exps = [ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, oracle_scoring(w)) for w in reqs]
    srv = gbsw.Server(0, slots=2)
    try:
        subs = []
        for w in reqs:
            t, out = srv.submit(gbsw_scoring(w), w.ref, w.ref_off, w.query, w.query_off)
            subs.append((t, out))
        print("case serving: %d requests, %d pairs submitted" % (len(reqs), sum(w.n for w in reqs)), flush=True)
        for (t, out), w, exp in zip(subs, reqs, exps):
            info = srv.wait(t, timeout=120)
            assert info["status"] == gbsw.GBSW_OK, info
            got = {f: out[k][:w.n] for k, f in enumerate(FIELDS)}
            assert_same(got, exp, "served %s (%d pairs)" % (w.name, w.n))
    finally:
        srv.close()
