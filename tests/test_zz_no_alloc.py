"""GPU test of the allocation-free entry point gbsw_align_noalloc: its results
equal gbsw_align's (the existing entry point it mirrors) on every pair, on two
calls in a row (the second smaller, so the static scratch holds the first
call's leftovers), in both forms and for both alphabets; the device's free
memory does not change across its calls; a call too large for the scratch is
rejected before any device work."""
import ctypes

import numpy as np
import pytest

import gbsw
from gbsw import synth
from gbsw.scoring import BLOSUM62

pytestmark = pytest.mark.gpu

FIELDS = gbsw.FIELDS


def device_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def scoring_of(w):
    return gbsw.dna_scoring(*w.scoring) if w.alphabet == gbsw.DNA else gbsw.protein_scoring(BLOSUM62, *w.scoring)


def assert_same(got, exp, label):
    for f in FIELDS:
        bad = np.nonzero(got[f] != exp[f])[0]
        assert len(bad) == 0, "%s: %d pairs differ in %s, first %d: got %d want %d" % (
            label, len(bad), f, bad[0], got[f][bad[0]], exp[f][bad[0]])


def test_noalloc_equals_gbsw_align_two_calls():
    ctx = gbsw.Context(0)
    try:
        cases = [
            (synth.batched_dna(20_000, seed=0x5EED0601), synth.batched_dna(3_000, seed=0x5EED0602)),
            (synth.plain_protein(4_000, seed=0x5EED0603), synth.plain_protein(700, seed=0x5EED0604)),
            # multi-strip pairs: boundary rows in the scratch
            (synth.long_protein(40, seed=0x5EED0605), synth.long_protein(9, seed=0x5EED0606)),
        ]
        # warm-up: kernel code objects load on first use
        w0 = cases[0][1]
        ctx.align_packed(w0.ref, w0.ref_off, w0.query, w0.query_off, scoring_of(w0), "batched", noalloc=True)
        ctx.align_packed(w0.ref, w0.ref_off, w0.query, w0.query_off, scoring_of(w0), "plain", noalloc=True)
        for big, small in cases:
            exp = {}
            for w in (big, small):
                for mode in ("batched", "plain"):
                    exp[(w.name, w.n, mode)] = ctx.align_packed(w.ref, w.ref_off, w.query, w.query_off,
                                                               scoring_of(w), mode)
            for mode in ("batched", "plain"):
                free0 = device_free_bytes()
                for w in (big, small):
                    print("case noalloc", w.name, w.n, mode, flush=True)
                    got = ctx.align_packed(w.ref, w.ref_off, w.query, w.query_off, scoring_of(w), mode, noalloc=True)
                    assert_same(got, exp[(w.name, w.n, mode)], "noalloc %s %d %s" % (w.name, w.n, mode))
                assert device_free_bytes() == free0, "device memory changed across gbsw_align_noalloc calls"
    finally:
        ctx.close()


def test_noalloc_rejects_calls_beyond_the_scratch():
    ctx = gbsw.Context(0)
    try:
        n = 6_000_000  # about 60 bytes of per-pair arrays each: more than the 256 MiB scratch
        rb = np.full(n, ord("A"), dtype=np.uint8)
        ro = np.arange(n + 1, dtype=np.uint64)
        with pytest.raises(gbsw.GbswError) as e:
            ctx.align_packed(rb, ro, rb, ro, gbsw.dna_scoring(), "batched", noalloc=True)
        assert e.value.code == gbsw.E_RANGE and "scratch" in str(e.value)
        # the context still works afterwards
        w = synth.plain_dna(500, seed=0x5EED0607)
        got = ctx.align_packed(w.ref, w.ref_off, w.query, w.query_off, scoring_of(w), "batched", noalloc=True)
        assert_same(got, ctx.align_packed(w.ref, w.ref_off, w.query, w.query_off, scoring_of(w), "batched"), "after")
    finally:
        ctx.close()
