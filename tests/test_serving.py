"""GPU tests of the serving loop (gbsw_server_*): every request's results equal
the oracle's and gbsw_align's on that request alone, whatever it was batched
with, split across or queued behind.

Small slots and a small pinned arena force the interesting paths: requests
split over several micro-batches, micro-batches holding many requests, arena
wrap-around and submit back-pressure.
"""
import random
import threading
import time

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


def scorings(alphabet, scoring):
    if alphabet == gbsw.DNA:
        return gbsw.dna_scoring(*scoring), ol.dna_scoring(*scoring)
    return gbsw.protein_scoring(BLOSUM62, *scoring), ol.aa_scoring(BLOSUM62, *scoring)


def expect(ctx, w, scoring):
    """Oracle results, checked equal to gbsw_align on the request alone."""
    gsc, osc = scorings(w.alphabet, scoring)
    exp = ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, osc)
    alone = ctx.align_packed(w.ref, w.ref_off, w.query, w.query_off, gsc, "batched")
    for f in FIELDS:
        assert np.array_equal(alone[f], exp[f]), (w.name, "alone", f)
    return exp


def check(out, exp, n, label):
    for k, f in enumerate(FIELDS):
        bad = np.nonzero(out[k][:n] != exp[f][:n])[0]
        assert len(bad) == 0, "%s: %s differs at %d pairs, first %d: got %d want %d" % (
            label, f, len(bad), bad[0], out[k][bad[0]], exp[f][bad[0]])


def test_serving_mix_matches_alone_and_oracle(ctx):
    reqs = synth.serving_requests(48, seed=0x5EED0400, max_pairs=300)
    # a second DNA scoring scheme makes a third scoring class
    schemes = []
    for i, w in enumerate(reqs):
        if w.alphabet == gbsw.DNA and i % 3 == 0:
            schemes.append((1, -3, -3, -1))
        else:
            schemes.append(w.scoring)
    exps = [expect(ctx, w, s) for w, s in zip(reqs, schemes)]
    srv = gbsw.Server(0, slots=3, pairs_per_slot=256, residues_per_slot=1 << 20,
                      boundary_bytes_per_slot=64 << 20, staging_bytes=1 << 20)
    rng = random.Random(7)
    tickets = {}
    infos = {}

    def submitter():
        for i, (w, s) in enumerate(zip(reqs, schemes)):
            gsc, _ = scorings(w.alphabet, s)
            t, out = srv.submit(gsc, w.ref, w.ref_off, w.query, w.query_off)
            tickets[t] = (i, out)
            if rng.random() < 0.5:
                time.sleep(rng.random() * 0.003)

    th = threading.Thread(target=submitter)
    th.start()
    deadline = time.time() + 120
    while (th.is_alive() or len(infos) < len(reqs)) and time.time() < deadline:
        for info in srv.poll():
            infos[info["ticket"]] = info
        time.sleep(0.001)
    th.join()
    assert len(infos) == len(reqs), "only %d of %d requests completed" % (len(infos), len(reqs))
    for t, info in infos.items():
        i, out = tickets[t]
        assert info["status"] == gbsw.GBSW_OK, (i, info)
        assert info["pairs"] == reqs[i].n and info["cells"] == reqs[i].cells()
        assert info["latency_ms"] >= info["queue_ms"] >= 0
        check(out, exps[i], reqs[i].n, "request %d (%s)" % (i, reqs[i].name))
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["pending"] == 0
    assert st["cells"] == sum(w.cells() for w in reqs)
    assert max(info["parts"] for info in infos.values()) > 1, "no request was split"
    srv.close()


def test_serving_burst_is_batched(ctx):
    """A burst of small requests shares micro-batches (continuous batching),
    and results stay exact."""
    w = synth.plain_dna(400, seed=0x5EED0401)
    exp = expect(ctx, w, w.scoring)
    srv = gbsw.Server(0, slots=2)
    sc, _ = scorings(w.alphabet, w.scoring)
    subs = []
    for b in range(0, 400, 5):
        sub = w.subset(np.arange(b, b + 5))
        t, out = srv.submit(sc, sub.ref, sub.ref_off, sub.query, sub.query_off)
        subs.append((t, out, b))
    for t, out, b in subs:
        info = srv.wait(t)
        assert info["status"] == gbsw.GBSW_OK
        check(out, {f: exp[f][b:b + 5] for f in FIELDS}, 5, "burst request at %d" % b)
    st = srv.stats()
    assert st["micro_batches"] < len(subs), st
    srv.close()


def test_serving_zero_copy_unaligned(ctx):
    """Sequences in gbsw_host_alloc memory, at offsets that are not multiples of
    16, are read in place by the device."""
    w = synth.plain_protein(300, seed=0x5EED0402)
    exp = expect(ctx, w, w.scoring)
    srv = gbsw.Server(0, slots=2, pairs_per_slot=128)
    sc, _ = scorings(w.alphabet, w.scoring)
    nr, nq = len(w.ref), len(w.query)
    hb = gbsw.HostBuffer(nr + nq + 64)
    a = hb.array
    r0, q0 = 3, 3 + nr + 13
    a[r0:r0 + nr] = w.ref
    a[q0:q0 + nq] = w.query
    got = []
    for b in range(0, 300, 100):
        ro = (w.ref_off[b:b + 101] + np.uint64(r0)).astype(np.uint64)
        qo = (w.query_off[b:b + 101] + np.uint64(q0)).astype(np.uint64)
        t, out = srv.submit(sc, a, ro, a, qo)
        got.append((t, out, b))
    for t, out, b in got:
        info = srv.wait(t)
        assert info["status"] == gbsw.GBSW_OK
        check(out, {f: exp[f][b:b + 100] for f in FIELDS}, 100, "zero-copy request at %d" % b)
    srv.close()
    hb.close()


def test_serving_residue_errors_are_per_request(ctx):
    """A request with a byte outside its alphabet fails alone; the requests
    sharing its micro-batch are unaffected."""
    good = synth.plain_dna(64, seed=0x5EED0403)
    exp = expect(ctx, good, good.scoring)
    srv = gbsw.Server(0, slots=1)
    sc, _ = scorings(gbsw.DNA, good.scoring)
    bad_r, bad_ro = gbsw.pack(["ACGTACGTXACGT", "ACGT"])
    bad_q, bad_qo = gbsw.pack(["ACGTACGT", "ACG"])
    pr_r, pr_ro = gbsw.pack(["MKVLAAGIVa"])
    pr_q, pr_qo = gbsw.pack(["KVLAAG"])
    t1, out1 = srv.submit(sc, good.ref, good.ref_off, good.query, good.query_off)
    t2, out2 = srv.submit(sc, bad_r, bad_ro, bad_q, bad_qo)
    t3, out3 = srv.submit(gbsw.protein_scoring(), pr_r, pr_ro, pr_q, pr_qo)
    i1, i2, i3 = srv.wait(t1), srv.wait(t2), srv.wait(t3)
    assert i1["status"] == gbsw.GBSW_OK
    check(out1, exp, good.n, "good request")
    assert i2["status"] == gbsw.E_RESIDUE and not out2.any()
    assert i3["status"] == gbsw.E_RESIDUE
    # empty request completes at once
    e_r, e_ro = gbsw.pack([])
    t4, _ = srv.submit(sc, e_r, e_ro, e_r, e_ro)
    assert srv.wait(t4, timeout=10)["status"] == gbsw.GBSW_OK
    with pytest.raises(gbsw.GbswError):
        srv.wait(t4)  # reported tickets are released
    srv.close()


def test_serving_custom_dna_alphabet(ctx):
    """Raw-byte equality with a declared alphabet: lower case is its own symbol."""
    rng = np.random.default_rng(11)
    alpha = np.frombuffer(b"ACGTacgt", dtype=np.uint8)
    refs = [bytes(alpha[rng.integers(0, 8, size=int(rng.integers(20, 200)))]) for _ in range(50)]
    qrys = [r[5:5 + int(rng.integers(10, 60))] for r in refs]
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    exp = ol.align_batch(rb, ro, qb, qo, ol.dna_scoring(2, -3, -5, -2))
    srv = gbsw.Server(0, slots=2, dna_alphabet="ACGTacgt")
    t, out = srv.submit(gbsw.dna_scoring(2, -3, -5, -2), rb, ro, qb, qo)
    assert srv.wait(t)["status"] == gbsw.GBSW_OK
    check(out, exp, 50, "custom alphabet")
    srv.close()


def test_serving_large_requests_uploaded_by_dma(ctx):
    """Requests of 256 KiB of residues or more are uploaded by DMA into the
    device staging ring at submit (from gbsw_host_alloc memory at unaligned
    offsets, and from pageable memory through the pinned arena), small ones
    read over PCIe by the gather; mixed in one stream, every request equals
    the oracle, also when the staging ring is small enough to wrap and block."""
    w = synth.batched_dna(6000, seed=0x5EED0404)
    exp = expect(ctx, w, w.scoring)
    sc, _ = scorings(w.alphabet, w.scoring)
    nr, nq = len(w.ref), len(w.query)
    hb = gbsw.HostBuffer(nr + nq + 64)
    a = hb.array
    r0, q0 = 5, 5 + nr + 7
    a[r0:r0 + nr] = w.ref
    a[q0:q0 + nq] = w.query
    srv = gbsw.Server(0, slots=2, device_staging_bytes=3 << 20, staging_bytes=4 << 20)
    got = []
    # request sizes around the threshold (about 0.75 KiB of residues per pair)
    cuts = [0, 2000, 3500, 3510, 3800, 4500, 4530, 6000]
    for i, (b, e) in enumerate(zip(cuts[:-1], cuts[1:])):
        if i % 2 == 0:
            ro = (w.ref_off[b:e + 1] + np.uint64(r0)).astype(np.uint64)
            qo = (w.query_off[b:e + 1] + np.uint64(q0)).astype(np.uint64)
            t, out = srv.submit(sc, a, ro, a, qo)
        else:
            sub = w.subset(np.arange(b, e))
            t, out = srv.submit(sc, sub.ref, sub.ref_off, sub.query, sub.query_off)
        got.append((t, out, b, e))
    for t, out, b, e in got:
        info = srv.wait(t, timeout=60)
        assert info["status"] == gbsw.GBSW_OK
        check(out, {f: exp[f][b:e] for f in FIELDS}, e - b, "request [%d, %d)" % (b, e))
    srv.close()
    hb.close()
