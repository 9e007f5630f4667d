"""GPU parity: the HIP engine through the C ABI against the CPU oracle.

Bar: bit-exact on all five outputs (score, ref_begin, ref_end, query_begin,
query_end) for every pair, in both execution forms (plain and batched).
"""
import os
import random

import numpy as np
import pytest

import oracle_lib as ol
import gbsw
from gbsw import synth
from gbsw.scoring import BLOSUM62

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIELDS = gbsw.FIELDS


@pytest.fixture(scope="module")
def ctx():
    c = gbsw.Context(0)
    yield c
    c.close()


def oracle_scoring(alphabet, scoring):
    if alphabet == gbsw.DNA:
        return ol.dna_scoring(*scoring)
    return ol.aa_scoring(BLOSUM62, *scoring)


def gbsw_scoring(alphabet, scoring):
    if alphabet == gbsw.DNA:
        return gbsw.dna_scoring(*scoring)
    return gbsw.protein_scoring(BLOSUM62, *scoring)


def oracle_results(rb, ro, qb, qo, alphabet, scoring):
    out = ol.align_batch(rb, ro, qb, qo, oracle_scoring(alphabet, scoring))
    return {f: out[f] for f in FIELDS}


def assert_same(got, exp, label, refs=None):
    for f in FIELDS:
        bad = np.nonzero(got[f] != exp[f])[0]
        if len(bad):
            k = int(bad[0])
            detail = {g: (int(got[g][k]), int(exp[g][k])) for g in FIELDS}
            raise AssertionError("%s: %d pairs differ in %s; first pair %d (got, want) %s" %
                                 (label, len(bad), f, k, detail))


def run_both(ctx, rb, ro, qb, qo, alphabet, scoring, label):
    exp = oracle_results(rb, ro, qb, qo, alphabet, scoring)
    for mode in ("plain", "batched"):
        print("case", label, mode, "pairs", len(ro) - 1, flush=True)
        got = ctx.align_packed(rb, ro, qb, qo, gbsw_scoring(alphabet, scoring), mode)
        assert_same(got, exp, "%s/%s" % (label, mode))
    return exp


def rand_pairs(seed, n, lo, hi, alphabets, query_lo=None, query_hi=None):
    rng = random.Random(seed)
    refs, qrys = [], []
    for _ in range(n):
        alpha = rng.choice(alphabets)
        refs.append("".join(rng.choice(alpha) for _ in range(rng.randint(lo, hi))))
        qrys.append("".join(rng.choice(alpha) for _ in range(rng.randint(query_lo or lo, query_hi or hi))))
    return refs, qrys


def test_fixtures(ctx):
    import json
    with open(os.path.join(HERE, "golden", "tie_fixtures.json")) as fh:
        fx = json.load(fh)
    refs = [c[0] for c in fx["cases"]]
    qrys = [c[1] for c in fx["cases"]]
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    exp = run_both(ctx, rb, ro, qb, qo, gbsw.DNA, tuple(fx["scoring"]), "tie_fixtures")
    for k, c in enumerate(fx["cases"]):
        assert tuple(int(exp[f][k]) for f in FIELDS) == tuple(c[2])


@pytest.mark.parametrize("scheme", [(1, -3, -3, -1), (3, -3, -6, -1), (2, -1, -2, -2), (5, -4, -10, 0)])
def test_random_dna_small_alphabets(ctx, scheme):
    refs, qrys = rand_pairs(hash(scheme) & 0xFFFF, 600, 0, 90, ["AC", "ACG", "ACGT", "AAC"])
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    run_both(ctx, rb, ro, qb, qo, gbsw.DNA, scheme, "dna_small_%s" % (scheme,))


def test_random_protein_small(ctx):
    refs, qrys = rand_pairs(77, 500, 0, 120, ["ARND", "WYFHA", "ARNDCQEGHILKMFPSTWYV", "BZX*AC"])
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    run_both(ctx, rb, ro, qb, qo, gbsw.PROTEIN, (-6, -1), "protein_small")


def test_tie_heavy_cases(ctx):
    from test_oracle import _tie_heavy_cases
    cases = _tie_heavy_cases(31, 400) + _tie_heavy_cases(32, 400)
    rb, ro = gbsw.pack([c[0] for c in cases])
    qb, qo = gbsw.pack([c[1] for c in cases])
    run_both(ctx, rb, ro, qb, qo, gbsw.DNA, (1, -3, -3, -1), "tie_heavy")


def test_edge_cases(ctx):
    refs = ["", "A", "", "A", "C", "ACGT", "NNNN", "acgtACGT", "AAAA", "ACGTN" * 40, "A" * 700, "T"]
    qrys = ["", "", "A", "A", "A", "ACGT", "NNAN", "ACGTacgt", "CCCC", "ACGT" * 60, "A" * 700, "A" * 1500]
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    exp = run_both(ctx, rb, ro, qb, qo, gbsw.DNA, (1, -3, -3, -1), "edges")
    assert exp["score"][0] == 0 and exp["score"][3] == 1 and exp["score"][8] == 0


def test_single_pair_and_odd_counts(ctx):
    """Batches of 1, 2, 3 and 5 pairs (the reference needs >= 2 per device)."""
    refs, qrys = rand_pairs(5, 5, 20, 60, ["ACGT"])
    for n in (1, 2, 3, 5):
        rb, ro = gbsw.pack(refs[:n])
        qb, qo = gbsw.pack(qrys[:n])
        run_both(ctx, rb, ro, qb, qo, gbsw.DNA, (3, -3, -6, -1), "n=%d" % n)


def test_multi_strip_dna(ctx):
    """Shorter sequences longer than one strip: boundary columns pass through
    global memory between strips (plain K=32 covers 1024 columns)."""
    refs, qrys = rand_pairs(9, 40, 1100, 2600, ["ACGT", "AC"])
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    run_both(ctx, rb, ro, qb, qo, gbsw.DNA, (3, -3, -6, -1), "multi_strip_dna")


def test_multi_strip_protein_mutated(ctx):
    """Long-protein generator, cut to refs <= 3000 and queries <= 1500 so every
    score is provably inside int16, crossing up to six 256-column strips."""
    w = synth.long_protein(24)
    refs = [bytes(w.pair(k)[0][:3000]) for k in range(w.n)]
    qrys = [bytes(w.pair(k)[1][:1500]) for k in range(w.n)]
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    run_both(ctx, rb, ro, qb, qo, w.alphabet, w.scoring, "long_protein_24_cut")


def test_split_form_runs_multi_strip_buckets(ctx):
    """Multi-strip DNA and protein buckets run as split items (several
    workgroups per pair, pipelined on the boundary rows), and match the oracle;
    the small-K geometry model makes many strips per pair."""
    refs, qrys = rand_pairs(21, 300, 600, 2000, ["ACGT", "AC"], 300, 1200)
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    exp = oracle_results(rb, ro, qb, qo, gbsw.DNA, (3, -3, -6, -1))
    w = synth.long_protein(24, seed=0x5EED0207)
    prefs = [bytes(w.pair(k)[0][:2500]) for k in range(w.n)]
    pqrys = [bytes(w.pair(k)[1][:1200]) for k in range(w.n)]
    prb, pro = gbsw.pack(prefs)
    pqb, pqo = gbsw.pack(pqrys)
    pexp = oracle_results(prb, pro, pqb, pqo, gbsw.PROTEIN, w.scoring)
    for label, args, alphabet, scoring, want in (("dna", (rb, ro, qb, qo), gbsw.DNA, (3, -3, -6, -1), exp),
                                                 ("protein", (prb, pro, pqb, pqo), gbsw.PROTEIN, w.scoring, pexp)):
        print("case split", label, flush=True)
        b = ctx.batch(alphabet, *args)
        b.align(gbsw_scoring(alphabet, scoring), "batched")
        assert_same(b.results(), want, "split/" + label)
        kinds = {L["kind"] & 0xFF for L in b.launches()}
        assert kinds & {gbsw.KIND_SPLIT, gbsw.KIND_SPLIT_F16}, kinds
        b.close()


def test_config0_plain_dna_full(ctx):
    w = synth.plain_dna()
    run_both(ctx, w.ref, w.ref_off, w.query, w.query_off, w.alphabet, w.scoring, "config0_full")


def test_config1_plain_protein_full(ctx):
    w = synth.plain_protein()
    run_both(ctx, w.ref, w.ref_off, w.query, w.query_off, w.alphabet, w.scoring, "config1_full")


def test_config2_batched_dna_subset(ctx):
    w = synth.batched_dna(20_000, seed=123)
    run_both(ctx, w.ref, w.ref_off, w.query, w.query_off, w.alphabet, w.scoring, "config2_20k")


def test_resident_batch_reuse(ctx):
    """One resident batch aligned twice with different scorings and modes."""
    w = synth.plain_dna(3000, seed=4)
    b = ctx.batch(w.alphabet, w.ref, w.ref_off, w.query, w.query_off)
    for scheme in [(3, -3, -6, -1), (1, -3, -3, -1)]:
        exp = oracle_results(w.ref, w.ref_off, w.query, w.query_off, 0, scheme)
        for mode in ("batched", "plain"):
            b.align(gbsw.dna_scoring(*scheme), mode)
            assert_same(b.results(), exp, "reuse/%s/%s" % (scheme, mode))
            st = b.stats()
            assert st["cells"] == w.cells() and st["device_ms"] > 0
    b.close()


def test_errors_are_reported(ctx):
    with pytest.raises(gbsw.GbswError) as e:
        ctx.align(["ACD"], ["ACa"], gbsw.protein_scoring())  # 'a' = 97 >= 91
    assert e.value.code == -2
    with pytest.raises(gbsw.GbswError) as e:
        ctx.align(["ACGT"], ["ACGT"], gbsw.dna_scoring(1, -1, 2, -1))  # positive gap score
    assert e.value.code == -3
    # 32 distinct DNA bytes (found by the device scan of the uploaded batch)
    many = "".join(chr(c) for c in range(65, 97))
    with pytest.raises(gbsw.GbswError) as e:
        ctx.align([many], ["ACGT"], gbsw.dna_scoring(3, -3, -6, -1))
    assert e.value.code == -2
    # 31 distinct bytes are fine, in their byte order
    got = ctx.align([many[:31]], [many[5:20]], gbsw.dna_scoring(3, -3, -6, -1))
    assert int(got["score"][0]) == 45 and int(got["ref_begin"][0]) == 5


def test_drop_in_python_driver(ctx):
    reads, contigs = ["ACGTTACGT", "ACGT"], ["GGGGGACGTGGGG", "TTACGTTT"]
    r = gbsw.kernel_driver_dna(reads, contigs, [1, -3, -3, -1])
    assert (int(r["score"][0]), int(r["ref_begin"][0]), int(r["ref_end"][0]),
            int(r["query_begin"][0]), int(r["query_end"][0])) == (4, 5, 9, 5, 9)
    a = gbsw.kernel_driver_aa(["MKV"], ["AMKVA"], BLOSUM62, -6, -1)
    assert int(a["score"][0]) == 5 + 5 + 4


def test_config3_long_protein_full_length(ctx):
    """Long-protein pairs at full length (query 1024-4096, ref 4096-16384):
    several strips, int16 cells with saturation detection."""
    w = synth.long_protein(48, seed=303)
    run_both(ctx, w.ref, w.ref_off, w.query, w.query_off, w.alphabet, w.scoring, "config3_48")


def test_int16_saturation_falls_back_to_int32(ctx):
    """Scores beyond the int16 range: identical W-rich proteins (11 per residue)
    and DNA with a large match score, mixed with ordinary pairs.  Saturated pairs
    are recomputed with int32 cells and must still match the oracle exactly."""
    rng = random.Random(17)
    refs, qrys = [], []
    for k in range(24):
        if k % 3 == 0:
            core = "".join(rng.choice("WWWWC") for _ in range(rng.randint(3000, 4200)))
            refs.append("AKR" * 50 + core + "GG" * 40)
            qrys.append(core[: len(core) - rng.randint(0, 50)])
        else:
            refs.append("".join(rng.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(rng.randint(200, 900))))
            qrys.append(refs[-1][rng.randint(0, 50):][:rng.randint(100, 400)])
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    exp = run_both(ctx, rb, ro, qb, qo, gbsw.PROTEIN, (-6, -1), "saturation_protein")
    assert exp["score"].max() > 32767
    refs, qrys = rand_pairs(18, 30, 300, 900, ["ACGT"])
    for k in range(0, 30, 5):
        qrys[k] = refs[k][10:10 + 700] if len(refs[k]) > 720 else refs[k]
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    exp = run_both(ctx, rb, ro, qb, qo, gbsw.DNA, (100, -3, -6, -1), "saturation_dna")
    assert exp["score"].max() > 32767


def test_dropin_cpp_caller(tmp_path):
    """A C++ program written against the reference's driver API, linked to
    libgbsw.so, prints the reference's five columns (main.cpp:130-137)."""
    import subprocess
    from test_abi import build_dropin_caller
    exe = build_dropin_caller(tmp_path)
    full = subprocess.run([exe, "dna"], capture_output=True, text=True, check=True).stdout.splitlines()
    # the driver's own report lines (driver.cpp:240-242) come first; rows are tab-separated
    out = [l for l in full if "\t" in l]
    assert any(l.startswith("Total Execution Time (seconds):") for l in full)
    assert out[0] == "4\t5\t8\t5\t8"          # score, ref_begin, ref_end-1, query_begin, query_end-1
    assert out[1] == "4\t0\t3\t0\t3"
    sc = ol.dna_scoring(1, -3, -3, -1)
    s, rb, re_, qb, qe = ol.align("TTTTACGTACGTTTTT", "ACGTACGT", sc)
    assert out[2] == "%d\t%d\t%d\t%d\t%d" % (s, rb, re_ - 1, qb, qe - 1)
    outa = [l for l in subprocess.run([exe, "aa"], capture_output=True, text=True, check=True).stdout.splitlines()
            if "\t" in l]
    assert outa[1].split("\t")[0] == "33"      # WWW vs AWWWA: 3 x 11


def test_dropin_cpp_caller_large_call_runs_pipelined(tmp_path):
    """A drop-in call of 131,072 pairs or more runs as sub-batches on two
    contexts (upload of one beside the alignment of another): its five columns
    equal the oracle's on every pair."""
    import subprocess
    from test_abi import build_dropin_caller
    exe = build_dropin_caller(tmp_path)
    w = synth.plain_dna(140_000, seed=0x5EED0301)
    refs = [w.ref[w.ref_off[k]:w.ref_off[k + 1]].tobytes().decode() for k in range(w.n)]
    qs = [w.query[w.query_off[k]:w.query_off[k + 1]].tobytes().decode() for k in range(w.n)]
    fr, fq = tmp_path / "refs.txt", tmp_path / "queries.txt"
    fr.write_text("\n".join(refs) + "\n")
    fq.write_text("\n".join(qs) + "\n")
    print("case dropin pipelined", w.n, flush=True)
    lines = subprocess.run([exe, "file", str(fr), str(fq)], capture_output=True, text=True, check=True,
                           timeout=300).stdout.splitlines()
    rows = np.array([[int(x) for x in l.split("\t")] for l in lines if "\t" in l], dtype=np.int64)
    assert rows.shape == (w.n, 5)
    exp = ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, ol.dna_scoring(3, -3, -6, -1))
    want = np.stack([exp["score"], exp["ref_begin"], exp["ref_end"] - 1, exp["query_begin"], exp["query_end"] - 1], 1)
    bad = np.nonzero((rows != want).any(1))[0]
    assert len(bad) == 0, (len(bad), int(bad[0]), rows[bad[0]], want[bad[0]])


def test_batched_options_give_identical_results(ctx):
    """Serialised or concurrent buckets, f16 kernels on or off, multi-strip
    pairs split over workgroups or not: the same results (the f16 kernels run
    only where their exact range holds)."""
    for w in (synth.batched_dna(3000, seed=0x5EED0202), synth.plain_protein(600, seed=0x5EED0203)):
        exp = oracle_results(w.ref, w.ref_off, w.query, w.query_off, w.alphabet, w.scoring)
        sc = gbsw_scoring(w.alphabet, w.scoring)
        try:
            for streams, f16, split in ((4, 1, 1), (1, 1, 1), (8, 0, 1), (1, 0, 0), (4, 1, 0), (8, 1, 1)):
                print("case options", w.name, "streams", streams, "f16", f16, "split", split, flush=True)
                ctx.set_option(gbsw.OPT_STREAMS, streams)
                ctx.set_option(gbsw.OPT_F16, f16)
                ctx.set_option(gbsw.OPT_SPLIT, split)
                got = ctx.align_packed(w.ref, w.ref_off, w.query, w.query_off, sc, "batched")
                assert_same(got, exp, "%s streams=%d f16=%d split=%d" % (w.name, streams, f16, split))
        finally:
            ctx.set_option(gbsw.OPT_STREAMS, 4)
            ctx.set_option(gbsw.OPT_F16, 1)
            ctx.set_option(gbsw.OPT_SPLIT, 1)
    with pytest.raises(gbsw.GbswError):
        ctx.set_option(gbsw.OPT_STREAMS, 9)


def test_graph_replays_match_eager_results(ctx):
    """A resident batch replays the hipGraph captured at its first align: every replay,
    with eager aligns of other data and synchronisations between the replays,
    equals the eager result; a new scoring captures again, and going back to the
    first scoring captures again too (the graph is keyed by scoring and options)."""
    w = synth.batched_dna(3000, seed=0x5EED0208)
    wp = synth.plain_protein(500, seed=0x5EED0209)
    refs, qrys = rand_pairs(22, 60, 600, 1800, ["ACGT"], 300, 1100)   # multi-strip, split form
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    sc1, sc2 = gbsw.dna_scoring(*w.scoring), gbsw.dna_scoring(2, -3, -5, -2)
    exp1 = oracle_results(w.ref, w.ref_off, w.query, w.query_off, gbsw.DNA, w.scoring)
    exp2 = oracle_results(w.ref, w.ref_off, w.query, w.query_off, gbsw.DNA, (2, -3, -5, -2))
    expm = oracle_results(rb, ro, qb, qo, gbsw.DNA, w.scoring)
    expp = oracle_results(wp.ref, wp.ref_off, wp.query, wp.query_off, gbsw.PROTEIN, wp.scoring)
    a = ctx.batch(gbsw.DNA, w.ref, w.ref_off, w.query, w.query_off)
    m = ctx.batch(gbsw.DNA, rb, ro, qb, qo)
    try:
        states = []
        for rep, (sc, exp) in enumerate([(sc1, exp1)] * 4 + [(sc2, exp2)] * 3 + [(sc1, exp1)] * 2):
            print("case graph replay", rep, flush=True)
            a.align(sc, "batched")
            states.append(a.stats()["graph"])
            assert_same(a.results(), exp, "graph replay %d" % rep)
            # eager and graphed work of other batches between the replays
            got = ctx.align_packed(wp.ref, wp.ref_off, wp.query, wp.query_off, gbsw_scoring(gbsw.PROTEIN, wp.scoring),
                                   "batched")
            assert_same(got, expp, "eager protein between replays %d" % rep)
            m.align(sc1, "batched")
            assert_same(m.results(), expm, "multi-strip batch between replays %d" % rep)
        # captured at the first align (again if another batch grew the context's
        # buffers meanwhile) and replayed after; each scoring change captures again
        assert states[0] == 1 and states[4] == 1 and states[7] == 1, states
        assert states[2:4] == [2, 2] and states[5:7] == [2, 2] and states[8] == 2, states
        assert m.stats()["graph"] == 2
    finally:
        a.close()
        m.close()


def test_autotune_picks_a_model_and_keeps_results(ctx):
    """gbsw_batch_autotune times the distinct length -> geometry maps on the batch
    and keeps the fastest; the results stay the oracle's."""
    w = synth.batched_dna(4000, seed=0x5EED020A)
    exp = oracle_results(w.ref, w.ref_off, w.query, w.query_off, gbsw.DNA, w.scoring)
    sc = gbsw.dna_scoring(*w.scoring)
    b = ctx.batch(gbsw.DNA, w.ref, w.ref_off, w.query, w.query_off)
    try:
        print("case autotune", flush=True)
        model, ms = b.autotune(sc, repeats=1)
        assert 0 <= model < 6 and ms > 0
        assert_same(b.results(), exp, "autotune runs")
        b.align(sc, "batched")
        assert_same(b.results(), exp, "after autotune")
    finally:
        b.close()
        ctx.set_option(gbsw.OPT_GEOMETRY, 1)


def test_geometry_models_give_identical_results(ctx):
    """Both length -> geometry models (the plain form of protein follows the
    option too): the same results."""
    for w in (synth.batched_dna(3000, seed=0x5EED0204), synth.plain_protein(800, seed=0x5EED0205)):
        exp = oracle_results(w.ref, w.ref_off, w.query, w.query_off, w.alphabet, w.scoring)
        sc = gbsw_scoring(w.alphabet, w.scoring)
        try:
            for model in range(6):
                ctx.set_option(gbsw.OPT_GEOMETRY, model)
                for mode in ("plain", "batched"):
                    print("case geometry", w.name, "model", model, mode, flush=True)
                    got = ctx.align_packed(w.ref, w.ref_off, w.query, w.query_off, sc, mode)
                    assert_same(got, exp, "%s model=%d %s" % (w.name, model, mode))
        finally:
            ctx.set_option(gbsw.OPT_GEOMETRY, 1)


def test_protein_scores_beyond_int8_use_the_int16_profile(ctx):
    """BLOSUM62 x 12 (scores -48..132) does not fit the int8 profile: those
    launches take the int16 profile at a geometry that fits its LDS, BLOSUM62
    itself the int8 one; both match the oracle."""
    w = synth.plain_protein(400, seed=0x5EED0206)
    big = [12 * v for v in BLOSUM62]
    exp = ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, ol.aa_scoring(big, -30, -4))
    for mode in ("plain", "batched"):
        print("case scaled blosum", mode, flush=True)
        b = ctx.batch(w.alphabet, w.ref, w.ref_off, w.query, w.query_off)
        b.align(gbsw.protein_scoring(big, -30, -4), mode)
        assert_same(b.results(), {f: exp[f] for f in FIELDS}, "blosum62x12/" + mode)
        assert all(L["kind"] < gbsw.KIND_BYTE_PROFILE for L in b.launches())
        b.align(gbsw.protein_scoring(BLOSUM62, -6, -1), mode)
        assert all(L["kind"] >= gbsw.KIND_BYTE_PROFILE for L in b.launches())
        b.close()


def test_autotune_then_replay_uses_the_kept_map(ctx):
    """A batch large enough for its buckets to fill the grid (several maps give
    the same launch plans): after gbsw_batch_autotune has tried the maps, the
    results are overwritten by a plain run of another scoring, and the next
    batched align (a graph replay or a new capture) must use the kept map's
    bucket slices and give the oracle's results."""
    w = synth.batched_dna(200_000, seed=0x5EED020B)
    exp = oracle_results(w.ref, w.ref_off, w.query, w.query_off, gbsw.DNA, w.scoring)
    other = oracle_results(w.ref, w.ref_off, w.query, w.query_off, gbsw.DNA, (1, -3, -3, -1))
    sc = gbsw.dna_scoring(*w.scoring)
    b = ctx.batch(gbsw.DNA, w.ref, w.ref_off, w.query, w.query_off)
    try:
        print("case autotune then replay", flush=True)
        b.autotune(sc, repeats=1)
        for rep in range(2):
            b.align(gbsw.dna_scoring(1, -3, -3, -1), "plain")
            assert_same(b.results(), other, "overwrite %d" % rep)
            b.align(sc, "batched")
            assert_same(b.results(), exp, "after autotune %d" % rep)
    finally:
        b.close()
        ctx.set_option(gbsw.OPT_GEOMETRY, 1)


def test_f16_kernels_take_a_byte_profile_when_the_scores_allow(ctx):
    """The f16 kernels read a byte profile (the scores' f16 high bytes) when
    every score has at most three significant bits (3/-3, or BLOSUM62 with 9
    and 11 moved to 8 and 12), and the int16 profile otherwise (9/-11): all
    three give the oracle's results."""
    w = synth.batched_dna(3000, seed=0x5EED020C)
    f16_kinds = (gbsw.KIND_BATCHED_F16, gbsw.KIND_SPLIT_F16)
    b = ctx.batch(w.alphabet, w.ref, w.ref_off, w.query, w.query_off)
    try:
        for scoring, byte_profile in (((3, -3, -6, -1), True), ((9, -11, -13, -2), False)):
            print("case f16 profile", scoring, flush=True)
            exp = oracle_results(w.ref, w.ref_off, w.query, w.query_off, gbsw.DNA, scoring)
            b.align(gbsw.dna_scoring(*scoring), "batched")
            assert_same(b.results(), exp, "dna %s" % (scoring,))
            f16 = [L for L in b.launches() if L["kind"] & 0xFF in f16_kinds]
            assert f16, b.launches()
            assert all(bool(L["kind"] & gbsw.KIND_BYTE_PROFILE) == byte_profile for L in f16), f16
    finally:
        b.close()
    wp = synth.plain_protein(600, seed=0x5EED020D)
    table = [8 if v == 9 else 12 if v == 11 else v for v in BLOSUM62]
    exp = ol.align_batch(wp.ref, wp.ref_off, wp.query, wp.query_off, ol.aa_scoring(table, -6, -1))
    b = ctx.batch(wp.alphabet, wp.ref, wp.ref_off, wp.query, wp.query_off)
    try:
        print("case f16 profile protein", flush=True)
        b.align(gbsw.protein_scoring(table, -6, -1), "batched")
        assert_same(b.results(), {f: exp[f] for f in FIELDS}, "protein, exact f16 bytes")
        kinds = {L["kind"] for L in b.launches()}
        assert kinds & {gbsw.KIND_BATCHED_F16 | gbsw.KIND_BYTE_PROFILE, gbsw.KIND_SPLIT_F16 | gbsw.KIND_BYTE_PROFILE}, kinds
    finally:
        b.close()
