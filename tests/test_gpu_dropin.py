"""GPU tests of the C++ drop-in (gpu_bsw_driver::kernel_driver_dna / kernel_driver_aa,
driver.hpp:32-37) called from a program written against the reference's API:
its five `short` columns equal the oracle's on every pair, its `factor` memory
bound splits a call into sub-batches without changing results, and results
beyond the `short` range are reported as GBSW_E_RANGE instead of wrapping.
"""
import os
import random
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
import gbsw
from gbsw import synth
from gbsw.scoring import BLOSUM62

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    from test_abi import build_dropin_caller
    return build_dropin_caller(tmp_path_factory.mktemp("dropin"))


def write_lines(path, buf, off):
    """One sequence per line: the bytes with a newline after each pair's end."""
    lines = np.insert(np.asarray(buf, dtype=np.uint8), np.asarray(off[1:], dtype=np.int64), ord("\n"))
    lines.tofile(str(path))


def run_caller(args, env=None, timeout=300):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=e)


def columns(stdout):
    rows = [l for l in stdout.splitlines() if "\t" in l]
    status = [int(l.split()[1]) for l in stdout.splitlines() if l.startswith("status ")]
    return np.array([[int(x) for x in l.split("\t")] for l in rows], dtype=np.int64).reshape(-1, 5), status[-1]


def expected_columns(exp):
    """The reference sample's five columns (main.cpp:130-137): score,
    ref_begin, ref_end - 1, query_begin, query_end - 1."""
    return np.stack([exp["score"], exp["ref_begin"], exp["ref_end"] - 1, exp["query_begin"],
                     exp["query_end"] - 1], 1).astype(np.int64)


def assert_rows(rows, want, label):
    assert rows.shape == want.shape, (label, rows.shape, want.shape)
    bad = np.nonzero((rows != want).any(1))[0]
    assert len(bad) == 0, "%s: %d pairs differ, first %d: got %s want %s" % (
        label, len(bad), int(bad[0]), rows[bad[0]], want[bad[0]])


def test_kernel_driver_aa_protein_20k_matches_oracle(caller, tmp_path):
    """All 20,000 pairs of the protein configuration (BLOSUM62, -6/-1) through
    gpu_bsw_driver::kernel_driver_aa: every pair's five columns equal the oracle's."""
    w = synth.plain_protein()
    fr, fq, fm = tmp_path / "refs.txt", tmp_path / "queries.txt", tmp_path / "blosum62.txt"
    write_lines(fr, w.ref, w.ref_off)
    write_lines(fq, w.query, w.query_off)
    fm.write_text(" ".join(str(v) for v in BLOSUM62) + "\n")
    print("case dropin aa 20k", flush=True)
    out = run_caller([caller, "fileaa", str(fr), str(fq), str(fm), "-6", "-1"])
    assert out.returncode == 0, out.stderr[-2000:]
    rows, status = columns(out.stdout)
    assert status == 0
    exp = ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, ol.aa_scoring(BLOSUM62, -6, -1))
    assert_rows(rows, expected_columns(exp), "kernel_driver_aa 20k")
    assert "Total Execution Time (seconds):" in out.stdout


def test_kernel_driver_aa_scaled_table_matches_oracle(caller, tmp_path):
    """A caller's own 576-entry table (BLOSUM62 x 3, gap -11/-2) through
    kernel_driver_aa on 2,000 protein pairs."""
    w = synth.plain_protein(2000, seed=0x5EED0501)
    table = [3 * v for v in BLOSUM62]
    fr, fq, fm = tmp_path / "refs.txt", tmp_path / "queries.txt", tmp_path / "m.txt"
    write_lines(fr, w.ref, w.ref_off)
    write_lines(fq, w.query, w.query_off)
    fm.write_text(" ".join(str(v) for v in table) + "\n")
    out = run_caller([caller, "fileaa", str(fr), str(fq), str(fm), "-11", "-2"])
    assert out.returncode == 0, out.stderr[-2000:]
    rows, status = columns(out.stdout)
    assert status == 0
    exp = ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, ol.aa_scoring(table, -11, -2))
    assert_rows(rows, expected_columns(exp), "kernel_driver_aa scaled table")


@pytest.mark.parametrize("factor", [1e-5, 2e-7])
def test_factor_bounds_sub_batches(caller, tmp_path, factor):
    """factor x device memory (driver.cpp:36, 51-55) bounds a call's footprint:
    with a tiny factor the 3,000-pair call runs as several sub-batches (about
    3 and 130 on a 288 GB device) and its results are the oracle's."""
    w = synth.plain_dna(3000, seed=0x5EED0502)
    fr, fq = tmp_path / "refs.txt", tmp_path / "queries.txt"
    write_lines(fr, w.ref, w.ref_off)
    write_lines(fq, w.query, w.query_off)
    print("case dropin factor", factor, flush=True)
    out = run_caller([caller, "file", str(fr), str(fq), repr(factor)])
    assert out.returncode == 0, out.stderr[-2000:]
    rows, status = columns(out.stdout)
    assert status == 0
    exp = ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, ol.dna_scoring(3, -3, -6, -1))
    assert_rows(rows, expected_columns(exp), "factor %g" % factor)


def _saturating_proteins():
    """W-rich identical cores (11 per residue): scores above 32,767."""
    rng = random.Random(17)
    refs, qrys = [], []
    for k in range(6):
        if k % 3 == 0:
            core = "".join(rng.choice("WWWWC") for _ in range(rng.randint(3000, 4200)))
            refs.append("AKR" * 50 + core + "GG" * 40)
            qrys.append(core[: len(core) - rng.randint(0, 50)])
        else:
            refs.append("".join(rng.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(rng.randint(200, 900))))
            qrys.append(refs[-1][rng.randint(0, 50):][:rng.randint(100, 400)])
    return refs, qrys


def test_results_beyond_short_are_range_errors(caller, tmp_path):
    """Scores above 32,767 do not fit alignment_results' shorts: the call reports
    GBSW_E_RANGE (exit with a GPUassert message by default; with
    GBSW_DRIVER_NOEXIT the status and the values clamped to the short range),
    never wrapped values."""
    refs, qrys = _saturating_proteins()
    fr, fq, fm = tmp_path / "refs.txt", tmp_path / "queries.txt", tmp_path / "blosum62.txt"
    fr.write_text("\n".join(refs) + "\n")
    fq.write_text("\n".join(qrys) + "\n")
    fm.write_text(" ".join(str(v) for v in BLOSUM62) + "\n")
    args = [caller, "fileaa", str(fr), str(fq), str(fm), "-6", "-1"]
    out = run_caller(args)
    assert out.returncode == 1 and "GPUassert" in out.stderr and "short range" in out.stderr, out.stderr
    out = run_caller(args, env={"GBSW_DRIVER_NOEXIT": "1"})
    assert out.returncode == 0, out.stderr[-2000:]
    rows, status = columns(out.stdout)
    assert status == gbsw.E_RANGE
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    exp = ol.align_batch(rb, ro, qb, qo, ol.aa_scoring(BLOSUM62, -6, -1))
    want = np.clip(expected_columns(exp), -32768, 32767)
    # the end columns print end - 1 of the clamped end
    want[:, 2] = np.clip(exp["ref_end"], -32768, 32767) - 1
    want[:, 4] = np.clip(exp["query_end"], -32768, 32767) - 1
    assert exp["score"].max() > 32767
    assert_rows(rows, want, "clamped")


def test_kernel_driver_aa_asymmetric_table(caller, tmp_path):
    """kernel_driver_aa with an asymmetric 576-entry table (the reference takes
    any table, driver.cpp:304-308): the five columns equal the oracle's."""
    from test_gpu_asymmetric import skewed_blosum
    w = synth.plain_protein(3000, seed=0x5EED0503)
    table = skewed_blosum(11)
    fr, fq, fm = tmp_path / "refs.txt", tmp_path / "queries.txt", tmp_path / "m.txt"
    write_lines(fr, w.ref, w.ref_off)
    write_lines(fq, w.query, w.query_off)
    fm.write_text(" ".join(str(v) for v in table) + "\n")
    out = run_caller([caller, "fileaa", str(fr), str(fq), str(fm), "-6", "-1"])
    assert out.returncode == 0, out.stderr[-2000:]
    rows, status = columns(out.stdout)
    assert status == 0
    exp = ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, ol.aa_scoring(table, -6, -1))
    assert_rows(rows, expected_columns(exp), "kernel_driver_aa asymmetric table")
