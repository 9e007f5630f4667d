"""program_gpu (gpu-bsw_amd/csrc/program_gpu.cpp): the sample program's command
line (main.cpp:236-245) and the regression script beside it.

CPU tests check argument handling and the script's reporting of missing inputs;
GPU tests run the program on FASTA files and compare its five columns with the
oracle (score, ref_begin, ref_end-1, query_begin, query_end-1, main.cpp:130-137).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
import gbsw
from gbsw import synth
from gbsw.scoring import BLOSUM62

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(REPO, "gpu-bsw_amd", "program_gpu")
SCRIPT = os.path.join(REPO, "gpu-bsw_amd", "regression_test.sh")


def need_prog():
    if not os.path.exists(PROG):
        pytest.fail("program_gpu is not built; run __graft_entry__.build() or make -C gpu-bsw_amd")


def test_usage_errors():
    need_prog()
    for args in ([], ["rna", "a", "b", "c"], ["dna", "a", "b", "c", "--scores=1,2"], ["aa", "a", "b", "c", "--x"]):
        r = subprocess.run([PROG] + args, capture_output=True, text=True)
        assert r.returncode == 2 and "usage" in r.stderr


def test_missing_input_files(tmp_path):
    need_prog()
    r = subprocess.run([PROG, "dna", str(tmp_path / "no.fa"), str(tmp_path / "no.fa"), str(tmp_path / "o")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "cannot read" in r.stderr


def test_regression_script_reports_missing_inputs(tmp_path):
    need_prog()
    r = subprocess.run(["bash", SCRIPT, str(tmp_path), str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 2 and "inputs missing" in r.stdout


def write_fasta(path, seqs, headers=True):
    with open(path, "wb") as f:
        for k, s in enumerate(seqs):
            if headers:
                f.write(b">%d\n" % k)
            f.write(s + b"\n")


def expected_lines(refs, qrys, sc):
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    o = ol.align_batch(rb, ro, qb, qo, sc)
    return ["%d\t%d\t%d\t%d\t%d" % (o["score"][k], o["ref_begin"][k], o["ref_end"][k] - 1, o["query_begin"][k],
                                    o["query_end"][k] - 1) for k in range(len(refs))]


def run_prog(args):
    r = subprocess.run([PROG] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


@pytest.mark.gpu
def test_program_gpu_dna_and_protein(tmp_path):
    need_prog()
    w = synth.plain_dna(300, seed=0x5EED0500)
    refs = [w.pair(k)[0] for k in range(w.n)] + [b"ACGT"]
    qrys = [w.pair(k)[1] for k in range(w.n)] + [b""]          # an empty query line
    write_fasta(tmp_path / "r.fa", refs)
    write_fasta(tmp_path / "q.fa", qrys, headers=False)
    for extra, scores in (([], (1, -3, -3, -1)), (["--scores=3,-3,-6,-1", "--plain"], (3, -3, -6, -1))):
        print("case program_gpu dna", extra, flush=True)
        out = run_prog(["dna", str(tmp_path / "r.fa"), str(tmp_path / "q.fa"), str(tmp_path / "o.txt")] + extra)
        cells = sum(len(r) * len(q) for r, q in zip(refs, qrys))
        assert "Total Cells:%d" % cells in out
        assert "Total Alignments:%d" % len(refs) in out
        got = open(tmp_path / "o.txt").read().splitlines()
        assert got == expected_lines(refs, qrys, ol.dna_scoring(*scores))
    p = synth.plain_protein(200, seed=0x5EED0501)
    refs = [p.pair(k)[0] for k in range(p.n)]
    qrys = [p.pair(k)[1] for k in range(p.n)]
    write_fasta(tmp_path / "pr.fa", refs)
    write_fasta(tmp_path / "pq.fa", qrys)
    for extra, gaps in (([], (-6, -1)), (["--gaps=-10,-2"], (-10, -2))):
        print("case program_gpu aa", extra, flush=True)
        run_prog(["aa", str(tmp_path / "pr.fa"), str(tmp_path / "pq.fa"), str(tmp_path / "po.txt")] + extra)
        got = open(tmp_path / "po.txt").read().splitlines()
        assert got == expected_lines(refs, qrys, ol.aa_scoring(BLOSUM62, *gaps))


@pytest.mark.gpu
def test_regression_script_on_a_checkout_layout(tmp_path):
    """The script on a directory laid out like a GPU-BSW checkout, with expected
    scores from the oracle under the sample program's scoring."""
    need_prog()
    src = tmp_path / "src"
    (src / "test-data").mkdir(parents=True)
    (src / "regression_test").mkdir()
    d = synth.plain_dna(150, seed=0x5EED0502)
    p = synth.plain_protein(100, seed=0x5EED0503)
    for kind, w, sc in (("dna", d, ol.dna_scoring(1, -3, -3, -1)), ("protein", p, ol.aa_scoring(BLOSUM62, -6, -1))):
        refs = [w.pair(k)[0] for k in range(w.n)]
        qrys = [w.pair(k)[1] for k in range(w.n)]
        write_fasta(src / "test-data" / ("%s-reference.fasta" % kind), refs)
        write_fasta(src / "test-data" / ("%s-query.fasta" % kind), qrys)
        scores = [ln.split("\t")[0] for ln in expected_lines(refs, qrys, sc)]
        (src / "regression_test" / ("test_regression_%s_good_result" % kind)).write_text("\n".join(scores) + "\n")
    r = subprocess.run(["bash", SCRIPT, str(src), str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "REGRESSION TESTS PASS" in r.stdout
