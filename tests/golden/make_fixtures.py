"""Writes the small fixtures under tests/golden/ from the reference's own test data.

Run in the build container (where the read-only reference checkout exists):
    python tests/golden/make_fixtures.py /root/reference

Only data files are read (no reference code is imported or run):
  test-data/protein-query.fasta                         (30,000 single-line records)
  regression_test/test_regression_protein_good_result   (30,000 golden top scores)

Output:
  protein_query_golden.tsv : first N records as "<query>\t<golden score>"
  reference_tables.json    : the numbers of the reference's residue encoding
                             (driver.cpp:252-258) and of the BLOSUM62 matrix its
                             sample program passes (main.cpp:70-92), parsed as data
The protein reference sequences and all DNA inputs are missing from the
reference checkout (.MISSING_LARGE_BLOBS), so the protein goldens can only bound
scores, and the DNA goldens (regression_test/test_regression_dna_good_result),
whose inputs are all missing, pin nothing and are not kept.
"""
import json
import os
import re
import sys

N_PROTEIN = 3000


def read_single_line_fasta(path):
    """Same reader convention as main.cpp:22-36: skip lines starting with '>',
    every other line is one sequence."""
    seqs = []
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line.startswith(">"):
                continue
            seqs.append(line)
    return seqs


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    queries = read_single_line_fasta(os.path.join(ref_root, "test-data", "protein-query.fasta"))
    with open(os.path.join(ref_root, "regression_test", "test_regression_protein_good_result")) as fh:
        golden = [int(x.split()[0]) for x in fh if x.strip()]
    assert len(queries) == len(golden) == 30000, (len(queries), len(golden))
    with open(os.path.join(here, "protein_query_golden.tsv"), "w") as out:
        for q, g in zip(queries[:N_PROTEIN], golden[:N_PROTEIN]):
            out.write("%s\t%d\n" % (q, g))
    drv = open(os.path.join(ref_root, "driver.cpp")).read()
    a = drv.index("short encoding_matrix[] = {")
    enc = [int(x) for x in re.findall(r"-?\d+", drv[drv.index("{", a) + 1:drv.index("}", a)])]
    main = open(os.path.join(ref_root, "main.cpp")).read()
    b = main.index("short scores_matrix[] = {")
    blosum = [int(x) for x in re.findall(r"-?\d+", main[main.index("{", b) + 1:main.index("}", b)])]
    assert len(enc) == 91 and len(blosum) == 576
    with open(os.path.join(here, "reference_tables.json"), "w") as out:
        json.dump({"source": "numbers parsed from driver.cpp:252-258 (encoding_matrix) and main.cpp:70-92 "
                             "(scores_matrix) of the reference", "encoding": enc, "blosum62": blosum}, out)
    print("wrote %d protein golden rows and the reference tables" % N_PROTEIN)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
