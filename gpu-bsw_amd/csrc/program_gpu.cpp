// program_gpu.cpp -- the command line of GPU-BSW's sample program
// (main.cpp:236-245) on this engine:
//
//   program_gpu {dna|aa} <ref.fasta> <query.fasta> <out> [--scores=m,x,o,e] [--gaps=o,e] [--plain]
//
// Input: FASTA files with each sequence on one line -- a line starting with '>'
// is a header, every other line is one sequence, in file order (main.cpp:20-37,
// 160-197).  Pair k aligns reference k with query k; the pair count is the
// number of references, and a query file with fewer sequences is an error.
// Output file: one line per pair, "score\tref_begin\tref_end-1\tquery_begin\tquery_end-1"
// (main.cpp:130-137); stdout: the driver's report lines (driver.cpp:240-242) and
// "Total Cells:<sum of |ref| x |query|>" (main.cpp:141-147).
// Scoring as in the sample program: DNA +1/-3/-3/-1 (main.cpp:210), protein
// BLOSUM62 with gap open -6 and extend -1 (main.cpp:70-123); --scores and
// --gaps override them.  Results are the engine's int32 outputs, so scores above
// 32,767 print correctly (the reference's short results cannot hold them).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gbsw.h"

namespace {

const char kAaOrder[] = "ARNDCQEGHILKMFPSTWYVBZX*";

// BLOSUM62 (Henikoff & Henikoff 1992) as published by NCBI, rows and columns in
// kAaOrder.
const char kBlosum62[] =
    "A  4 -1 -2 -2  0 -1 -1  0 -2 -1 -1 -1 -1 -2 -1  1  0 -3 -2  0 -2 -1  0 -4\n"
    "R -1  5  0 -2 -3  1  0 -2  0 -3 -2  2 -1 -3 -2 -1 -1 -3 -2 -3 -1  0 -1 -4\n"
    "N -2  0  6  1 -3  0  0  0  1 -3 -3  0 -2 -3 -2  1  0 -4 -2 -3  3  0 -1 -4\n"
    "D -2 -2  1  6 -3  0  2 -1 -1 -3 -4 -1 -3 -3 -1  0 -1 -4 -3 -3  4  1 -1 -4\n"
    "C  0 -3 -3 -3  9 -3 -4 -3 -3 -1 -1 -3 -1 -2 -3 -1 -1 -2 -2 -1 -3 -3 -2 -4\n"
    "Q -1  1  0  0 -3  5  2 -2  0 -3 -2  1  0 -3 -1  0 -1 -2 -1 -2  0  3 -1 -4\n"
    "E -1  0  0  2 -4  2  5 -2  0 -3 -3  1 -2 -3 -1  0 -1 -3 -2 -2  1  4 -1 -4\n"
    "G  0 -2  0 -1 -3 -2 -2  6 -2 -4 -4 -2 -3 -3 -2  0 -2 -2 -3 -3 -1 -2 -1 -4\n"
    "H -2  0  1 -1 -3  0  0 -2  8 -3 -3 -1 -2 -1 -2 -1 -2 -2  2 -3  0  0 -1 -4\n"
    "I -1 -3 -3 -3 -1 -3 -3 -4 -3  4  2 -3  1  0 -3 -2 -1 -3 -1  3 -3 -3 -1 -4\n"
    "L -1 -2 -3 -4 -1 -2 -3 -4 -3  2  4 -2  2  0 -3 -2 -1 -2 -1  1 -4 -3 -1 -4\n"
    "K -1  2  0 -1 -3  1  1 -2 -1 -3 -2  5 -1 -3 -1  0 -1 -3 -2 -2  0  1 -1 -4\n"
    "M -1 -1 -2 -3 -1  0 -2 -3 -2  1  2 -1  5  0 -2 -1 -1 -1 -1  1 -3 -1 -1 -4\n"
    "F -2 -3 -3 -3 -2 -3 -3 -3 -1  0  0 -3  0  6 -4 -2 -2  1  3 -1 -3 -3 -1 -4\n"
    "P -1 -2 -2 -1 -3 -1 -1 -2 -2 -3 -3 -1 -2 -4  7 -1 -1 -4 -3 -2 -2 -1 -2 -4\n"
    "S  1 -1  1  0 -1  0  0  0 -1 -2 -2  0 -1 -2 -1  4  1 -3 -2 -2  0  0  0 -4\n"
    "T  0 -1  0 -1 -1 -1 -1 -2 -2 -1 -1 -1 -1 -2 -1  1  5 -2 -2  0 -1 -1  0 -4\n"
    "W -3 -3 -4 -4 -2 -2 -3 -2 -2 -3 -2 -3 -1  1 -4 -3 -2 11  2 -3 -4 -3 -2 -4\n"
    "Y -2 -2 -2 -3 -2 -1 -2 -3  2 -1 -1 -2 -1  3 -3 -2 -2  2  7 -1 -3 -2 -1 -4\n"
    "V  0 -3 -3 -3 -1 -2 -2 -3 -3  3  1 -2  1 -1 -2 -2  0 -3 -1  4 -3 -2 -1 -4\n"
    "B -2 -1  3  4 -3  0  1 -1  0 -3 -4  0 -3 -3 -2  0 -1 -4 -3 -3  4  1 -1 -4\n"
    "Z -1  0  0  1 -3  3  4 -2  0 -3 -3  1 -1 -3 -1  0 -1 -3 -2 -2  1  4 -1 -4\n"
    "X  0 -1 -1 -1 -2 -1 -1 -1 -1 -1 -1 -1 -1 -1 -2  0  0 -2 -1 -1 -1 -1 -1 -4\n"
    "* -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4  1\n";

bool parse_blosum62(int16_t* m)
{
    std::istringstream in(kBlosum62);
    for (int r = 0; r < 24; ++r) {
        char name = 0;
        if (!(in >> name) || name != kAaOrder[r]) return false;
        for (int c = 0; c < 24; ++c) {
            int v = 0;
            if (!(in >> v)) return false;
            m[r * 24 + c] = (int16_t)v;
        }
    }
    return true;
}

// Sequences of a single-line FASTA file; false if the file cannot be opened.
bool read_fasta(const char* path, std::vector<std::string>& seqs)
{
    std::ifstream f(path);
    if (!f.is_open()) return false;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line[0] == '>') continue;
        seqs.push_back(line);
    }
    return true;
}

void pack(const std::vector<std::string>& s, size_t n, std::vector<char>& bytes, std::vector<uint64_t>& off)
{
    off.assign(n + 1, 0);
    for (size_t k = 0; k < n; ++k) off[k + 1] = off[k] + s[k].size();
    bytes.resize(off[n] ? off[n] : 1);
    for (size_t k = 0; k < n; ++k)
        if (!s[k].empty()) std::memcpy(bytes.data() + off[k], s[k].data(), s[k].size());
}

bool parse_ints(const char* text, int* v, int n)
{
    for (int i = 0; i < n; ++i) {
        char* end = nullptr;
        const long x = std::strtol(text, &end, 10);
        if (end == text) return false;
        v[i] = (int)x;
        text = end;
        if (i + 1 < n) {
            if (*text != ',') return false;
            ++text;
        }
    }
    return *text == '\0';
}

int usage()
{
    std::fprintf(stderr,
                 "usage: program_gpu {dna|aa} <ref.fasta> <query.fasta> <out> [--scores=m,x,o,e] [--gaps=o,e] "
                 "[--plain]\n");
    return 2;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 5) return usage();
    const std::string kind = argv[1];
    if (kind != "dna" && kind != "aa") return usage();
    const bool protein = kind == "aa";
    gbsw_scoring sc;
    std::memset(&sc, 0, sizeof(sc));
    int dna[4] = {1, -3, -3, -1}, gaps[2] = {-6, -1};
    int32_t mode = GBSW_BATCHED;
    for (int i = 5; i < argc; ++i) {
        if (std::strncmp(argv[i], "--scores=", 9) == 0) {
            if (!parse_ints(argv[i] + 9, dna, 4)) return usage();
        } else if (std::strncmp(argv[i], "--gaps=", 7) == 0) {
            if (!parse_ints(argv[i] + 7, gaps, 2)) return usage();
        } else if (std::strcmp(argv[i], "--plain") == 0) {
            mode = GBSW_PLAIN;
        } else {
            return usage();
        }
    }
    if (protein) {
        sc.alphabet = GBSW_PROTEIN;
        if (!parse_blosum62(sc.matrix)) {
            std::fprintf(stderr, "program_gpu: internal BLOSUM62 table is malformed\n");
            return 1;
        }
        sc.gap_open = (int16_t)gaps[0];
        sc.gap_extend = (int16_t)gaps[1];
    } else {
        sc.alphabet = GBSW_DNA;
        sc.match = (int16_t)dna[0];
        sc.mismatch = (int16_t)dna[1];
        sc.gap_open = (int16_t)dna[2];
        sc.gap_extend = (int16_t)dna[3];
    }

    std::vector<std::string> refs, queries;
    if (!read_fasta(argv[2], refs) || !read_fasta(argv[3], queries)) {
        std::fprintf(stderr, "program_gpu: cannot read %s or %s\n", argv[2], argv[3]);
        return 1;
    }
    const size_t n = refs.size();
    if (queries.size() < n) {
        std::fprintf(stderr, "program_gpu: %zu references but only %zu queries\n", n, queries.size());
        return 1;
    }
    std::vector<char> rb, qb;
    std::vector<uint64_t> ro, qo;
    pack(refs, n, rb, ro);
    pack(queries, n, qb, qo);
    size_t max_ref = 0, max_qry = 0;
    long long cells = 0;
    for (size_t k = 0; k < n; ++k) {
        max_ref = std::max(max_ref, refs[k].size());
        max_qry = std::max(max_qry, queries[k].size());
        cells += (long long)refs[k].size() * (long long)queries[k].size();
    }

    const auto t0 = std::chrono::high_resolution_clock::now();
    gbsw_context* ctx = nullptr;
    std::vector<int32_t> out(5 * (n ? n : 1));
    int rc = gbsw_context_create(0, &ctx);
    if (rc == GBSW_OK)
        rc = gbsw_align(ctx, &sc, mode, (int64_t)n, rb.data(), ro.data(), qb.data(), qo.data(), out.data(),
                        out.data() + n, out.data() + 2 * n, out.data() + 3 * n, out.data() + 4 * n);
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (rc != GBSW_OK) {
        std::fprintf(stderr, "program_gpu: %s (gbsw status %d)\n", gbsw_last_error(), rc);
        gbsw_context_destroy(ctx);
        return 1;
    }
    gbsw_context_destroy(ctx);
    const std::chrono::duration<double> dt = t1 - t0;
    std::cout << "Total Alignments:" << n << "\n"
              << "Max Reference Size:" << max_ref << "\n"
              << "Max Query Size:" << max_qry << "\n"
              << "Total Execution Time (seconds):" << dt.count() << std::endl;

    std::ofstream f(argv[4]);
    if (!f.is_open()) {
        std::fprintf(stderr, "program_gpu: cannot write %s\n", argv[4]);
        return 1;
    }
    for (size_t k = 0; k < n; ++k)
        f << out[k] << "\t" << out[n + k] << "\t" << out[2 * n + k] - 1 << "\t" << out[3 * n + k] << "\t"
          << out[4 * n + k] - 1 << "\n";
    f.close();
    std::cout << "Total Cells:" << cells << std::endl;
    return 0;
}
