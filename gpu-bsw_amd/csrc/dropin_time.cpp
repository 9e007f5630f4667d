// dropin_time.cpp -- end-to-end timing of the gpu_bsw_driver drop-in, the way the
// reference reports it (its driver prints "Total Execution Time", driver.cpp:34,
// 240-242, covering packing, transfers, both passes and the copy back).
//
//   dropin_time REFS QUERIES REPS
//
// REFS and QUERIES hold one sequence per line (pair k = line k of each).  The DNA
// pairs (+3/-3/-6/-1) are aligned REPS times through kernel_driver_dna; the
// driver prints its summary lines each time; the program then prints
// "checksum <sum of scores> <sum of the four coordinates>" of the last call.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/gpu_bsw_driver.hpp"

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: dropin_time REFS QUERIES REPS\n");
        return 2;
    }
    std::vector<std::string> contigs, reads;
    std::ifstream fr(argv[1]), fq(argv[2]);
    std::string line;
    while (std::getline(fr, line)) contigs.push_back(line);
    while (std::getline(fq, line)) reads.push_back(line);
    if (contigs.empty() || reads.size() < contigs.size()) {
        std::fprintf(stderr, "dropin_time: need one query per reference\n");
        return 2;
    }
    short scores[] = {3, -3, -6, -1};
    long long sum_score = 0, sum_coord = 0;
    for (int rep = 0; rep < std::atoi(argv[3]); ++rep) {
        gpu_bsw_driver::alignment_results res;
        gpu_bsw_driver::kernel_driver_dna(reads, contigs, &res, scores, 1.0);
        sum_score = sum_coord = 0;
        for (size_t k = 0; k < contigs.size(); ++k) {
            sum_score += res.top_scores[k];
            sum_coord += res.ref_begin[k] + res.ref_end[k] + res.query_begin[k] + res.query_end[k];
        }
        free_alignments(&res);
    }
    std::printf("checksum %lld %lld\n", sum_score, sum_coord);
    return 0;
}
