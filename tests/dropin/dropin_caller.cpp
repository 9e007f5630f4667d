// A caller written against GPU-BSW's driver API (driver.hpp / utils.hpp), built
// against this repository's headers and libgbsw.so.  Usage:
//   dropin_caller dna|aa             -> aligns a few fixed pairs and prints five columns
//   dropin_caller file REFS QUERIES  -> DNA pairs, one sequence per line of each file
//                                       (+3/-3/-6/-1), five columns per pair
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "driver.hpp"
#include "utils.hpp"

int main(int argc, char** argv)
{
    const bool aa = argc > 1 && std::strcmp(argv[1], "aa") == 0;
    const bool file = argc > 3 && std::strcmp(argv[1], "file") == 0;
    std::vector<std::string> contigs, reads;
    gpu_bsw_driver::alignment_results res;
    if (file) {
        std::ifstream fr(argv[2]), fq(argv[3]);
        std::string line;
        while (std::getline(fr, line)) contigs.push_back(line);
        while (std::getline(fq, line)) reads.push_back(line);
        short scores[] = {3, -3, -6, -1};
        gpu_bsw_driver::kernel_driver_dna(reads, contigs, &res, scores, 1.0);
    } else if (!aa) {
        contigs = {"GGGGGACGTGGGG", "ACGT", "TTTTACGTACGTTTTT"};
        reads = {"ACGTTACGT", "ACGTTACGT", "ACGTACGT"};
        short scores[] = {1, -3, -3, -1};
        gpu_bsw_driver::kernel_driver_dna(reads, contigs, &res, scores, 0.5);
    } else {
        contigs = {"MKVLAAGIVGLLLAW", "AWWWA"};
        reads = {"KVLAAGIV", "WWW"};
        static const char order[] = "ARNDCQEGHILKMFPSTWYVBZX*";
        std::vector<short> m(576, -1);
        for (int i = 0; i < 24; ++i) m[i * 24 + i] = (order[i] == 'W') ? 11 : 4;
        gpu_bsw_driver::kernel_driver_aa(reads, contigs, &res, m.data(), -6, -1, 0.5);
    }
    for (size_t k = 0; k < contigs.size(); ++k)
        std::printf("%d\t%d\t%d\t%d\t%d\n", res.top_scores[k], res.ref_begin[k], res.ref_end[k] - 1,
                    res.query_begin[k], res.query_end[k] - 1);
    std::printf("max_len %u\n", getMaxLength(contigs));
    free_alignments(&res);
    return 0;
}
