// A caller written against GPU-BSW's driver API (driver.hpp / utils.hpp), built
// against this repository's headers and libgbsw.so.  Usage:
//   dropin_caller dna|aa             -> aligns a few fixed pairs and prints five columns
//   dropin_caller file REFS QUERIES [FACTOR]
//                                    -> DNA pairs, one sequence per line of each file
//                                       (+3/-3/-6/-1), five columns per pair
//   dropin_caller fileaa REFS QUERIES MATRIX OPEN EXTEND [FACTOR]
//                                    -> protein pairs through kernel_driver_aa; MATRIX is
//                                       a file of 576 scores (ARNDCQEGHILKMFPSTWYVBZX*
//                                       order, row-major)
// The last line is "status N" (gpu_bsw_driver::last_status()).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "driver.hpp"
#include "utils.hpp"

static void read_lines(const char* path, std::vector<std::string>& v)
{
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) v.push_back(line);
}

int main(int argc, char** argv)
{
    const char* mode = argc > 1 ? argv[1] : "dna";
    std::vector<std::string> contigs, reads;
    gpu_bsw_driver::alignment_results res;
    if (std::strcmp(mode, "file") == 0 && argc > 3) {
        read_lines(argv[2], contigs);
        read_lines(argv[3], reads);
        short scores[] = {3, -3, -6, -1};
        const float factor = argc > 4 ? (float)std::atof(argv[4]) : 1.0f;
        gpu_bsw_driver::kernel_driver_dna(reads, contigs, &res, scores, factor);
    } else if (std::strcmp(mode, "fileaa") == 0 && argc > 6) {
        read_lines(argv[2], contigs);
        read_lines(argv[3], reads);
        std::vector<short> m(576, 0);
        std::ifstream fm(argv[4]);
        for (int i = 0; i < 576; ++i) {
            int v = 0;
            fm >> v;
            m[i] = (short)v;
        }
        const float factor = argc > 7 ? (float)std::atof(argv[7]) : 1.0f;
        gpu_bsw_driver::kernel_driver_aa(reads, contigs, &res, m.data(), (short)std::atoi(argv[5]),
                                         (short)std::atoi(argv[6]), factor);
    } else if (std::strcmp(mode, "aa") != 0) {
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
    std::printf("status %d\n", gpu_bsw_driver::last_status());
    free_alignments(&res);
    return 0;
}
