// CPU check of the split form's placement guard (gbsw_kernel_args.hpp):
// simulates the per-workgroup compare-and-swap into the kXcds-slot table for a
// workgroup -> XCC mapping and prints the host's decision (1 = rerun unsplit).
//   placement_check MODE GRID
//   MODE: rr8 (w % 8), rr4 (w % 4: four XCCs per partition), one (a single
//         XCC), shift (w % 8 for the first half of the grid, (w + 1) % 8 after),
//         swap (w % 8 except one workgroup placed on the XCC of its neighbour)
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../gpu-bsw_amd/csrc/gbsw_kernel_args.hpp"

using namespace gbsw;

static uint32_t xcc_of(const char* mode, int w, int grid)
{
    if (!std::strcmp(mode, "rr8")) return (uint32_t)(w % 8);
    if (!std::strcmp(mode, "rr4")) return (uint32_t)(w % 4);
    if (!std::strcmp(mode, "one")) return 0u;
    if (!std::strcmp(mode, "shift")) return (uint32_t)((w < grid / 2 ? w : w + 1) % 8);
    if (!std::strcmp(mode, "swap")) return (uint32_t)((w == grid / 3 ? w + 1 : w) % 8);
    std::fprintf(stderr, "unknown mode %s\n", mode);
    std::exit(2);
}

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    const int grid = std::atoi(argv[2]);
    uint32_t table[kXcds] = {0};
    int32_t status[kStatusWords] = {0};
    for (int w = 0; w < grid; ++w) {
        const int x = w % kXcds;
        const uint32_t xcc = xcc_of(argv[1], w, grid);
        const uint32_t prev = table[x];          // the compare-and-swap's old value
        if (prev == 0u) table[x] = xcc + 1u;
        if (!placement_consistent(prev, xcc)) status[kStatusMisplaced] |= 1;
    }
    std::printf("%d\n", split_rerun_needed(status) ? 1 : 0);
    return 0;
}
