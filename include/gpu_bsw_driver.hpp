// gpu_bsw_driver.hpp -- source-compatible drop-in for GPU-BSW's driver API.
//
// Replaces driver.hpp:18-42 of the reference (mgawan/GPU-BSW): the same namespace,
// struct and function signatures, so a program written against the reference
// recompiles unchanged and links libgbsw.so instead.
//
//   kernel_driver_dna   <- driver.hpp:32-33 (driver.cpp:15-243)
//   kernel_driver_aa    <- driver.hpp:36-37 (driver.cpp:245-489)
//   alignment_results   <- driver.hpp:21-27
//   get_tot_gpu_mem     <- driver.hpp:29-30 (driver.cpp:10-14)
//   free_alignments     <- utils.hpp:25 (utils.cpp:21-28)
//
// Behaviour kept: reads[k] (query) pairs with contigs[k] (ref); the driver
// allocates the five result arrays in pinned host memory and the caller frees
// them with free_alignments; ends are 1-based, begins 0-based; errors print a
// message and exit the process (utils.hpp:12-21), unless GBSW_DRIVER_NOEXIT is set
// in the environment, in which case the results are zero-filled and
// gpu_bsw_driver::last_status() reports the error; `factor` bounds the device
// memory a call uses to factor x the device's memory (driver.cpp:36, 51-55): the
// pairs run in as many sub-batches as that takes.
// Behaviour changed: one device (the current HIP device, or GBSW_DEVICE), no
// OpenMP side effects; a result value outside the `short` range (scores or
// coordinates above 32767, which the reference wraps) is an error,
// GBSW_E_RANGE, with the arrays holding the values clamped to that range.
#ifndef GPU_BSW_DRIVER_HPP
#define GPU_BSW_DRIVER_HPP

#include <cstddef>
#include <string>
#include <vector>

namespace gpu_bsw_driver {

// for storing the alignment results
struct alignment_results {
    short* ref_begin;
    short* query_begin;
    short* ref_end;
    short* query_end;
    short* top_scores;
};

size_t get_tot_gpu_mem(int id);

void kernel_driver_dna(std::vector<std::string> reads, std::vector<std::string> contigs,
                       alignment_results* alignments, short scores[4], float factor = 1.0);

void kernel_driver_aa(std::vector<std::string> reads, std::vector<std::string> contigs,
                      alignment_results* alignments, short scoring_matrix[], short openGap,
                      short extendGap, float factor = 1.0);

// Status of the last driver call on this thread (0 = success, else a gbsw.h code).
int last_status();

} // namespace gpu_bsw_driver

void free_alignments(gpu_bsw_driver::alignment_results* alignments);

#endif
