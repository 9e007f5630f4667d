// gbsw_kernels.hpp -- host-visible entry points of gbsw_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "gbsw_kernel_args.hpp"

namespace gbsw {

enum KernelKind { kPlainForward = 0, kPlainReverse = 1, kBatched = 2 };

bool kernel_supported(int K);
const int* kernel_k_list(int* count); // K values with compiled kernels, ascending
int kernel_occupancy(int kind, int K, int lds_bytes);
hipError_t kernel_launch(int kind, int K, int grid, int lds_bytes, hipStream_t stream, const KernelArgs& args);
hipError_t zero_u32(uint32_t* p, int n, hipStream_t stream);

} // namespace gbsw
