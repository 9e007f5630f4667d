// gbsw_gather.hip -- device-side staging of a serving micro-batch.
//
// One launch moves every request part of a micro-batch from pinned host memory
// (the server's arena, or caller buffers from gbsw_host_alloc) into the slot's
// device-resident sequence buffers, packed back to back in micro-batch order, and
// checks each byte against the alphabet on the way.  This replaces per-request
// copy calls (the reference issues a cudaMemcpyAsync per array, stream half and
// chunk, utils.cpp:30-54).
//
// A tile reads the aligned 16-byte blocks that cover its source range.  Each of
// them holds at least one byte of the range, so no read leaves the pages of the
// source buffer; the bytes are then realigned through LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gbsw_kernels.hpp"

namespace gbsw {

constexpr int kGatherThreads = 256;

__global__ __launch_bounds__(kGatherThreads) void k_gather(const GatherTile* tiles, const uint8_t* accept,
                                                           int32_t* part_flags)
{
    __shared__ uint8_t ok[256];
    __shared__ uint4 stage[kGatherTileBytes / 16 + 1];
    for (int i = threadIdx.x; i < 256; i += kGatherThreads) ok[i] = accept[i];
    const GatherTile t = tiles[blockIdx.x];
    const uintptr_t s = (uintptr_t)t.src;
    const uintptr_t s0 = s & ~(uintptr_t)15;
    const uintptr_t s1 = (s + t.len + 15) & ~(uintptr_t)15;
    const uint32_t nblk = (uint32_t)((s1 - s0) / 16); // <= kGatherTileBytes / 16 + 1
    const uint4* src = (const uint4*)s0;
    for (uint32_t i = threadIdx.x; i < nblk; i += kGatherThreads) stage[i] = src[i];
    __syncthreads();
    const uint8_t* sb = (const uint8_t*)stage + (s - s0);
    uint32_t bad = 0;
    for (uint32_t i = threadIdx.x; i < t.len; i += kGatherThreads) {
        const uint8_t b = sb[i];
        t.dst[i] = b;
        bad |= ok[b] ^ 1u;
    }
    if (__syncthreads_or((int)bad) && threadIdx.x == 0) atomicOr(&part_flags[t.part], 1);
}

hipError_t gather(const GatherTile* tiles, int ntiles, const uint8_t* accept, int32_t* part_flags,
                  hipStream_t stream)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)ntiles), dim3(kGatherThreads), 0, stream, tiles, accept, part_flags);
    return hipGetLastError();
}

} // namespace gbsw
