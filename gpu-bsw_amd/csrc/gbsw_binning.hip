// gbsw_binning.hip -- device-side length binning of a batch (batched form).
//
// Every pair gets a 22-bit key: bucket id of its shorter length (6 bits), then
// 31 - (strips - 1) (5 bits; strips of the bucket's width its shorter length
// spans, capped at 32), then 2047 - (longer length / 32) (11 bits).  A stable
// radix sort of (key, pair index) yields the work order: buckets in ascending
// order, inside a bucket the most strips first and then the longest pairs
// (longest-processing-time order for the persistent work queue), so the four
// pairs of a wave have like strip and step counts.  The reference has no
// counterpart: it launches pairs in input order (driver.cpp:176-185).
#include <algorithm>

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "gbsw_kernels.hpp"

namespace gbsw {

__global__ void k_bin_keys(const uint64_t* ref_off, const uint64_t* qry_off, int64_t n, const uint8_t* bucket_of_len,
                           StripWidths widths, uint32_t* keys, int32_t* vals)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t la = (uint32_t)(ref_off[p + 1] - ref_off[p]);
    const uint32_t lb = (uint32_t)(qry_off[p + 1] - qry_off[p]);
    const uint32_t m = la < lb ? la : lb, M = la < lb ? lb : la;
    const uint32_t b = bucket_of_len[m < 65535u ? m : 65535u] & 63u;
    const uint32_t w = widths.w[b] ? widths.w[b] : 1u;
    const uint32_t strips = m ? (m + w - 1) / w : 1u;
    const uint32_t sk = 31u - (strips < 32u ? strips - 1u : 31u);
    const uint32_t mk = 2047u - ((M >> 5) < 2047u ? (M >> 5) : 2047u);
    keys[p] = (b << 16) | (sk << 11) | mk;
    vals[p] = (int32_t)p;
}

// Which byte values occur in p[0, n): bit b of mask[b / 32].  Each thread reads
// 16 bytes at a time (p is 16-byte aligned) into a private 256-bit mask (eight
// registers, set without indexing), then the masks are OR-ed per workgroup in
// LDS and into `mask`, which the caller zeroes.
__global__ void k_byte_presence(const uint8_t* p, uint64_t n, uint32_t* mask)
{
    __shared__ uint32_t m[8];
    if (threadIdx.x < 8) m[threadIdx.x] = 0;
    __syncthreads();
    uint32_t loc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto add = [&](uint32_t b) {
        const uint32_t w = b >> 5, bit = 1u << (b & 31);
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) loc[j] |= (w == j) ? bit : 0u;
    };
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 16;
    for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16; i < n; i += stride) {
        if (i + 16 <= n) {
            const uint4 v = *(const uint4*)(p + i);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int k = 0; k < 4; ++k) add((w[q] >> (8 * k)) & 0xFFu);
        } else {
            for (uint64_t j = i; j < n; ++j) add(p[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if (loc[j]) atomicOr(&m[j], loc[j]);
    __syncthreads();
    if (threadIdx.x < 8 && m[threadIdx.x]) atomicOr(mask + threadIdx.x, m[threadIdx.x]);
}

hipError_t byte_presence(const uint8_t* p, uint64_t n, uint32_t* mask, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const uint64_t threads = (n + 15) / 16;
    const unsigned blocks = (unsigned)std::min<uint64_t>(2048, (threads + 255) / 256);
    hipLaunchKernelGGL(k_byte_presence, dim3(blocks), dim3(256), 0, stream, p, n, mask);
    return hipGetLastError();
}

hipError_t bin_keys(const uint64_t* ref_off, const uint64_t* qry_off, int64_t n, const uint8_t* bucket_of_len,
                    const StripWidths& widths, uint32_t* keys, int32_t* vals, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ref_off, qry_off, n,
                       bucket_of_len, widths, keys, vals);
    return hipGetLastError();
}

// Stable LSD radix sort of (key, value) on bits [0, end_bit).  With temp == nullptr
// only the scratch size is returned in temp_bytes.
hipError_t sort_pairs(void* temp, size_t& temp_bytes, const uint32_t* kin, uint32_t* kout, const int32_t* vin,
                      int32_t* vout, int n, int end_bit, hipStream_t stream)
{
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, kin, kout, vin, vout, n, 0, end_bit, stream);
}

} // namespace gbsw
