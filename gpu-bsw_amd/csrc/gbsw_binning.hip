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
