// pj_scan.h -- block-level and device-wide exclusive scans (u32), shared by
// the partition passes of the join and of PartitionOp. Kept in an anonymous
// namespace so each translation unit owns its copy of the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pj {
namespace {

// Exclusive scan of one value per thread across the block. wsum needs
// NT/64 + 1 words of LDS. Returns the thread's exclusive prefix; *total gets
// the block sum. Ends with a barrier, so wsum may be reused after it.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t* total)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d)
            x += y;
    }
    if (lane == 63)
        wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        uint32_t w = lane < NW ? wsum[lane] : 0u;
        uint32_t s = w;
#pragma unroll
        for (int d = 1; d < NW; d <<= 1) {
            uint32_t y = __shfl_up(s, d, 64);
            if (lane >= d)
                s += y;
        }
        if (lane < NW)
            wsum[lane] = s - w;
        if (lane == NW - 1)
            wsum[NW] = s;
    }
    __syncthreads();
    uint32_t r = x - v + wsum[wid];
    *total = wsum[NW];
    __syncthreads();
    return r;
}

// In-place exclusive scan of a[0..n) in LDS; a[n] receives the total.
template <int NT>
__device__ __forceinline__ void block_scan_array(uint32_t* a, uint32_t n, uint32_t* wsum)
{
    const uint32_t seg = (n + NT - 1) / NT;
    const uint32_t b = threadIdx.x * seg;
    uint32_t s = 0;
    for (uint32_t i = 0; i < seg; ++i)
        if (b + i < n)
            s += a[b + i];
    uint32_t total;
    uint32_t pre = block_excl_scan<NT>(s, wsum, &total);
    for (uint32_t i = 0; i < seg; ++i)
        if (b + i < n) {
            uint32_t v = a[b + i];
            a[b + i] = pre;
            pre += v;
        }
    if (threadIdx.x == 0)
        a[n] = total;
    __syncthreads();
}

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanSeg = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ a, uint64_t m,
                                                               uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t wsum[kScanThreads / 64 + 1];
    const uint64_t b = (uint64_t)blockIdx.x * kScanSeg;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        uint64_t i = b + (uint64_t)k * kScanThreads + threadIdx.x;
        if (i < m)
            s += a[i];
    }
    uint32_t total;
    block_excl_scan<kScanThreads>(s, wsum, &total);
    if (threadIdx.x == 0)
        bsum[blockIdx.x] = total;
}

// Exclusive scan of bsum[0..nb) by one block; bsum[nb] = total.
__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ bsum, uint32_t nb)
{
    __shared__ uint32_t wsum[1024 / 64 + 1];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < nb; b += 1024) {
        uint32_t i = b + threadIdx.x;
        uint32_t v = i < nb ? bsum[i] : 0u;
        uint32_t total;
        uint32_t pre = block_excl_scan<1024>(v, wsum, &total);
        if (i < nb)
            bsum[i] = pre + carry;
        __syncthreads();
        if (threadIdx.x == 0)
            carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        bsum[nb] = carry;
}

// a[i] = exclusive prefix of a (in place), using the scanned block sums.
__global__ __launch_bounds__(kScanThreads) void k_scan_down(uint32_t* __restrict__ a, uint64_t m,
                                                             const uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t wsum[kScanThreads / 64 + 1];
    const uint64_t b = (uint64_t)blockIdx.x * kScanSeg + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = b + k < m ? a[b + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t pre = block_excl_scan<kScanThreads>(s, wsum, &total) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (b + k < m)
            a[b + k] = pre;
        pre += v[k];
    }
}

// The whole exclusive scan of a[0..m) (m <= kScanSmall) by one block, then
// the partition starts of a histogram laid out a[p * nblk + blk]:
// part_start[p] = a[p * nblk], part_start[F] = the total. Replaces the
// three-kernel scan and k_part_bounds for small histograms.
constexpr uint32_t kScanSmallThreads = 1024;
constexpr uint32_t kScanSmallPer = 16;
constexpr uint32_t kScanSmall = kScanSmallThreads * kScanSmallPer;

__global__ __launch_bounds__(kScanSmallThreads) void k_scan_small(uint32_t* __restrict__ a, uint32_t m, uint32_t F,
                                                                   uint32_t nblk, uint32_t* __restrict__ part_start)
{
    __shared__ uint32_t wsum[kScanSmallThreads / 64 + 1];
    const uint32_t per = (m + kScanSmallThreads - 1) / kScanSmallThreads;
    const uint32_t b = threadIdx.x * per;
    uint32_t v[kScanSmallPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanSmallPer; ++k) {
        v[k] = k < per && b + k < m ? a[b + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t pre = block_excl_scan<kScanSmallThreads>(sum, wsum, &total);
#pragma unroll
    for (uint32_t k = 0; k < kScanSmallPer; ++k)
        if (k < per && b + k < m) {
            a[b + k] = pre;
            pre += v[k];
        }
    __syncthreads();   // the scanned entries of other threads are read below
    for (uint32_t p = threadIdx.x; p <= F; p += kScanSmallThreads)
        part_start[p] = p < F ? a[(uint64_t)p * nblk] : total;
}

}  // namespace
}  // namespace pj
